"""Batched form and autotuner on the GPU (build additions, no reference
counterpart).

A batch merges several prepared problems into one block set; every kernel
launch covers all of them while each problem keeps its own PCG scalars,
smoother weight and convergence test.  Each problem's iteration count,
final residual and pressure must equal those of the plain solve on its own
handle (the same kernels over the same blocks; the FP64 dot products are
summed in a different order, which the FP32 rounding of alpha and beta
absorbs, so iterates are checked bit for bit and the FP64 residual history
to 1e-12 relative).  The autotuner's choices
(graph chunk, smoother kernel per block size) must not change any result."""
import numpy as np
import pytest

import msbgx

pytestmark = pytest.mark.gpu


def prepared(cfg):
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    return g


def pressure(g, base=0, n_blocks=None):
    lv, _ = g.block_info(0)
    n_blocks = n_blocks if n_blocks is not None else g.n_blocks
    out = []
    for b in range(n_blocks):
        e = g.bsx >> int(lv[base + b])
        out.append(g.get_block(msbgx.CH_PRESSURE, 0, base + b, e).copy())
    return out


def small_batch(n_problems, seed):
    """C3-style problems at 64^3 and 128^3 (sizes kept small for the test)."""
    cfgs = []
    for c in msbgx.config_c3(seed=seed, n_problems=4 * n_problems):
        if c.n <= 128:
            cfgs.append(c)
        if len(cfgs) == n_problems:
            break
    return cfgs


@pytest.mark.parametrize("opts", [0, msbgx.OPT_NO_GRAPH, msbgx.OPT_NO_COMPACT])
def test_batched_problems_equal_individual_solves(opts):
    """Default: converged problems leave the launches (list compaction)."""
    print(f"batch vs individual, options={opts}", flush=True)
    cfgs = small_batch(6, seed=7)
    assert len(cfgs) == 6
    singles = [prepared(c) for c in cfgs]
    batch = msbgx.BatchGrid.create(singles)
    batch.set_options(opts)
    it_b, rr_b = batch.solve(max_iter=200, tol=1e-6)
    for q, (c, g) in enumerate(zip(cfgs, singles)):
        it, rr = g.solve_pressure_flip(max_iter=200, tol=1e-6)
        print(f"  problem {q} {c.name} mode {c.level_mode}: single {it} {rr:.3e}  batch {it_b[q]} {rr_b[q]:.3e}",
              flush=True)
        assert it == it_b[q]
        assert rr == rr_b[q]
        # the FP64 history itself carries the summation-order difference
        np.testing.assert_allclose(g.residual_history(), batch.problem_history(q), rtol=1e-12, atol=0)
        ps = pressure(g)
        pb = pressure(batch, batch.bases[q], g.n_blocks)
        for b in range(g.n_blocks):
            np.testing.assert_array_equal(ps[b], pb[b], err_msg=f"problem {q} block {b}")


def test_batch_with_an_empty_problem():
    """A problem without liquid has |b| = 0: zero iterations, zero residual,
    and the other problems are unaffected."""
    print("batch with empty problem", flush=True)
    cfgs = small_batch(2, seed=11)
    empty = msbgx.Config("empty", 64, 16, 2, msbgx.make_scene(64, spheres=[]), level_mode=-1)
    cfgs.insert(1, empty)
    singles = [prepared(c) for c in cfgs]
    batch = msbgx.BatchGrid.create(singles)
    it_b, rr_b = batch.solve(max_iter=200, tol=1e-6)
    assert it_b[1] == 0 and rr_b[1] == 0.0
    for q in (0, 2):
        it, rr = singles[q].solve_pressure_flip(max_iter=200, tol=1e-6)
        assert it == it_b[q] and rr == rr_b[q]


def test_batch_rejects_single_grid_operations():
    cfgs = small_batch(2, seed=3)
    singles = [prepared(c) for c in cfgs]
    batch = msbgx.BatchGrid.create(singles)
    with pytest.raises(msbgx.MsbgxError):
        batch.solve_pressure_flip()
    with pytest.raises(msbgx.MsbgxError):
        batch.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    # mismatched block size is rejected
    other = msbgx.Config("bsx8", 64, 8, 2, msbgx.make_scene(64, spheres=[(0.5, 0.5, 0.5, 0.2)]), level_mode=-1)
    with pytest.raises(msbgx.MsbgxError):
        msbgx.BatchGrid.create([singles[0], prepared(other)])


@pytest.mark.parametrize("name", ["c1_128", "c4_128_multires"])
def test_launch_choices_keep_results_bit_identical(name):
    """Every smoother kernel and graph chunk gives the same iterates."""
    print(f"launch config invariance {name}", flush=True)
    cfg = msbgx.config_c1(n=128) if name == "c1_128" else msbgx.config_c4(n=128)
    g = prepared(cfg)
    it0, rr0 = g.solve_pressure_flip(max_iter=300, tol=1e-6)
    p0 = pressure(g)
    for chunk, mask in ((1, 0), (8, 0), (4, 0b11111), (16, 0b10)):
        g.set_launch_config(chunk, mask)
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr = g.solve_pressure_flip(max_iter=300, tol=1e-6)
        print(f"  chunk {chunk} mask {mask:05b}: {it} {rr:.3e} ({g.last_solve_ms():.2f} ms)", flush=True)
        assert it == it0 and rr == rr0
        for a, b in zip(p0, pressure(g)):
            np.testing.assert_array_equal(a, b)
    chunk, mask = g.autotune(max_iter=300, tol=1e-6)
    print(f"  autotune: chunk {chunk} mask {mask:05b}", flush=True)
    it, rr = g.solve_pressure_flip(max_iter=300, tol=1e-6)
    assert it == it0 and rr == rr0
    for a, b in zip(p0, pressure(g)):
        np.testing.assert_array_equal(a, b)


def test_batch_autotune_keeps_results():
    print("batch autotune", flush=True)
    cfgs = small_batch(4, seed=5)
    singles = [prepared(c) for c in cfgs]
    batch = msbgx.BatchGrid.create(singles)
    it0, rr0 = batch.solve(max_iter=200, tol=1e-6)
    p0 = pressure(batch)
    chunk, mask = batch.autotune(max_iter=200, tol=1e-6)
    print(f"  autotune: chunk {chunk} mask {mask:05b}", flush=True)
    it, rr = batch.solve(max_iter=200, tol=1e-6)
    np.testing.assert_array_equal(it, it0)
    np.testing.assert_array_equal(rr, rr0)
    for a, b in zip(p0, pressure(batch)):
        np.testing.assert_array_equal(a, b)


def test_full_batch_of_64_problems():
    """The whole batched workload: 64 problems of sizes {64,128,192,256}^3,
    one or two resolutions.  Every problem's iterations, residual and
    pressure equal its own solve's, and a 192^3 and a 256^3 problem (one of
    each resolution kind where the draw has them) match the oracle: same
    iteration count, pressure within 1e-5 relative L2."""
    from oracle_lib import build_oracle
    print("full batch of 64 problems", flush=True)
    cfgs = msbgx.config_c3(seed=1234, n_problems=64)
    mi = 400
    singles = [prepared(c) for c in cfgs]
    single_res = []
    for q, g in enumerate(singles):
        it, rr = g.solve_pressure_flip(max_iter=mi, tol=1e-6)
        single_res.append((it, rr, pressure(g)))
    print("  individual solves done", flush=True)
    batch = msbgx.BatchGrid.create(singles)
    it_b, rr_b = batch.solve(max_iter=mi, tol=1e-6)
    for q, (c, g) in enumerate(zip(cfgs, singles)):
        it, rr, ps = single_res[q]
        assert it == it_b[q] and rr == rr_b[q], f"problem {q} {c.name}: single {it} {rr} batch {it_b[q]} {rr_b[q]}"
        pb = pressure(batch, batch.bases[q], g.n_blocks)
        for b in range(g.n_blocks):
            np.testing.assert_array_equal(ps[b], pb[b], err_msg=f"problem {q} block {b}")
    print(f"  batch equals the individual solves; {int((it_b < mi).sum())}/64 converged", flush=True)
    picks = []
    for n in (192, 256):
        for mode in (-1, 0):
            cand = [q for q, c in enumerate(cfgs) if c.n == n and c.level_mode == mode and it_b[q] < mi]
            if cand:
                picks.append(cand[0])
    assert any(cfgs[q].n == 256 for q in picks) and any(cfgs[q].n == 192 for q in picks)
    for q in picks:
        c = cfgs[q]
        o = build_oracle(c)
        o.prepare_scene(c.scene)
        it_o, rr_o, _ = o.solve(max_iter=mi, tol=1e-6)
        lv, _ = singles[q].block_info(0)
        num = den = 0.0
        pb = pressure(batch, batch.bases[q], singles[q].n_blocks)
        for b in range(singles[q].n_blocks):
            if lv[b] != 0:
                continue
            r = o.get_block(msbgx.CH_PRESSURE, 0, b, c.bsx).astype(np.float64)
            num += float(((pb[b].astype(np.float64) - r) ** 2).sum())
            den += float((r ** 2).sum())
        rel = np.sqrt(num / max(den, 1e-300))
        print(f"  problem {q} {c.name} mode {c.level_mode}: batch {it_b[q]} oracle {it_o}, pressure rel {rel:.2e}",
              flush=True)
        assert it_b[q] == it_o and rel <= 1e-5
        del o
    batch.destroy()
    for g in singles:
        g.destroy()
