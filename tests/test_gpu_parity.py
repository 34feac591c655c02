"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Per-kernel results must be bit-identical (both sides run the same FP32
operation sequence with explicit FMAs); whole solves must reach the same
iteration count and final residual, with pressure within 1e-5 relative L2
(only the FP64 reduction order differs)."""
import numpy as np
import pytest

import msbgx
from parity_util import blocks_of, compare_channel, pair

pytestmark = pytest.mark.gpu

SMALL_CONFIGS = {
    "c1_64": lambda: msbgx.config_c1(n=64),
    "c1_128": lambda: msbgx.config_c1(n=128),
    "c4_128_multires": lambda: msbgx.config_c4(n=128),
    "c2_128_bsx16": lambda: msbgx.config_c2(n=128, bsx=16),
    "c2_256_bsx32": lambda: msbgx.config_c2(n=256, bsx=32),
    "c5_128_particles_bsx32": lambda: msbgx.config_c5(n=128),
}


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", list(SMALL_CONFIGS))
def test_metadata_bit_exact(name):
    _flush(f"metadata {name}")
    cfg = SMALL_CONFIGS[name]()
    g, o = pair(cfg)
    assert g.omega() == o.omega()
    for mg in range(cfg.levels):
        lg, fg = g.block_info(mg)
        lo, fo = o.block_info(mg)
        assert np.array_equal(lg, lo)
        assert np.array_equal(fg, fo), f"block flags differ at MG {mg}"
        assert np.array_equal(g.relax_list(mg), o.relax_list(mg))
        blocks, lv = blocks_of(g, mg)
        for ch in (msbgx.CH_FLAGS, msbgx.CH_FACE_U, msbgx.CH_FACE_V, msbgx.CH_FACE_W, msbgx.CH_DIAGONAL):
            nmis, _, _ = compare_channel(g, o, ch, mg, blocks, lv)
            assert nmis == 0, f"channel {ch} MG {mg}: {nmis} mismatches"
    blocks, lv = blocks_of(g, 0)
    nmis, _, _ = compare_channel(g, o, msbgx.CH_DIVERGENCE, 0, blocks, lv)
    assert nmis == 0


@pytest.mark.parametrize("name", list(SMALL_CONFIGS))
def test_kernels_bit_exact(name):
    """relax / residual / restriction / prolongation, one at a time."""
    _flush(f"kernels {name}")
    cfg = SMALL_CONFIGS[name]()
    g, o = pair(cfg)
    for side in (g, o):
        side.copy_channel(msbgx.CH_DIVERGENCE, msbgx.CH_R)
        side.zero_channel(msbgx.CH_Z, 0)
    blocks, lv = blocks_of(g, 0)
    for side in (g, o):
        side.relax(0, False, 3, msbgx.CH_Z, msbgx.CH_R)
    nmis, md, mv = compare_channel(g, o, msbgx.CH_Z, 0, blocks, lv)
    assert nmis == 0 and mv > 0, f"forward relax: {nmis} mismatches, maxdiff {md}"
    for side in (g, o):
        side.multiply(0, True, msbgx.CH_Z, msbgx.CH_R, msbgx.CH_Q)
    fluid, _ = blocks_of(g, 0, "fluid")
    nmis, md, _ = compare_channel(g, o, msbgx.CH_Q, 0, fluid, lv)
    assert nmis == 0, f"residual: {nmis} mismatches, maxdiff {md}"
    for side in (g, o):
        side.relax(0, True, 2, msbgx.CH_Z, msbgx.CH_R)
    nmis, md, _ = compare_channel(g, o, msbgx.CH_Z, 0, blocks, lv)
    assert nmis == 0, f"reverse relax: {nmis} mismatches, maxdiff {md}"
    for L in range(1, cfg.levels):
        for side in (g, o):
            side.zero_channel(msbgx.CH_R, L) if side is o else None
            side.downsample(L, msbgx.CH_Q, msbgx.CH_R)
            side.zero_channel(msbgx.CH_Z, L)
            side.relax(L, False, 2, msbgx.CH_Z, msbgx.CH_R)
            side.multiply(L, True, msbgx.CH_Z, msbgx.CH_R, msbgx.CH_Q)
        fl_, lvL = blocks_of(g, L, "fluid")
        nmis, md, _ = compare_channel(g, o, msbgx.CH_R, L, fl_, lvL)
        assert nmis == 0, f"restriction to MG {L}: {nmis} mismatches"
        nmis, md, _ = compare_channel(g, o, msbgx.CH_Z, L, fl_, lvL)
        assert nmis == 0, f"relax at MG {L}: {nmis} mismatches"
    for L in range(cfg.levels - 2, -1, -1):
        for side in (g, o):
            side.prolongate(L, msbgx.CH_Z)
        fl_, lvL = blocks_of(g, L, "fluid")
        nmis, md, _ = compare_channel(g, o, msbgx.CH_Z, L, fl_, lvL)
        assert nmis == 0, f"prolongation to MG {L}: {nmis} mismatches"


@pytest.mark.parametrize("name", list(SMALL_CONFIGS))
def test_vcycle_bit_exact(name):
    _flush(f"vcycle {name}")
    cfg = SMALL_CONFIGS[name]()
    g, o = pair(cfg, options=msbgx.OPT_LITERAL_VCYCLE)
    for side in (g, o):
        side.copy_channel(msbgx.CH_DIVERGENCE, msbgx.CH_R)
        side.vcycle(3, 3, 24)
        side.vcycle(3, 3, 24)   # second cycle: coarse rhs of same-level blocks is zero
    blocks, lv = blocks_of(g, 0)
    nmis, md, mv = compare_channel(g, o, msbgx.CH_Z, 0, blocks, lv)
    assert nmis == 0 and mv > 0, f"V-cycle: {nmis} mismatches, maxdiff {md}"


@pytest.mark.parametrize("name", list(SMALL_CONFIGS))
@pytest.mark.parametrize("literal", [False, True])
def test_solve_parity(name, literal):
    _flush(f"solve {name} literal={literal}")
    cfg = SMALL_CONFIGS[name]()
    g, o = pair(cfg, options=msbgx.OPT_LITERAL_VCYCLE if literal else 0)
    max_iter = 200
    it_g, rr_g = g.solve_pressure_flip(max_iter=max_iter, tol=1e-6)
    it_o, rr_o, hist_o = o.solve(max_iter=max_iter, tol=1e-6)
    _flush(f"  gpu iters {it_g} rel {rr_g:.3e} | oracle iters {it_o} rel {rr_o:.3e} | {g.last_solve_ms():.2f} ms")
    assert it_g == it_o
    assert abs(rr_g - rr_o) <= 1e-3 * rr_o + 1e-12
    hist_g = g.residual_history()
    assert np.allclose(hist_g, hist_o[:len(hist_g)], rtol=1e-6)
    blocks, lv = blocks_of(g, 0)
    num = den = 0.0
    for b in blocks:
        if lv[b] != 0:
            continue
        a = g.get_block(msbgx.CH_PRESSURE, 0, int(b), cfg.bsx).astype(np.float64)
        r = o.get_block(msbgx.CH_PRESSURE, 0, int(b), cfg.bsx).astype(np.float64)
        num += float(((a - r) ** 2).sum())
        den += float((r ** 2).sum())
    rel = np.sqrt(num / max(den, 1e-300))
    _flush(f"  pressure rel L2 {rel:.3e}")
    assert rel <= 1e-5
