"""Compact MIC(0)-PCG comparator on the GPU (src/flip_sim.cpp:858-1151).

* Random liquid, face coefficients and divergence, single-level and
  multires, 16^3 and 32^3 blocks: the MIC(0) factor and one application
  M^-1 rhs (hyperplane wavefronts) are bit-identical to the oracle's
  sequential loops; the full solve takes the same iteration count (+-1: the
  FP64 dots are summed in another order) and its solution agrees to 1e-5.
* 256^3 sphere scene: the compact solution is minus the MSBG solution (the
  two paths solve A p = -div/dt and A p = +div/dt with the same operator on a
  single-level grid), to the solver tolerances.
"""
import numpy as np
import pytest

import msbgx
import oracle_lib as orc

pytestmark = pytest.mark.gpu

EPS = 1e-6


def dense_case(n, bsx, levels, seed, multires):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    mass = np.zeros((n, n, n), np.float32)
    for _ in range(3):
        c = rng.uniform(0.25, 0.75, 3) * n
        r = rng.uniform(0.1, 0.25) * n
        mass[(xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r] = 1.0
    mass[:, : n // 6] = 1.0                                # a pool along y = 0
    flags = np.where(mass > EPS, 0, msbgx.CELL_VOID).astype(np.uint16)
    beta = [rng.uniform(0.5, 1.5, mass.shape).astype(np.float32) for _ in range(3)]
    div = np.where(mass > EPS, rng.normal(0, 1, mass.shape), 0).astype(np.float32)
    g = msbgx.MultiresSparseGrid.create(n, n, n, bsx, levels)
    g.set_domain_boundary_spec(*([msbgx.DBC_SOLID] * 6))
    nb = (n // bsx) ** 3
    lv = rng.choice([0, levels - 1], size=nb, p=[0.6, 0.4]).astype(np.int32) if multires else np.zeros(nb, np.int32)
    g.set_refinement_map(g.regularize_refinement_map(lv))
    g.prepare_pressure_solve_flip(dense=(flags, *beta, div))
    level, _ = g.block_info(0)
    return g, level, mass, beta, div


def dense_pressure(g, level, n, bsx):
    out = np.zeros((n, n, n), np.float32)
    nbx = n // bsx
    for b in range(len(level)):
        if level[b] != 0:
            continue
        bx, by, bz = b % nbx, (b // nbx) % nbx, b // (nbx * nbx)
        out[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = \
            g.get_block(msbgx.CH_PRESSURE, 0, b, bsx).reshape(bsx, bsx, bsx)
    return out


@pytest.mark.parametrize("n,bsx,levels,multires", [(64, 16, 3, False), (64, 16, 3, True), (64, 32, 2, True)])
def test_compact_matches_oracle(n, bsx, levels, multires):
    print(f"compact MIC(0)-PCG vs oracle: n={n} bsx={bsx} multires={multires}", flush=True)
    g, level, mass, beta, div = dense_case(n, bsx, levels, seed=n + bsx + multires, multires=multires)
    pre, z = g.compact_mic0()
    xo, ito, rro, pre_o, z_o = orc.compact_pcg(level, bsx, mass, EPS, beta, div, tol=1e-6, max_iter=1000,
                                               want_factor=True)
    assert len(pre) == len(pre_o) > 0
    np.testing.assert_array_equal(pre.view(np.uint32), pre_o.view(np.uint32))
    np.testing.assert_array_equal(z.view(np.uint32), z_o.view(np.uint32))
    print(f"  {len(pre)} rows: factor and M^-1 rhs bit-identical", flush=True)
    it, rr = g.solve_pressure_compact(max_iter=1000, tol=1e-6)
    x = dense_pressure(g, level, n, bsx)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"  solve: {it} iterations (oracle {ito}), rel {rr:.2e}; |x - x_oracle| / |x_oracle| = {err:.2e}, "
          f"{g.last_solve_ms():.1f} ms", flush=True)
    assert abs(it - ito) <= 1 and rr <= 1e-6 and err <= 1e-5


def test_compact_is_minus_msbg_at_256():
    cfg = msbgx.config_c1()
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    blocks = g.level0_blocks()
    g.solve_pressure_refined(tol=1e-10)
    pm = np.concatenate([g.get_block(msbgx.CH_PRESSURE, 0, int(b), cfg.bsx) for b in blocks]).astype(np.float64)
    it, rr = g.solve_pressure_compact(max_iter=2000, tol=1e-8)
    ms = g.last_solve_ms()
    pc = np.concatenate([g.get_block(msbgx.CH_PRESSURE, 0, int(b), cfg.bsx) for b in blocks]).astype(np.float64)
    err = np.linalg.norm(pc + pm) / np.linalg.norm(pm)
    print(f"compact at 256^3: {it} iterations to {rr:.1e} in {ms:.0f} ms; |p_compact + p_msbg| / |p_msbg| = {err:.2e}",
          flush=True)
    assert rr <= 1e-8 and err <= 1e-5
