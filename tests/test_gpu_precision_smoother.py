"""Precision form: the 16-bit-operand smoother on the GPU is bit-identical to
the oracle's 16-bit relax (operands rounded to FP16 / BF16 with
round-to-nearest-even, FP32 arithmetic), for the column kernels of 16^3 and
smaller blocks and for both 32^3 kernels (register-resident and
record-streaming), on single-level and multires grids; and the MG-PCG with
16-bit smoother operands takes the oracle's iterations with the same pressure."""
import numpy as np
import pytest

import msbgx
import oracle_lib as orc
from parity_util import blocks_of, compare_channel

pytestmark = pytest.mark.gpu

CASES = [("c1_64", lambda: msbgx.config_c1(n=64)),
         ("c4_128_multires", lambda: msbgx.config_c4(n=128)),
         ("c2_128_32", lambda: msbgx.config_c2(n=128, bsx=32)),
         ("c1_128_32", lambda: msbgx.Config("c1_128_b32", 128, 32, 3, msbgx.config_c1(n=128).scene, level_mode=-1)),
         ("c5_128_32", lambda: msbgx.config_c5(n=128))]


def pair(cfg, prec, variant=0):
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    g.set_precision(prec)
    if variant:
        g.set_launch_config(4, variant)
    o = orc.build_oracle(cfg)
    o.prepare_scene(cfg.scene)
    o.set_smoother_precision(prec)
    return g, o


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("prec", [msbgx.PREC_FP16, msbgx.PREC_BF16], ids=["fp16", "bf16"])
def test_16bit_relax_bit_exact(name, make, prec):
    cfg = make()
    variants = [0, 1] if cfg.bsx == 32 else [0]
    for variant in variants:
        print(f"16-bit relax {name} prec={prec} variant={variant}", flush=True)
        g, o = pair(cfg, prec, variant)
        for side in (g, o):
            side.copy_channel(msbgx.CH_DIVERGENCE, msbgx.CH_R)
            side.zero_channel(msbgx.CH_Z, 0)
            side.relax(0, False, 2, msbgx.CH_Z, msbgx.CH_R)
        blocks, lv = blocks_of(g, 0)
        nmis, md, mv = compare_channel(g, o, msbgx.CH_Z, 0, blocks, lv)
        assert nmis == 0 and mv > 0, f"forward relax: {nmis} mismatches (max diff {md})"
        for side in (g, o):
            side.relax(0, True, 1, msbgx.CH_Z, msbgx.CH_R)
        nmis, md, _ = compare_channel(g, o, msbgx.CH_Z, 0, blocks, lv)
        assert nmis == 0, f"reverse relax: {nmis} mismatches (max diff {md})"
        # one coarse level: restrict the residual and relax there
        for side in (g, o):
            side.multiply(0, True, msbgx.CH_Z, msbgx.CH_R, msbgx.CH_Q)
            if side is o:
                side.zero_channel(msbgx.CH_R, 1)
            side.downsample(1, msbgx.CH_Q, msbgx.CH_R)
            side.zero_channel(msbgx.CH_Z, 1)
            side.relax(1, False, 2, msbgx.CH_Z, msbgx.CH_R)
        fl1, lv1 = blocks_of(g, 1, "fluid")
        nmis, md, _ = compare_channel(g, o, msbgx.CH_Z, 1, fl1, lv1)
        assert nmis == 0, f"MG 1 relax: {nmis} mismatches (max diff {md})"
        g.destroy()


@pytest.mark.parametrize("name,make", [CASES[0], CASES[2], CASES[4]], ids=["c1_64", "c2_128_32", "c5_128_32"])
def test_16bit_solve_matches_oracle(name, make):
    cfg = make()
    multires = cfg.level_mode >= 0
    max_iter, tol = (30, 1e-12) if multires else (300, 1e-6)
    g, o = pair(cfg, msbgx.PREC_FP16)
    it_g, rr_g = g.solve_pressure_flip(max_iter=max_iter, tol=tol)
    it_o, rr_o, _ = o.solve(max_iter=max_iter, tol=tol)
    lv, _ = g.block_info(0)
    num = den = 0.0
    for b in range(g.n_blocks):
        if lv[b] != 0:
            continue
        a = g.get_block(msbgx.CH_PRESSURE, 0, b, cfg.bsx).astype(np.float64)
        r = o.get_block(msbgx.CH_PRESSURE, 0, b, cfg.bsx).astype(np.float64)
        num += float(((a - r) ** 2).sum())
        den += float((r ** 2).sum())
    rel = np.sqrt(num / max(den, 1e-300))
    print(f"16-bit solve {name}: gpu {it_g} it rel {rr_g:.3e}; oracle {it_o} it rel {rr_o:.3e}; "
          f"pressure rel L2 {rel:.2e}", flush=True)
    assert it_g == it_o and abs(rr_g - rr_o) <= 1e-3 * rr_o + 1e-12 and rel <= 1e-5


@pytest.mark.parametrize("bsx0", [8, 16, 32])
def test_16bit_vcycle_fp64_coarsest_bit_exact(bsx0):
    """Precision form: the V-cycle's coarsest "solve" (24 relax iterations)
    runs in FP64 on both sides.  Under the reference's coarse metadata (Q1)
    the workloads' coarsest levels relax no block, so the grid here is built
    for it: a 4x4x4 raster, two MG levels, block 0 at level 0 and the others
    at level 1, OPEN walls, dense inputs -- the coarsest level relaxes the
    blocks whose doubled-coordinate samples reach level-1 data, and its
    right-hand side is the restricted MG 0 residual (coarsest block edges 4,
    8 and 16).  Two V-cycles give bit-identical z at both levels."""
    n = 4 * bsx0
    print(f"16-bit V-cycle with FP64 coarsest level, bsx0 {bsx0}", flush=True)
    rng = np.random.default_rng(5 + bsx0)
    flags = np.zeros((n, n, n), np.uint16)
    w = [rng.uniform(0.5, 1.5, (n, n, n)).astype(np.float32) for _ in range(3)]
    rhs = rng.standard_normal((n, n, n)).astype(np.float32)
    lv = np.ones(64, np.int32)
    lv[0] = 0
    g = msbgx.MultiresSparseGrid.create(n, n, n, bsx0, 2)
    g.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    g.set_refinement_map(lv)
    g.prepare_pressure_solve_flip(dense=(flags, *w, rhs))
    g.set_precision(msbgx.PREC_FP16)
    g.set_options(msbgx.OPT_LITERAL_VCYCLE)
    o = orc.Oracle(n, n, n, bsx0, 2)
    o.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    o.set_refinement_map(lv)
    o.prepare_dense((n, n, n), flags, *w, rhs)
    o.set_smoother_precision(msbgx.PREC_FP16)
    assert len(g.relax_list(1)) > 0 and np.array_equal(g.relax_list(1), o.relax_list(1))
    for side in (g, o):
        side.copy_channel(msbgx.CH_DIVERGENCE, msbgx.CH_R)
        side.vcycle(3, 3, 24)
        side.vcycle(3, 3, 24)
    for mg in (1, 0):
        blocks, lvb = blocks_of(g, mg, "fluid")
        nmis, md, mv = compare_channel(g, o, msbgx.CH_Z, mg, blocks, lvb)
        assert nmis == 0 and mv > 0, f"MG {mg}: {nmis} mismatches (max diff {md})"
    g.destroy()
