"""GPU parity at the workloads' own sizes (and the largest sizes the oracle
finishes in seconds for the configurations it cannot solve whole).

* 256^3 sphere, 16^3 blocks, single level: the whole solve against the
  oracle -- same iteration count and final residual, pressure within 1e-5
  relative L2.
* 512^3 slab + 64 spheres, 32^3 base blocks, 3 levels: metadata, one relax
  pass (FP32 and FP16/BF16 smoother operands) and one residual at MG 0, bit
  for bit.
* the same scene at 256^3 with FP16 / BF16 smoother operands: the refined
  solve against the oracle's FP64 solution of the level-0 system.
* 1024^3 particle workload's scene at 256^3 (32^3 blocks, 4 levels), FP16
  smoother operands, learned coarse correction: converges to 1e-6 and its
  pressure is as close to the FP64 solution as the plain solve's.
"""
import numpy as np
import pytest

import msbgx
from parity_util import blocks_of, compare_channel, pair

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _pressure_rel(g, o, cfg):
    lv, _ = g.block_info(0)
    num = den = 0.0
    for b in range(g.n_blocks):
        if lv[b] != 0:
            continue
        a = g.get_block(msbgx.CH_PRESSURE, 0, b, cfg.bsx).astype(np.float64)
        r = o.get_block(msbgx.CH_PRESSURE, 0, b, cfg.bsx).astype(np.float64)
        num += float(((a - r) ** 2).sum())
        den += float((r ** 2).sum())
    return np.sqrt(num / max(den, 1e-300))


def _err_vs(g, x_ref, cfg):
    num = den = 0.0
    for b, v in x_ref.items():
        a = g.get_block(msbgx.CH_PRESSURE, 0, b, cfg.bsx).astype(np.float64)
        num += float(((a - v) ** 2).sum())
        den += float((v ** 2).sum())
    return np.sqrt(num / max(den, 1e-300))


def test_sphere_256_whole_solve_matches_oracle():
    cfg = msbgx.config_c1(256)
    _flush(f"full-size solve {cfg.name} n={cfg.n}")
    g, o = pair(cfg)
    it_g, rr_g = g.solve_pressure_flip(max_iter=200, tol=1e-6)
    it_o, rr_o, hist_o = o.solve(max_iter=200, tol=1e-6)
    _flush(f"  gpu {it_g} iterations rel {rr_g:.3e} ({g.last_solve_ms():.1f} ms) | oracle {it_o} rel {rr_o:.3e}")
    assert it_g == it_o and rr_g <= 1e-6
    assert abs(rr_g - rr_o) <= 1e-3 * rr_o
    hist_g = g.residual_history()
    np.testing.assert_allclose(hist_g, hist_o[:len(hist_g)], rtol=1e-6)
    rel = _pressure_rel(g, o, cfg)
    _flush(f"  pressure rel L2 {rel:.2e}")
    assert rel <= 1e-5
    g.destroy()


@pytest.mark.parametrize("prec", [msbgx.PREC_FP32, msbgx.PREC_FP16, msbgx.PREC_BF16], ids=["fp32", "fp16", "bf16"])
def test_slab_spheres_512_kernels_bit_exact(prec):
    cfg = msbgx.config_c2(512, bsx=32)
    _flush(f"full-size kernels {cfg.name} n={cfg.n} prec={prec}")
    g, o = pair(cfg)
    assert g.omega() == o.omega()
    lg, fg = g.block_info(0)
    lo, fo = o.block_info(0)
    assert np.array_equal(lg, lo) and np.array_equal(fg, fo)
    assert np.array_equal(g.relax_list(0), o.relax_list(0))
    blocks, lv = blocks_of(g, 0, "relax")
    nmis, _, _ = compare_channel(g, o, msbgx.CH_DIAGONAL, 0, blocks, lv)
    assert nmis == 0
    nmis, _, _ = compare_channel(g, o, msbgx.CH_DIVERGENCE, 0, blocks, lv)
    assert nmis == 0
    g.set_precision(prec)
    o.set_smoother_precision(prec)
    for side in (g, o):
        side.copy_channel(msbgx.CH_DIVERGENCE, msbgx.CH_R)
        side.zero_channel(msbgx.CH_Z, 0)
        side.relax(0, False, 1, msbgx.CH_Z, msbgx.CH_R)
    nmis, md, mv = compare_channel(g, o, msbgx.CH_Z, 0, blocks, lv)
    _flush(f"  relax over {len(blocks)} blocks: {nmis} mismatches, max |z| {mv:.3e}")
    assert nmis == 0 and mv > 0
    for side in (g, o):
        side.multiply(0, True, msbgx.CH_Z, msbgx.CH_R, msbgx.CH_Q)
    fluid, _ = blocks_of(g, 0, "fluid")
    nmis, md, _ = compare_channel(g, o, msbgx.CH_Q, 0, fluid, lv)
    _flush(f"  residual over {len(fluid)} blocks: {nmis} mismatches")
    assert nmis == 0
    g.destroy()


def test_slab_spheres_256_refined_16bit_matches_fp64_oracle():
    cfg = msbgx.config_c2(256, bsx=32)
    _flush(f"refined 16-bit {cfg.name} n={cfg.n}")
    g, o = pair(cfg)
    it_ref, rel_ref, x_ref = o.solve_fp64(tol=1e-13, max_iter=50000)
    _flush(f"  oracle FP64: {it_ref} iterations, rel {rel_ref:.2e}")
    assert rel_ref <= 1e-12
    for prec in (msbgx.PREC_FP16, msbgx.PREC_BF16):
        g.set_precision(prec)
        outer, inner, rel, xb = g.solve_pressure_refined(tol=1e-10, max_outer=30, inner_tol=1e-5, want_x=True)
        num = np.sqrt(sum(float(((xb.get(b, 0.0) - v) ** 2).sum()) for b, v in x_ref.items()))
        den = np.sqrt(sum(float((v ** 2).sum()) for v in x_ref.values()))
        _flush(f"  prec={prec}: outer {outer} inner {inner} FP64 rel residual {rel:.2e}, "
               f"error vs FP64 oracle {num / den:.2e}")
        assert rel <= 1e-10 and num / den <= 1e-8
    g.destroy()


def test_particle_scene_256_fp16_learned_coarse():
    cfg = msbgx.config_c5(256)
    _flush(f"FP16 + learned coarse {cfg.name} n={cfg.n} bsx={cfg.bsx} levels={cfg.levels}")
    g, o = pair(cfg)
    _, rel_ref, x_ref = o.solve_fp64(tol=1e-12, max_iter=50000)
    assert rel_ref <= 1e-11
    g.set_precision(msbgx.PREC_FP16)
    it_p, rr_p = g.solve_pressure_flip(max_iter=1000, tol=1e-6)
    e_p = _err_vs(g, x_ref, cfg)
    net = msbgx.LearnedNet(seed=3, mode=msbgx.NET_COARSE)
    net.prolongation = msbgx.PROLONG_LINEAR
    net.scales = 3
    net.collect(g, max_pairs=200, seed=5, tol=1e-7, max_iter=1000)
    net.train(300, batch=32, lr=1e-3, seed=1)
    g.zero_channel(msbgx.CH_PRESSURE, 0)
    it_l, rr_l, fb = net.solve(g, max_iter=1000, tol=1e-6)
    e_l = _err_vs(g, x_ref, cfg)
    _flush(f"  plain {it_p} iterations rel {rr_p:.2e} error {e_p:.2e} | learned {it_l} rel {rr_l:.2e} "
           f"fallback {fb} error {e_l:.2e}")
    assert rr_p <= 1e-6 and rr_l <= 1e-6
    # both stop at the same relative residual, so their errors are of one size
    assert e_l <= 10 * e_p + 1e-7
    net.destroy()
    g.destroy()
