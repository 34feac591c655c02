"""The drop-in boundary as the reference's caller uses it, and the captured
PCG graphs' keys.

* preparePressureSolveFLIP(const FlipPressureCallbacks&) (src/msbg.h:829-853,
  src/msbg3.cpp:2472-2583) through msbgx_prepare_pressure_solve with ctypes
  CFUNCTYPE callbacks: metadata at every MG level, relax lists and the solve
  are bit-identical to the dense-array prepare and to the oracle; the user
  pointer reaches every callback unchanged.
* The C++ facade include/msbgx_facade.hpp, compiled into
  msbg-flip_amd/facade_check and run as a separate program: the same
  pressure bit for bit, its error behaviour (create -> nullptr on invalid
  arguments, solve before prepare -> state error).
* Captured graphs: a solve that grows max_iter (the history buffer is
  reallocated), changes nPre/nPost/nCoarse, toggles MSBGX_OPT_LITERAL_VCYCLE
  or alternates plain and learned solves on one handle gives the iterates of
  the same sequence run without graphs.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import msbgx
import oracle_lib as orc
from parity_util import blocks_of, compare_channel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "msbg-flip_amd", "facade_check")
USER_TOKEN = 0x5EED1234


def dense_inputs(n, bsx, levels, multires, seed):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    fluid = yy < 0.3 * n
    for _ in range(2):
        c = rng.uniform(0.25, 0.75, 3) * n
        r = rng.uniform(0.12, 0.2) * n
        fluid |= (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r
    flags = np.where(fluid, 0, msbgx.CELL_VOID).astype(np.uint16)
    solid = (np.abs(xx - 0.5 * n) < 0.08 * n) & (yy < 0.2 * n) & (np.abs(zz - 0.5 * n) < 0.08 * n)
    flags[solid] = msbgx.CELL_SOLID
    w = [rng.uniform(0.5, 1.5, flags.shape).astype(np.float32) for _ in range(3)]
    rhs = np.where(flags == 0, rng.normal(0, 1, flags.shape), 0).astype(np.float32)
    nb = (n // bsx) ** 3
    lv = (rng.choice([0, levels - 1], size=nb, p=[0.6, 0.4]) if multires else np.zeros(nb)).astype(np.int32)
    return flags, w, rhs, lv


def make_callbacks(n, flags, w, rhs, bad):
    """CFUNCTYPE callbacks over the dense arrays, sampled at the level-local
    coordinates they receive (as the dense prepare does)."""
    fl = flags.ravel().tolist()
    wl = [a.ravel().tolist() for a in w]
    rl = rhs.ravel().tolist()
    ct, ft, rt = (f[1] for f in msbgx.Callbacks._fields_[:3])

    def cell(user, x, y, z):
        if user != USER_TOKEN:
            bad[0] += 1
        return fl[x + n * (y + n * z)]

    def face(user, d, x, y, z):
        if user != USER_TOKEN:
            bad[0] += 1
        return wl[d][x + n * (y + n * z)]

    def rhs_cb(user, x, y, z):
        if user != USER_TOKEN:
            bad[0] += 1
        return rl[x + n * (y + n * z)]

    keep = [ct(cell), ft(face), rt(rhs_cb)]
    cb = msbgx.Callbacks(keep[0], keep[1], keep[2], C.c_void_p(USER_TOKEN))
    return cb, keep


def grid_for(n, bsx, levels, lv, options=0):
    g = msbgx.MultiresSparseGrid.create(n, n, n, bsx, levels)
    g.set_domain_boundary_spec(*([msbgx.DBC_SOLID] * 6))
    lg = g.regularize_refinement_map(lv)
    g.set_refinement_map(lg)
    if options:
        g.set_options(options)
    return g, lg


def level0_pressure(g, bsx):
    lv, _ = g.block_info(0)
    return {b: g.get_block(msbgx.CH_PRESSURE, 0, b, bsx) for b in range(g.n_blocks) if lv[b] == 0}


CB_CASES = [(64, 16, 3, False), (64, 16, 3, True), (64, 32, 2, True)]


@pytest.mark.parametrize("n,bsx,levels,multires", CB_CASES)
def test_callback_prepare_matches_dense_and_oracle(n, bsx, levels, multires):
    print(f"callback prepare n={n} bsx={bsx} levels={levels} multires={multires}", flush=True)
    flags, w, rhs, lv = dense_inputs(n, bsx, levels, multires, seed=n + bsx + levels + int(multires))
    bad = [0]
    cb, keep = make_callbacks(n, flags, w, rhs, bad)
    gc, lg = grid_for(n, bsx, levels, lv)
    gc.prepare_pressure_solve_flip(callbacks=cb)
    gd, _ = grid_for(n, bsx, levels, lv)
    gd.prepare_pressure_solve_flip(dense=(flags, *w, rhs))
    o = orc.Oracle(n, n, n, bsx, levels)
    o.set_domain_boundary_spec(*([msbgx.DBC_SOLID] * 6))
    o.set_refinement_map(lg)
    o.prepare_dense((n, n, n), flags, *w, rhs)
    assert bad[0] == 0, f"{bad[0]} callbacks received another user pointer"
    assert gc.omega() == gd.omega() == o.omega()
    for mg in range(levels):
        for other in (gd, o):
            lo, fo = other.block_info(mg)
            lc, fc = gc.block_info(mg)
            assert np.array_equal(lc, lo) and np.array_equal(fc, fo), f"block info MG {mg}"
            assert np.array_equal(gc.relax_list(mg), other.relax_list(mg)), f"relax list MG {mg}"
            blocks, lvb = blocks_of(gc, mg)
            for ch in (msbgx.CH_FLAGS, msbgx.CH_FACE_U, msbgx.CH_FACE_V, msbgx.CH_FACE_W, msbgx.CH_DIAGONAL):
                nmis, _, _ = compare_channel(gc, other, ch, mg, blocks, lvb)
                assert nmis == 0, f"channel {ch} MG {mg}: {nmis} mismatches"
    blocks, lvb = blocks_of(gc, 0)
    nmis, _, mv = compare_channel(gc, o, msbgx.CH_DIVERGENCE, 0, blocks, lvb)
    assert nmis == 0 and mv > 0
    max_iter, tol = (30, 1e-12) if multires else (300, 1e-6)
    it_c, rr_c = gc.solve_pressure_flip(max_iter=max_iter, tol=tol)
    it_d, rr_d = gd.solve_pressure_flip(max_iter=max_iter, tol=tol)
    it_o, rr_o, _ = o.solve(max_iter=max_iter, tol=tol)
    pc, pd = level0_pressure(gc, bsx), level0_pressure(gd, bsx)
    assert it_c == it_d == it_o and rr_c == rr_d
    for b in pc:
        assert np.array_equal(pc[b].view(np.uint32), pd[b].view(np.uint32)), f"block {b}"
    num = sum(float(((pc[b].astype(np.float64) - o.get_block(msbgx.CH_PRESSURE, 0, b, bsx)) ** 2).sum()) for b in pc)
    den = sum(float((o.get_block(msbgx.CH_PRESSURE, 0, b, bsx).astype(np.float64) ** 2).sum()) for b in pc)
    rel = np.sqrt(num / max(den, 1e-300))
    print(f"  callbacks {it_c} iterations rel {rr_c:.2e}; oracle {it_o} rel {rr_o:.2e}; pressure rel L2 {rel:.1e}",
          flush=True)
    assert rel <= 1e-5


@pytest.mark.parametrize("multires", [False, True])
def test_cpp_facade_program(tmp_path, multires):
    assert os.path.exists(FACADE), "build() compiles msbg-flip_amd/facade_check"
    n, bsx, levels = 64, 16, 3
    flags, w, rhs, lv = dense_inputs(n, bsx, levels, multires, seed=77 + int(multires))
    max_iter, tol = (30, 1e-12) if multires else (300, 1e-6)
    src = tmp_path / "in.bin"
    with open(src, "wb") as f:
        np.array([n, n, n, bsx, levels, 3, 3, 24, max_iter], np.int32).tofile(f)
        np.array([tol], np.float32).tofile(f)
        lv.astype(np.int32).tofile(f)
        flags.tofile(f)
        for a in (*w, rhs):
            a.tofile(f)
    dst = tmp_path / "out.bin"
    print(f"facade program multires={multires}", flush=True)
    r = subprocess.run([FACADE, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(dst, np.uint8)
    it_f = int(raw[:4].view(np.int32)[0])
    rr_f = float(raw[4:8].view(np.float32)[0])
    bad_user, failed = raw[8:16].view(np.int32)
    p_f = raw[16:].view(np.float32).reshape(-1, bsx ** 3)
    assert bad_user == 0 and failed == 0
    gd, _ = grid_for(n, bsx, levels, lv)
    gd.prepare_pressure_solve_flip(dense=(flags, *w, rhs))
    it_d, rr_d = gd.solve_pressure_flip(max_iter=max_iter, tol=tol)
    assert it_f == it_d and rr_f == rr_d
    pd = level0_pressure(gd, bsx)
    assert len(pd) == int((gd.block_info(0)[0] == 0).sum()) > 0
    for b, a in pd.items():
        assert np.array_equal(p_f[b].view(np.uint32), a.view(np.uint32)), f"block {b}"


def solve_seq(g, steps):
    out = []
    for kind, kw in steps:
        if kind == "plain":
            it, rr = g.solve_pressure_flip(**kw)
            out.append((it, rr, None, g.residual_history().copy(), level0_pressure(g, g.bsx)))
        elif kind == "opts":
            g.set_options(kw)
        else:
            it, rr, fb = kind.solve(g, **kw)
            out.append((it, rr, fb, g.residual_history().copy(), level0_pressure(g, g.bsx)))
    return out


def assert_same(a, b):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert x[0] == y[0] and x[1] == y[1] and x[2] == y[2], f"solve {k}: {x[:3]} vs {y[:3]}"
        assert np.array_equal(x[3], y[3]), f"solve {k}: residual history differs"
        for blk in x[4]:
            assert np.array_equal(x[4][blk].view(np.uint32), y[4][blk].view(np.uint32)), f"solve {k} block {blk}"


def test_graph_recaptured_when_history_grows_and_params_change():
    cfg = msbgx.config_c1(n=64)
    steps = [("plain", dict(max_iter=12, tol=1e-12)),
             ("plain", dict(max_iter=400, tol=1e-6)),
             ("plain", dict(n_pre=1, n_post=2, max_iter=400, tol=1e-6)),
             ("plain", dict(n_pre=2, n_post=2, n_coarse=8, max_iter=2000, tol=1e-7)),
             ("opts", msbgx.OPT_LITERAL_VCYCLE),
             ("plain", dict(max_iter=400, tol=1e-6))]
    runs = []
    for opts in (0, msbgx.OPT_NO_GRAPH):
        g = msbgx.build_grid(cfg)
        g.set_options(opts)
        g.prepare_pressure_solve_flip(scene=cfg.scene)
        seq = [(k, (v | opts) if k == "opts" else v) for k, v in steps]
        runs.append(solve_seq(g, seq))
        print("  iterations", [r[0] for r in runs[-1]], flush=True)
    assert_same(*runs)
    assert runs[0][1][0] != runs[0][2][0]   # the smoothing counts changed the iterates


def test_graph_keys_across_plain_and_learned_solves():
    cfg = msbgx.config_c4(n=64)
    runs = []
    for opts in (0, msbgx.OPT_NO_GRAPH):
        net = msbgx.LearnedNet(seed=5, mode=msbgx.NET_COARSE)
        net.set_stall(0, 0.0)
        g = msbgx.build_grid(cfg)
        g.set_options(opts)
        g.prepare_pressure_solve_flip(scene=cfg.scene)
        steps = [("plain", dict(max_iter=20, tol=1e-12)),
                 (net, dict(max_iter=60, tol=1e-6)),
                 ("plain", dict(max_iter=300, tol=1e-6)),
                 (net, dict(n_pre=2, n_post=2, max_iter=500, tol=1e-7))]
        runs.append(solve_seq(g, steps))
        print("  iterations", [r[0] for r in runs[-1]], flush=True)
        net.destroy()
    assert_same(*runs)
