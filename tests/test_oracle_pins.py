"""Pins of the CPU oracle against what the mathematics fixes (CPU only).

The reference ships no tests, fixtures or golden vectors for this path, so
each oracle function is checked here against closed forms, invariants, an
independent brute-force restatement on tiny inputs, or dense linear algebra:

  topology     hand-derived regularisation cases, 2:1 grading, idempotence
  operator     A(linear) = 0 and A(x^2) = -2h in the interior; A(1) counts
               Dirichlet neighbours; <Au,v> = <u,Av> on single-level grids
  smoother     pure-Python restatement of the 5-sweep segment schedule on one
               block; fixed point of A x = b; D = 0 cells untouched
  transfer     exact 2x2x2 means / parent copies on integer-valued data
  PCG          converged solution equals a dense numpy solve of the
               assembled operator
"""
import numpy as np
import pytest

import msbgx
from oracle_lib import Oracle, build_oracle, level_map_from_scene

CH_X, CH_B, CH_R, CH_Z, CH_D, CH_Q = 0, 1, 2, 3, 4, 5


def dense_problem(n, bsx, levels=None, void=None, w=None, rhs=None, bc=1, seed=0):
    """Oracle over dense inputs: every cell fluid unless `void` marks it."""
    if levels is None:
        levels = min(3, int(np.log2(bsx)))
    shape = (n, n, n)  # (z, y, x)
    flags = np.zeros(shape, np.uint16)
    if void is not None:
        flags[void] = msbgx.CELL_VOID
    ws = [np.ones(shape, np.float32) if w is None else w[k] for k in range(3)]
    r = np.zeros(shape, np.float32) if rhs is None else rhs
    o = Oracle(n, n, n, bsx, levels, threads=4)
    o.set_domain_boundary_spec(*([bc] * 6))
    o.set_refinement_map(np.zeros(o.n_blocks, np.int32))
    o.prepare_dense((n, n, n), flags, *ws, r)
    return o


def field_to_blocks(o, f, n, bsx, ch, mg=0):
    nb = n // bsx
    for bid in range(o.n_blocks):
        bx, by, bz = bid % nb, (bid // nb) % nb, bid // (nb * nb)
        blk = f[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx]
        o.set_block(ch, mg, bid, blk.reshape(-1))


def blocks_to_field(o, n, bsx, ch, mg=0):
    nb = n // bsx
    f = np.zeros((n, n, n), np.float32)
    for bid in range(o.n_blocks):
        bx, by, bz = bid % nb, (bid // nb) % nb, bid // (nb * nb)
        blk = o.get_block(ch, mg, bid, bsx)
        if blk is not None:
            f[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = blk.reshape(bsx, bsx, bsx)
    return f


# ---------------------------------------------------------------- topology

def test_regularize_single_fine_block():
    """4x4x4 raster, 3 levels: one level-0 block at the corner makes exactly its
    26-neighbours level 1 (src/msbg.cpp:1202-1420)."""
    o = Oracle(64, 64, 64, 16, 3)
    lv = np.full(64, 2, np.int32)
    lv[0] = 0
    out = o.regularize_refinement_map(lv).reshape(4, 4, 4)
    exp = np.full((4, 4, 4), 2)
    exp[:2, :2, :2] = 1
    exp[0, 0, 0] = 0
    assert np.array_equal(out, exp)


def test_regularize_four_levels_chain():
    """With 4 levels a level-0 block grades 0 -> 1 -> 2 -> 3 along a line."""
    o = Oracle(8 * 16, 16, 16, 16, 4)
    lv = np.full(8, 3, np.int32)
    lv[0] = 0
    out = o.regularize_refinement_map(lv)
    assert list(out) == [0, 1, 2, 3, 3, 3, 3, 3]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_regularize_grading_and_idempotence(seed):
    rng = np.random.default_rng(seed)
    o = Oracle(96, 96, 96, 16, 4)
    lv = rng.choice([0, 3], size=o.n_blocks, p=[0.1, 0.9]).astype(np.int32)
    out = o.regularize_refinement_map(lv).reshape(6, 6, 6)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                a = out[max(0, dz):6 + min(0, dz), max(0, dy):6 + min(0, dy), max(0, dx):6 + min(0, dx)]
                b = out[max(0, -dz):6 + min(0, -dz), max(0, -dy):6 + min(0, -dy), max(0, -dx):6 + min(0, -dx)]
                assert np.abs(a - b).max() <= 1
    assert np.array_equal(o.regularize_refinement_map(out.reshape(-1)), out.reshape(-1))
    assert np.all(out.reshape(-1) <= lv)


def test_block_levels_and_border_flags():
    """level = max(map, L); FINE_COARSE/COARSE_FINE over the 26-neighbourhood."""
    o = Oracle(64, 64, 64, 16, 3)
    lv = np.full(64, 2, np.int32)
    lv[0] = 0
    lv = o.regularize_refinement_map(lv)
    o.set_refinement_map(lv)
    for mg in range(3):
        L, F = o.block_info(mg)
        assert np.array_equal(L, np.maximum(lv, mg))
    L, F = o.block_info(0)
    assert F[0] & msbgx.BLK_FINE_COARSE and not F[0] & msbgx.BLK_COARSE_FINE
    assert F[1] & msbgx.BLK_COARSE_FINE and F[1] & msbgx.BLK_FINE_COARSE


def test_relax_lists_are_colour_conflict_free():
    cfg = msbgx.config_c4(n=128)
    o = build_oracle(cfg)
    o.prepare_scene(cfg.scene)
    nb = cfg.n // cfg.bsx
    for mg in range(cfg.levels):
        rl = o.relax_list(mg)
        assert np.all(np.diff(rl) > 0)
        coords = {int(b): (b % nb, (b // nb) % nb, b // (nb * nb)) for b in rl}
        for c in (0, 1):
            same = {b for b, (x, y, z) in coords.items() if (x + y + z) % 2 == c}
            for b in same:
                x, y, z = coords[b]
                for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    nbid = (x + d[0]) + nb * ((y + d[1]) + nb * (z + d[2]))
                    assert nbid not in same or x + d[0] >= nb or y + d[1] >= nb or z + d[2] >= nb


def test_single_level_grid_has_inert_coarse_levels():
    """All-level-0 grids: coarse flags sampled from empty same-level grids are
    VOID, so the coarse relaxation lists are empty (src/msbg3.cpp:2888-2931)."""
    cfg = msbgx.config_c1(n=64)
    o = build_oracle(cfg)
    o.prepare_scene(cfg.scene)
    assert len(o.relax_list(0)) > 0
    for mg in (1, 2):
        assert len(o.relax_list(mg)) == 0
        _, F = o.block_info(mg)
        assert np.all(F & msbgx.BLK_NO_FLUID)


def test_level_map_policy_hand_case():
    """One small sphere inside block (1,1,1) of a 4^3 raster: distance-0 seed
    block, d <= 1 -> level 0, d = 2 -> 1, else coarsest (src/flip_sim.cpp:480-548)."""
    s = msbgx.make_scene(64, spheres=[(0.375, 0.375, 0.375, 0.05)])
    lv = level_map_from_scene(s, 16, 3, 0, 1).reshape(4, 4, 4)
    z, y, x = np.indices((4, 4, 4))
    d = np.maximum(np.maximum(abs(x - 1), abs(y - 1)), abs(z - 1))
    exp = np.where(d <= 1, 0, np.where(d == 2, 1, 2))
    assert np.array_equal(lv, exp)


# ---------------------------------------------------------------- operator

def _apply(o, n, bsx, f):
    field_to_blocks(o, f, n, bsx, CH_Z)
    o.multiply(0, False, CH_Z, -1, CH_Q)
    return blocks_to_field(o, n, bsx, CH_Q)


@pytest.mark.parametrize("bsx", [4, 8, 16])
def test_operator_closed_forms(bsx):
    n = 32
    o = dense_problem(n, bsx)
    z, y, x = np.indices((n, n, n)).astype(np.float32)
    inner = (slice(1, -1),) * 3
    lin = 0.5 * x - 0.25 * y + 0.125 * z + 3.0
    assert np.abs(_apply(o, n, bsx, lin)[inner]).max() == 0.0
    quad = x * x
    assert np.all(_apply(o, n, bsx, quad)[inner] == -2.0)
    one = np.ones((n, n, n), np.float32)
    a1 = _apply(o, n, bsx, one)
    assert np.all(a1[inner] == 0.0)


def test_operator_counts_dirichlet_neighbours():
    """x = 1 on fluid, 0 on VOID: (A x)_i = number of VOID face neighbours."""
    n, bsx = 32, 8
    rng = np.random.default_rng(3)
    void = rng.random((n, n, n)) < 0.2
    o = dense_problem(n, bsx, void=void)
    x = np.where(void, 0.0, 1.0).astype(np.float32)
    ax = _apply(o, n, bsx, x)
    vp = np.pad(void, 1, constant_values=False)
    cnt = sum(np.roll(vp, s, axis=a)[1:-1, 1:-1, 1:-1] for a in range(3) for s in (1, -1))
    inner = (slice(1, -1),) * 3
    fluid = ~void
    assert np.array_equal(ax[inner][fluid[inner]], cnt[inner][fluid[inner]].astype(np.float32))
    assert np.all(ax[void] == 0.0)


def test_operator_symmetric_single_level():
    n, bsx = 32, 8
    rng = np.random.default_rng(5)
    void = rng.random((n, n, n)) < 0.15
    w = [rng.uniform(0.5, 1.5, (n, n, n)).astype(np.float32) for _ in range(3)]
    o = dense_problem(n, bsx, void=void, w=w)
    u = np.where(void, 0, rng.standard_normal((n, n, n))).astype(np.float32)
    v = np.where(void, 0, rng.standard_normal((n, n, n))).astype(np.float32)
    au, av = _apply(o, n, bsx, u), _apply(o, n, bsx, v)
    lhs = float((au.astype(np.float64) * v).sum())
    rhs = float((u.astype(np.float64) * av).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1.0)


# ---------------------------------------------------------------- smoother

def _fma(a, b, c):
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def brute_relax_block(x, bvec, D, W, omega, reverse, halo, wplus, S, h=1.0):
    """Pure-Python restatement of one block's 5 local sweeps at grid scale h
    (values h*x inside the block, halo values already scaled by h): SIMD
    segments of S lanes along x are Jacobi inside, Gauss-Seidel across;
    sweep k reversed iff reverse == (k & 1); the last sweep reads only the
    previous sweep's values and is the one stored, divided by h."""
    b = x.shape[0]
    X = np.zeros((b + 2, b + 2, b + 2), np.float32)
    X[1:-1, 1:-1, 1:-1] = x * np.float32(h)
    X[0, 1:-1, 1:-1], X[-1, 1:-1, 1:-1] = halo["z-"], halo["z+"]
    X[1:-1, 0, 1:-1], X[1:-1, -1, 1:-1] = halo["y-"], halo["y+"]
    X[1:-1, 1:-1, 0], X[1:-1, 1:-1, -1] = halo["x-"], halo["x+"]
    Wu, Wv, Ww = W
    om, omm = np.float32(omega), np.float32(1.0) - np.float32(omega)
    out = x.copy()
    for k in range(5):
        rev = reverse == bool(k & 1)
        last = k == 4
        zs = range(b - 1, -1, -1) if rev else range(b)
        for zz in zs:
            for yy in (range(b - 1, -1, -1) if rev else range(b)):
                segs = range(b // S - 1, -1, -1) if rev else range(b // S)
                for sg in segs:
                    ys = []
                    for l in range(S):
                        xx = sg * S + l
                        i, j, kk = zz + 1, yy + 1, xx + 1
                        w2 = Wu[zz, yy, xx + 1] if xx < b - 1 else wplus[0][zz, yy]
                        w4 = Wv[zz, yy + 1, xx] if yy < b - 1 else wplus[1][zz, xx]
                        w6 = Ww[zz + 1, yy, xx] if zz < b - 1 else wplus[2][yy, xx]
                        s = np.float32(X[i, j, kk - 1] * Wu[zz, yy, xx])
                        s = _fma(X[i, j, kk + 1], w2, s)
                        s = _fma(X[i, j - 1, kk], Wv[zz, yy, xx], s)
                        s = _fma(X[i, j + 1, kk], w4, s)
                        s = _fma(X[i - 1, j, kk], Ww[zz, yy, xx], s)
                        s = _fma(X[i + 1, j, kk], w6, s)
                        d = D[zz, yy, xx]
                        f0 = X[i, j, kk]
                        ys.append(_fma(np.float32(om * np.float32(bvec[zz, yy, xx] + s)), d, np.float32(omm * f0))
                                  if d != 0 else f0)
                    for l in range(S):
                        if last:
                            out[zz, yy, sg * S + l] = np.float32(ys[l] / np.float32(h))
                        else:
                            X[zz + 1, yy + 1, sg * S + l + 1] = ys[l]
    return out


@pytest.mark.parametrize("reverse", [False, True])
def test_smoother_brute_force_one_block(reverse):
    """A 1x1x1-block domain of edge 8 (SIMD segment 8) with Neumann walls."""
    n = bsx = 8
    rng = np.random.default_rng(11 + reverse)
    w = [rng.uniform(0.5, 1.5, (n, n, n)).astype(np.float32) for _ in range(3)]
    for k, ax in enumerate((2, 1, 0)):   # W = 0 on the faces at coordinate 0 (Neumann walls)
        sl = [slice(None)] * 3
        sl[ax] = 0
        w[k][tuple(sl)] = 0.0
    rhs = rng.standard_normal((n, n, n)).astype(np.float32)
    o = dense_problem(n, bsx, levels=2, w=w, rhs=rhs)
    x0 = rng.standard_normal((n, n, n)).astype(np.float32)
    field_to_blocks(o, x0, n, bsx, CH_Z)
    o.copy_channel(CH_B, CH_R)
    o.relax(0, reverse, 1, CH_Z, CH_R)
    got = blocks_to_field(o, n, bsx, CH_Z)
    D = o.get_block(10, 0, 0, bsx).reshape(n, n, n)
    zero = np.zeros((n, n), np.float32)
    halo = {k: zero for k in ("x-", "x+", "y-", "y+", "z-", "z+")}
    exp = brute_relax_block(x0, rhs, D, w, o.omega(), reverse, halo, (zero, zero, zero), 8)
    assert np.array_equal(got, exp)


def test_smoother_fixed_point_and_inactive_cells():
    n, bsx = 32, 8
    rng = np.random.default_rng(7)
    void = rng.random((n, n, n)) < 0.1
    o = dense_problem(n, bsx, void=void)
    xs = np.where(void, 0, rng.standard_normal((n, n, n))).astype(np.float32)
    b = _apply(o, n, bsx, xs)
    field_to_blocks(o, b, n, bsx, CH_R)
    field_to_blocks(o, xs, n, bsx, CH_Z)
    o.relax(0, False, 2, CH_Z, CH_R)
    got = blocks_to_field(o, n, bsx, CH_Z)
    assert np.abs(got - xs).max() <= 1e-4 * np.abs(xs).max()
    assert np.all(got[void] == 0.0)


# ---------------------------------------------------------------- transfer

def test_restriction_and_prolongation_exact():
    cfg = msbgx.config_c4(n=128)
    o = build_oracle(cfg)
    o.prepare_scene(cfg.scene)
    rng = np.random.default_rng(2)
    L0, _ = o.block_info(0)
    L1, _ = o.block_info(1)
    fine = {}
    for bid in range(o.n_blocks):
        e = cfg.bsx >> int(L0[bid])
        f = (8.0 * rng.integers(-64, 64, e ** 3)).astype(np.float32)
        fine[bid] = f
        o.set_block(CH_Q, 0, bid, f)
    o.zero_channel(CH_R, 1)
    o.downsample(1, CH_Q, CH_R)
    for bid in range(o.n_blocks):
        e0, e1 = cfg.bsx >> int(L0[bid]), cfg.bsx >> int(L1[bid])
        got = o.get_block(CH_R, 1, bid, e1).reshape(e1, e1, e1)
        f = fine[bid].reshape(e0, e0, e0)
        if L0[bid] >= 1:
            assert np.array_equal(got.reshape(-1), fine[bid])   # same level: the block itself
        else:
            exp = f.reshape(e1, 2, e1, 2, e1, 2).mean(axis=(1, 3, 5))
            assert np.array_equal(got, exp.astype(np.float32))
    # second restriction on the same channel set: same-level blocks read 0
    o.zero_channel(CH_R, 1)
    o.downsample(1, CH_Q, CH_R)
    for bid in range(o.n_blocks):
        if L0[bid] >= 1:
            e1 = cfg.bsx >> int(L1[bid])
            assert not np.any(o.get_block(CH_R, 1, bid, e1))
    # prolongation adds the parent value on fluid cells only
    coarse = {}
    for bid in range(o.n_blocks):
        e1 = cfg.bsx >> int(L1[bid])
        c = rng.integers(-50, 50, e1 ** 3).astype(np.float32)
        coarse[bid] = c
        o.set_block(CH_Z, 1, bid, c)
        o.set_block(CH_Z, 0, bid, np.zeros((cfg.bsx >> int(L0[bid])) ** 3, np.float32))
    o.prolongate(0, CH_Z)
    for bid in range(o.n_blocks):
        e0, e1 = cfg.bsx >> int(L0[bid]), cfg.bsx >> int(L1[bid])
        got = o.get_block(CH_Z, 0, bid, e0)
        fl = o.get_block(11, 0, bid, e0)
        c = coarse[bid].reshape(e1, e1, e1)
        exp = c if e0 == e1 else c.repeat(2, 0).repeat(2, 1).repeat(2, 2)
        exp = np.where((fl.reshape(e0, e0, e0) & (msbgx.CELL_SOLID | msbgx.CELL_VOID)) != 0, 0, exp)
        assert np.array_equal(got.reshape(e0, e0, e0), exp)


# ---------------------------------------------------------------- PCG

def test_pcg_solution_matches_dense_solve():
    """Assemble A on a tiny single-level grid from the operator and compare
    the converged PCG pressure with numpy.linalg.solve."""
    n, bsx = 8, 4
    rng = np.random.default_rng(9)
    void = np.zeros((n, n, n), bool)
    void[:, -2:, :] = True                       # air layer on top: Dirichlet
    w = [rng.uniform(0.5, 1.5, (n, n, n)).astype(np.float32) for _ in range(3)]
    for k, ax in enumerate((2, 1, 0)):
        sl = [slice(None)] * 3
        sl[ax] = 0
        w[k][tuple(sl)] = 0.0
    rhs = np.where(void, 0, rng.standard_normal((n, n, n))).astype(np.float32)
    o = dense_problem(n, bsx, levels=2, void=void, w=w, rhs=rhs)
    it, rr, _ = o.solve(max_iter=200, tol=1e-7)
    assert rr <= 1e-7 and it < 200
    x = blocks_to_field(o, n, bsx, CH_X)
    idx = np.flatnonzero(~void.reshape(-1))
    A = np.zeros((idx.size, idx.size))
    o2 = dense_problem(n, bsx, levels=2, void=void, w=w, rhs=rhs)
    for j, c in enumerate(idx):
        e = np.zeros(n ** 3, np.float32)
        e[c] = 1.0
        A[:, j] = _apply(o2, n, bsx, e.reshape(n, n, n)).reshape(-1)[idx]
    xs = np.linalg.solve(A, rhs.reshape(-1)[idx].astype(np.float64))
    assert np.linalg.norm(x.reshape(-1)[idx] - xs) <= 1e-5 * np.linalg.norm(xs)


def test_zero_rhs_returns_immediately():
    s = msbgx.make_scene(64, spheres=[])
    cfg = msbgx.Config("empty", 64, 16, 3, s, level_mode=-1)
    o = build_oracle(cfg)
    o.prepare_scene(s)
    it, rr, _ = o.solve(max_iter=10, tol=1e-6)
    assert it == 0 and rr == 0.0


def test_fp64_reference_solver_matches_dense_solve():
    """The oracle's FP64 solve (precision-form reference) on a tiny grid equals
    numpy.linalg.solve of the operator assembled from the FP32 oracle's apply
    (exact on this grid: the apply of a unit vector has no rounding)."""
    n, bsx = 8, 4
    rng = np.random.default_rng(13)
    void = np.zeros((n, n, n), bool)
    void[:, -2:, :] = True
    w = [np.ones((n, n, n), np.float32) for _ in range(3)]
    for k, ax in enumerate((2, 1, 0)):
        sl = [slice(None)] * 3
        sl[ax] = 0
        w[k][tuple(sl)] = 0.0
    rhs = np.where(void, 0, rng.standard_normal((n, n, n))).astype(np.float32)
    o = dense_problem(n, bsx, levels=2, void=void, w=w, rhs=rhs)
    it, rel, xb = o.solve_fp64(tol=1e-14, max_iter=1000)
    assert rel <= 1e-13
    nb = n // bsx
    x = np.zeros((n, n, n))
    for bid, blk in xb.items():
        bx, by, bz = bid % nb, (bid // nb) % nb, bid // (nb * nb)
        x[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = blk.reshape(bsx, bsx, bsx)
    # independent FP64 assembly: A_ii = 1/D_i, A_ij = -W(face ij) between fluid cells
    D = blocks_to_field(o, n, bsx, 10).astype(np.float64)
    fluid = ~void
    idx = np.flatnonzero(fluid.reshape(-1))
    pos = -np.ones(n ** 3, int)
    pos[idx] = np.arange(idx.size)
    A = np.zeros((idx.size, idx.size))
    strides = (1, n, n * n)
    for k, c in enumerate(idx):
        z, y, xx = np.unravel_index(c, (n, n, n))
        A[k, k] = 1.0 / D.reshape(-1)[c]
        for d, (coord, wk) in enumerate(((xx, w[0]), (y, w[1]), (z, w[2]))):
            if coord > 0 and pos[c - strides[d]] >= 0:
                A[k, pos[c - strides[d]]] -= float(wk.reshape(-1)[c])
            if coord < n - 1 and pos[c + strides[d]] >= 0:
                A[k, pos[c + strides[d]]] -= float(wk.reshape(-1)[c + strides[d]])
    xs = np.linalg.solve(A, rhs.reshape(-1)[idx].astype(np.float64))
    assert np.linalg.norm(x.reshape(-1)[idx] - xs) <= 1e-10 * np.linalg.norm(xs)
