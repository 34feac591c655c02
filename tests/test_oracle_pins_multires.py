"""Pins of the oracle's mixed-resolution and boundary semantics (CPU only).

Each expected value below is worked out in the test from the reference's
rules, not from the oracle:

  halo      a fine block next to a coarse block (and the other way round)
            along each axis and on each side: the staged neighbour value is
            h_own * (2/3) * [W_own] * x_coarse (fine side; W_own only when the
            coarse block is on the -side) and h_own * sum_4 (1/3) * [W_fine]
            * x_fine (coarse side; W_fine only when the fine block is on the
            +side), staged with face weight 1 (src/msbg3.cpp:46-161,
            690-742, 897-1044; DPMR_GRAD_COEFF_CF = 1/3, src/msbg.h:85-87)
  boundary  A*1 at every boundary cell: a -face at the domain boundary keeps
            the cell's own W against a zero ghost value, a +face counts 0 in
            the diagonal (src/msbg3.cpp:728-742, 2706-2722); the +face weight
            the reference stages there (0 non-OPEN, 1 OPEN) multiplies a zero
            ghost value and so changes no output
  coarse    coarse cell flags, face weights and diagonal of a hand-worked
            4-block raster, sampled from the same-level grid at doubled
            coordinates (src/msbg3.cpp:2882-3007, 2731-2777)
  omega     (6/7) / sqrt(max D / min D) on a grid whose D takes the values
            1/6 ... 1/3 (src/msbg3.cpp:3063-3094, src/msbg.cpp:819)
  smoother  brute force of the red-black block schedule on a 2x2x2 raster of
            4^3 / 2^3 / 8^3 blocks (SIMD segment widths 4, 2, 8) with halos
            from neighbour blocks and OPEN walls (src/msbg3.cpp:1109-1506,
            2171-2369, 1700-1743)
"""
import numpy as np
import pytest

import msbgx
from oracle_lib import Oracle
from test_oracle_pins import blocks_to_field, brute_relax_block, field_to_blocks

CH_X, CH_B, CH_R, CH_Z, CH_D, CH_Q = 0, 1, 2, 3, 4, 5
CH_W = (7, 8, 9)
CH_DIAG, CH_FLAGS = 10, 11
F32 = np.float32


def dyadic(rng, shape, lo=4, hi=24):
    """Face weights k/16 (exact float sums of a few terms)."""
    return (rng.integers(lo, hi, shape) / 16.0).astype(F32)


def perm(axis, a, t1, t2):
    """(x, y, z) of a point with coordinate a along `axis` and t1, t2 along
    the other two axes in increasing order."""
    c = [t1, t2]
    c.insert(axis, a)
    return tuple(c)


def block_view(blk, e):
    """Block data (x fastest) as [z][y][x]."""
    return blk.reshape(e, e, e)


def at(v, xyz):
    return v[xyz[2], xyz[1], xyz[0]]


# ------------------------------------------------------- mixed-resolution halo

@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("fine_first", [True, False])
def test_mixed_resolution_halo_by_hand(axis, fine_first):
    bsx, levels = 8, 2
    e_c = bsx // 2
    shape_b = [1, 1, 1]
    shape_b[axis] = 2
    sx, sy, sz = (s * bsx for s in shape_b)
    rng = np.random.default_rng(31 + 2 * axis + fine_first)
    flags = np.zeros((sz, sy, sx), np.uint16)
    w = [dyadic(rng, (sz, sy, sx)) for _ in range(3)]
    o = Oracle(sx, sy, sz, bsx, levels)
    o.set_domain_boundary_spec(*([msbgx.DBC_SOLID] * 6))
    bid_f, bid_c = (0, 1) if fine_first else (1, 0)
    lv = np.zeros(2, np.int32)
    lv[bid_c] = 1
    o.set_refinement_map(lv)
    o.prepare_dense((sx, sy, sz), flags, *w, np.zeros_like(w[0]))
    L, _ = o.block_info(0)
    assert L[bid_f] == 0 and L[bid_c] == 1
    wf = block_view(o.get_block(CH_W[axis], 0, bid_f, bsx), bsx)
    # the fine block samples the dense face array at its own cells
    f0 = [0, 0, 0]
    f0[axis] = bsx * bid_f
    assert np.array_equal(wf, w[axis][f0[2]:f0[2] + bsx, f0[1]:f0[1] + bsx, f0[0]:f0[0] + bsx])
    third, two_thirds = F32(1.0 / 3.0), F32(2.0 / 3.0)

    # 1. coarse values, fine block zero: the fine block's face layer sees the coarse neighbour
    xc = block_view(rng.integers(-40, 41, e_c ** 3).astype(F32), e_c)
    o.set_block(CH_Z, 0, bid_c, xc.reshape(-1))
    o.set_block(CH_Z, 0, bid_f, np.zeros(bsx ** 3, F32))
    o.multiply(0, False, CH_Z, -1, CH_Q)
    yf = block_view(o.get_block(CH_Q, 0, bid_f, bsx), bsx)
    exp = np.zeros((bsx, bsx, bsx), F32)
    layer = bsx - 1 if fine_first else 0            # the fine layer touching the coarse block
    clayer = 0 if fine_first else e_c - 1           # the coarse layer touching the fine block
    for t1 in range(bsx):
        for t2 in range(bsx):
            p = perm(axis, layer, t1, t2)
            a = two_thirds
            if not fine_first:                      # coarse block on the -side: own W weights it
                a = F32(a * at(wf, p))
            val = F32(a * at(xc, perm(axis, clayer, t1 // 2, t2 // 2)))   # h = 1 on level 0
            exp[p[2], p[1], p[0]] = -val
    assert np.array_equal(yf, exp), f"fine side: {np.count_nonzero(yf != exp)} cells differ"

    # 2. fine values, coarse block zero: the coarse face layer sees four fine cells each
    xf = block_view(rng.integers(-40, 41, bsx ** 3).astype(F32), bsx)
    o.set_block(CH_Z, 0, bid_f, xf.reshape(-1))
    o.set_block(CH_Z, 0, bid_c, np.zeros(e_c ** 3, F32))
    o.multiply(0, False, CH_Z, -1, CH_Q)
    yc = block_view(o.get_block(CH_Q, 0, bid_c, e_c), e_c)
    exp = np.zeros((e_c, e_c, e_c), F32)
    flayer = bsx - 1 if fine_first else 0           # the fine layer touching the coarse block
    for t1 in range(e_c):
        for t2 in range(e_c):
            s = 0.0
            for d1 in (0, 1):
                for d2 in (0, 1):
                    q = perm(axis, flayer, 2 * t1 + d1, 2 * t2 + d2)
                    a = third
                    if not fine_first:              # fine block on the +side: its -face W weights it
                        a = F32(a * at(wf, perm(axis, 0, 2 * t1 + d1, 2 * t2 + d2)))
                    s += float(F32(a * at(xf, q)))
            p = perm(axis, clayer, t1, t2)
            exp[p[2], p[1], p[0]] = -F32(s * 2.0)   # h = 2 on level 1
    assert np.array_equal(yc, exp), f"coarse side: {np.count_nonzero(yc != exp)} cells differ"


# ------------------------------------------------------------- domain boundary

@pytest.mark.parametrize("bc", [msbgx.DBC_OPEN, msbgx.DBC_SOLID])
def test_domain_boundary_faces_at_boundary_cells(bc):
    n, bsx = 16, 8
    rng = np.random.default_rng(40 + bc)
    w = [dyadic(rng, (n, n, n)) for _ in range(3)]
    o = Oracle(n, n, n, bsx, 2)
    o.set_domain_boundary_spec(*([bc] * 6))
    o.set_refinement_map(np.zeros(o.n_blocks, np.int32))
    o.prepare_dense((n, n, n), np.zeros((n, n, n), np.uint16), *w, np.zeros((n, n, n), F32))
    # D = 1 / (sum of the six face weights), a face outside the domain counting 0
    pad = [np.pad(wk, [(0, 1) if ax == 2 - k else (0, 0) for ax in range(3)]) for k, wk in enumerate(w)]
    tot = (w[0] + pad[0][:, :, 1:]) + (w[1] + pad[1][:, 1:, :]) + (w[2] + pad[2][1:, :, :])
    D = blocks_to_field(o, n, bsx, CH_DIAG)
    assert np.array_equal(D, (F32(1.0) / tot).astype(F32))
    # A * 1: interior faces cancel; a -face at the boundary leaves its own W
    # (Dirichlet ghost 0), a +face at the boundary leaves nothing
    field_to_blocks(o, np.ones((n, n, n), F32), n, bsx, CH_Z)
    o.multiply(0, False, CH_Z, -1, CH_Q)
    y = blocks_to_field(o, n, bsx, CH_Q).astype(np.float64)
    exp = np.zeros((n, n, n))
    exp[:, :, 0] += w[0][:, :, 0]
    exp[:, 0, :] += w[1][:, 0, :]
    exp[0, :, :] += w[2][0, :, :]
    assert np.abs(y - exp).max() <= 4e-7 * tot.max()
    boundary = np.zeros((n, n, n), bool)
    for ax in range(3):
        for s in (0, -1):
            sl = [slice(None)] * 3
            sl[ax] = s
            boundary[tuple(sl)] = True
    assert np.count_nonzero(np.abs(exp[boundary]) > 0) > 0


# ------------------------------------------------------ coarse MG metadata

@pytest.mark.parametrize("axis", [0, 1, 2])
def test_coarse_metadata_hand_raster(axis):
    """Four blocks of edge 4 in a row along `axis`, level map [0, 0, 1, 1],
    two MG levels.  At MG 1 every block is level 1 (edge 2); its cells sample
    the level-1 grid of MG 0 at doubled level-1 coordinates.  That grid holds
    blocks 2 and 3, which sampled the dense inputs at level-1 coordinates
    4..7 along the axis and 0..1 across it.  So coarse block 1 cell (c, 0, 0)
    reads dense cells 4 + 2c + {0, 1} along the axis and {0, 1} across; every
    other coarse cell reads outside that grid (VOID, weight 0)."""
    bsx, levels = 4, 2
    nb = [1, 1, 1]
    nb[axis] = 4
    sx, sy, sz = (k * bsx for k in nb)
    rng = np.random.default_rng(55 + axis)
    flags = np.full((sz, sy, sx), msbgx.CELL_VOID, np.uint16)
    w = [dyadic(rng, (sz, sy, sx)) for _ in range(3)]

    def put(arr, a, t1, t2, v):
        x, y, z = perm(axis, a, t1, t2)
        arr[z, y, x] = v

    for a in range(4, 6):        # children of coarse cell (0,0,0): solid and void, no fluid
        for t1 in (0, 1):
            for t2 in (0, 1):
                put(flags, a, t1, t2, msbgx.CELL_SOLID if (a + t1 + t2) % 2 else msbgx.CELL_VOID)
    put(flags, 7, 1, 0, 0)       # one fluid child of coarse cell (1,0,0)
    o = Oracle(sx, sy, sz, bsx, levels)
    o.set_domain_boundary_spec(*([msbgx.DBC_SOLID] * 6))
    o.set_refinement_map(np.array([0, 0, 1, 1], np.int32))
    o.prepare_dense((sx, sy, sz), flags, *w, np.zeros_like(w[0]))

    L1, F1 = o.block_info(1)
    assert np.array_equal(L1, [1, 1, 1, 1])
    no_fluid = (F1 & msbgx.BLK_NO_FLUID) != 0
    assert list(no_fluid) == [True, False, True, True]
    assert list(o.relax_list(1)) == [1]
    fl = block_view(o.get_block(CH_FLAGS, 1, 1, 2), 2)
    exp_fl = np.full((2, 2, 2), msbgx.CELL_VOID, np.uint16)
    p0, p1 = perm(axis, 0, 0, 0), perm(axis, 1, 0, 0)
    exp_fl[p0[2], p0[1], p0[0]] = msbgx.CELL_SOLID
    exp_fl[p1[2], p1[1], p1[0]] = 0
    got_fl = fl & (msbgx.CELL_SOLID | msbgx.CELL_VOID | 0x2)
    assert np.array_equal(got_fl, exp_fl)
    for b in (0, 2, 3):
        assert np.all(o.get_block(CH_FLAGS, 1, b, 2) & msbgx.CELL_VOID)

    # face weights: 1/4 of the four tangential children of the same-level grid
    def dense_w(k, a, t1, t2):
        x, y, z = perm(axis, a, t1, t2)
        return w[k][z, y, x]

    wc = [block_view(o.get_block(CH_W[k], 1, 1, 2), 2) for k in range(3)]
    for c in (0, 1):
        for k in range(3):
            s = 0.0
            for j in (0, 1):
                for m in (0, 1):
                    off = [0, 0, 0]
                    tang = [d for d in range(3) if d != k]
                    off[tang[0]], off[tang[1]] = j, m
                    along = 4 + 2 * c + off[axis]
                    across = [off[d] for d in range(3) if d != axis]
                    s += float(dense_w(k, along, across[0], across[1]))
            p = perm(axis, c, 0, 0)
            assert at(wc[k], p) == F32(0.25 * s), f"W{k} of coarse cell {c}"
        for t in ((0, 1, 0), (0, 0, 1), (1, 1, 1)):   # cells whose children are outside the grid
            q = perm(axis, c, t[1], t[2])
            for k in range(3):
                assert at(wc[k], q) == 0
    # diagonal: 1 / (2 (Wx + Wy + Wz)) on the fluid cell, 0 elsewhere
    d = block_view(o.get_block(CH_DIAG, 1, 1, 2), 2)
    ssum = F32(F32(at(wc[0], p1) + at(wc[1], p1)) + at(wc[2], p1))
    exp_d = np.zeros((2, 2, 2), F32)
    exp_d[p1[2], p1[1], p1[0]] = F32(1.0) / F32(ssum * F32(2.0))
    assert np.array_equal(d, exp_d)


# ------------------------------------------------------------------- omega

def test_omega_from_two_diagonal_extremes():
    """W = 1 everywhere with OPEN walls: interior cells sum 6 faces, cells on
    +faces of the domain lose one face each, the (+,+,+) corner keeps 3, so
    D ranges over 1/6 .. 1/3 and omega = (6/7) / sqrt(2)."""
    n, bsx = 16, 8
    o = Oracle(n, n, n, bsx, 2)
    o.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    o.set_refinement_map(np.zeros(o.n_blocks, np.int32))
    ones = np.ones((n, n, n), F32)
    o.prepare_dense((n, n, n), np.zeros((n, n, n), np.uint16), ones, ones, ones, np.zeros((n, n, n), F32))
    D = blocks_to_field(o, n, bsx, CH_DIAG)
    assert D.max() == F32(1.0 / 3.0) and D.min() == F32(1.0 / 6.0) and D[-1, -1, -1] == F32(1.0 / 3.0)
    assert o.omega() == F32(F32(6.0 / 7.0) / np.sqrt(F32(2.0)))
    # a uniform diagonal keeps the default weight
    o2 = Oracle(n, n, n, bsx, 2)
    o2.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    o2.set_refinement_map(np.zeros(o2.n_blocks, np.int32))
    flags = np.full((n, n, n), msbgx.CELL_VOID, np.uint16)
    flags[4:8, 4:8, 4:8] = 0
    o2.prepare_dense((n, n, n), flags, ones, ones, ones, np.zeros((n, n, n), F32))
    assert o2.omega() == F32(6.0 / 7.0)


# ---------------------------------------------------------------- smoother

def brute_relax_grid(x, b, w, D, omega, nb, e, h, relaxed, reverse, seg, bc_plus_w):
    """One relax iteration (both colours) over the blocks `relaxed` of a
    raster of nb = (nbx, nby, nbz) same-level blocks of edge e at grid scale
    h (fields [z][y][x] over the raster): colour 1 first (forward) or colour
    0 first (reverse); each block of a colour takes its halo, h * x, from the
    field as the previous colour left it (0 outside the domain) and the
    +face weights from the neighbouring cells (bc_plus_w outside), and runs
    the 5 local sweeps of brute_relax_block."""
    nbx, nby, nbz = nb
    x = x.copy()
    hh = F32(h)
    for colour in ((0, 1) if reverse else (1, 0)):
        snap = x.copy()
        pad = np.pad(snap, 1)
        for bid in relaxed:
            bx, by, bz = bid % nbx, (bid // nbx) % nby, bid // (nbx * nby)
            if (bx + by + bz) % 2 != colour:
                continue
            z0, y0, x0 = bz * e, by * e, bx * e
            sl = (slice(z0, z0 + e), slice(y0, y0 + e), slice(x0, x0 + e))
            P = (pad[z0:z0 + e + 2, y0:y0 + e + 2, x0:x0 + e + 2] * hh).astype(F32)
            halo = {"x-": P[1:-1, 1:-1, 0], "x+": P[1:-1, 1:-1, -1], "y-": P[1:-1, 0, 1:-1],
                    "y+": P[1:-1, -1, 1:-1], "z-": P[0, 1:-1, 1:-1], "z+": P[-1, 1:-1, 1:-1]}
            wp = []
            for k in range(3):
                hi = (x0, y0, z0)[k] + e
                if hi < (nbx, nby, nbz)[k] * e:
                    idx = list(sl)
                    idx[2 - k] = hi
                    wp.append(w[k][tuple(idx)])
                else:
                    wp.append(np.full((e, e), bc_plus_w, F32))
            x[sl] = brute_relax_block(snap[sl], b[sl], D[sl], [wk[sl] for wk in w], omega, reverse, halo, wp,
                                      seg, h)
    return x


def level_field(o, nb, e, ch, mg, dtype=F32):
    nbx, nby, nbz = nb
    f = np.zeros((nbz * e, nby * e, nbx * e), dtype)
    for bid in range(nbx * nby * nbz):
        bx, by, bz = bid % nbx, (bid // nbx) % nby, bid // (nbx * nby)
        blk = o.get_block(ch, mg, bid, e)
        if blk is not None:
            f[bz * e:(bz + 1) * e, by * e:(by + 1) * e, bx * e:(bx + 1) * e] = blk.reshape(e, e, e)
    return f


def set_level_field(o, f, nb, e, ch, mg):
    nbx, nby, nbz = nb
    for bid in range(nbx * nby * nbz):
        bx, by, bz = bid % nbx, (bid // nbx) % nby, bid // (nbx * nby)
        o.set_block(ch, mg, bid, f[bz * e:(bz + 1) * e, by * e:(by + 1) * e, bx * e:(bx + 1) * e].reshape(-1))


# (bsx0, raster edge in blocks, map level, MG level): the relaxed blocks have
# edge bsx0 >> level, hence SIMD segments of 8, 4 or 2 lanes, at h = 2^level
SMOOTHER_CASES = [(8, 2, 0, 0), (4, 2, 0, 0), (8, 4, 1, 1), (4, 4, 1, 1)]


@pytest.mark.parametrize("bsx0,nbe,lev,mg", SMOOTHER_CASES)
@pytest.mark.parametrize("reverse", [False, True])
def test_smoother_brute_force_raster_with_halos(bsx0, nbe, lev, mg, reverse):
    """Level-0 cases: every block is relaxed.  Level-1 cases (all blocks level
    1, relaxed at MG 1): the MG 1 metadata samples the MG 0 level-1 grid at
    doubled coordinates, so only the blocks with coordinates 0..1 hold fluid
    and their +face neighbours carry zero weights."""
    n = nbe * bsx0
    e = bsx0 >> lev
    seg = {8: 8, 4: 4, 2: 2}[e]
    rng = np.random.default_rng(70 + bsx0 + 7 * lev + 3 * reverse)
    w = [dyadic(rng, (n, n, n)) for _ in range(3)]
    rhs = rng.integers(-8, 9, (n, n, n)).astype(F32)
    o = Oracle(n, n, n, bsx0, 2)
    o.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    o.set_refinement_map(np.full(o.n_blocks, lev, np.int32))
    o.prepare_dense((n, n, n), np.zeros((n, n, n), np.uint16), *w, rhs)
    nb = (nbe, nbe, nbe)
    relaxed = list(o.relax_list(mg))
    assert len(relaxed) == (nbe ** 3 if lev == 0 else 8)
    D = level_field(o, nb, e, CH_DIAG, mg)
    W = [level_field(o, nb, e, CH_W[k], mg) for k in range(3)]
    m = nbe * e
    x0 = rng.integers(-16, 17, (m, m, m)).astype(F32)
    b0 = rng.integers(-8, 9, (m, m, m)).astype(F32)
    set_level_field(o, x0, nb, e, CH_Z, mg)
    set_level_field(o, b0, nb, e, CH_R, mg)
    o.relax(mg, reverse, 1, CH_Z, CH_R)
    got = level_field(o, nb, e, CH_Z, mg)
    exp = brute_relax_grid(x0, b0, W, D, o.omega(), nb, e, 2.0 ** lev, relaxed, reverse, seg, F32(1.0))
    assert np.array_equal(got, exp), f"{np.count_nonzero(got != exp)} cells differ"
    assert np.count_nonzero(got != x0) > 0
