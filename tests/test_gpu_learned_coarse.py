"""Learned preconditioner, coarse form, on the GPU (build addition, no
reference counterpart): the per-kernel apply ch_out += sum_f P_f N(R_f ch_in)
over the net's scales f = 1, 2, 4 equals a numpy restatement built on the
FP64 network oracle --

* R_1: mean over all cells of each PCG block (level-0 block with unknowns)
  on the dense block grid, 0 for every other block; R_2f: mean of R_f over
  the PCG blocks of each 2^3 cells;
* N: the network over 16^3 tiles of each grid, each read with a one-cell
  halo (0 outside the raster), all outputs kept;
* P_f: constant (the value of the cell holding the block) or trilinear
  between cell-centred centres of the scale-f cells, weights renormalised
  over the cells with PCG blocks among the eight nearest centres.

The trilinear restatement is itself pinned first (CPU test): it reproduces a
linear field of cell centres exactly where every neighbour takes part, and a
constant field everywhere, at scales 1 and 2.
"""
import numpy as np
import pytest

import msbgx
import nn_oracle_lib as nno


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(b), 1e-300))


def prolong_linear(coarse, pcg, b, bsx, bdims, f=1):
    """Trilinear, cell-centred, mask-renormalised value of `coarse` (dense
    grid of f^3-block cells [z][y][x]) on the cells of block index b of the
    block grid bdims ([z][y][x])."""
    nbx, nby, nbz = coarse.shape[2], coarse.shape[1], coarse.shape[0]
    b3 = (b % bdims[0], (b // bdims[0]) % bdims[1], b // (bdims[0] * bdims[1]))
    bx, by, bz = (v // f for v in b3)
    off = [(v - f * (v // f)) * bsx for v in b3]
    uu = [(o + np.arange(bsx) + 0.5) / (f * bsx) - 0.5 for o in off]
    lo = [np.where(u < 0, -1, 0) for u in uu]
    t = [u - l for u, l in zip(uu, lo)]
    num = np.zeros((bsx, bsx, bsx))
    den = np.zeros((bsx, bsx, bsx))
    for k in range(8):
        sx, sy, sz = k & 1, (k >> 1) & 1, k >> 2
        ox, oy, oz = lo[0] + sx, lo[1] + sy, lo[2] + sz             # per-axis neighbour offsets, shape (bsx,)
        wx = np.where(sx, t[0], 1 - t[0])
        wy = np.where(sy, t[1], 1 - t[1])
        wz = np.where(sz, t[2], 1 - t[2])
        X, Y, Z = bx + ox, by + oy, bz + oz
        inx, iny, inz = (X >= 0) & (X < nbx), (Y >= 0) & (Y < nby), (Z >= 0) & (Z < nbz)
        Xc, Yc, Zc = np.clip(X, 0, nbx - 1), np.clip(Y, 0, nby - 1), np.clip(Z, 0, nbz - 1)
        m = pcg[Zc[:, None, None], Yc[None, :, None], Xc[None, None, :]] & \
            inz[:, None, None] & iny[None, :, None] & inx[None, None, :]
        c = coarse[Zc[:, None, None], Yc[None, :, None], Xc[None, None, :]]
        w = wz[:, None, None] * wy[None, :, None] * wx[None, None, :] * m
        num += w * np.where(m, c, 0.0)
        den += w
    return num / den


@pytest.mark.parametrize("f", [1, 2])
def test_linear_prolongation_restatement_pins(f):
    """Closed forms: a linear function of the scale-f cell centres is
    reproduced at the fine cell centres of blocks whose cell's 26 neighbours
    all take part; a constant is reproduced everywhere, whatever the mask."""
    bdims = (10, 8, 12)
    cdims = tuple((v + f - 1) // f for v in bdims)
    nbx, nby, nbz = cdims
    bsx = 8
    z, y, x = np.meshgrid(np.arange(nbz), np.arange(nby), np.arange(nbx), indexing="ij")
    lin = 0.3 * x - 1.7 * y + 0.9 * z + 2.0
    rng = np.random.default_rng(3)
    pcg = rng.random((nbz, nby, nbx)) < 0.7
    full = np.ones_like(pcg)
    for b in range(bdims[0] * bdims[1] * bdims[2]):
        b3 = (b % bdims[0], (b // bdims[0]) % bdims[1], b // (bdims[0] * bdims[1]))
        bx, by, bz = (v // f for v in b3)
        if 1 <= bx < nbx - 1 and 1 <= by < nby - 1 and 1 <= bz < nbz - 1:
            got = prolong_linear(lin, full, b, bsx, bdims, f)
            # fine cell centres in units of scale-f cells, relative to cell index 0's centre
            c = [((v - f * (v // f)) * bsx + np.arange(bsx) + 0.5) / (f * bsx) - 0.5 for v in b3]
            exp = 0.3 * (bx + c[0])[None, None, :] - 1.7 * (by + c[1])[None, :, None] + \
                0.9 * (bz + c[2])[:, None, None] + 2.0
            np.testing.assert_allclose(got, exp, rtol=0, atol=1e-12)
        if pcg[bz, by, bx]:
            np.testing.assert_allclose(prolong_linear(np.full(pcg.shape, 4.25), pcg, b, bsx, bdims, f), 4.25,
                                       rtol=0, atol=1e-12)


def coarsen(cin, w):
    """Mean of cin over the PCG blocks of each 2^3 cells (weights w)."""
    nz, ny, nx = cin.shape
    cz, cy, cx = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    sv = np.zeros((cz, cy, cx))
    sw = np.zeros((cz, cy, cx))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                sv[z // 2, y // 2, x // 2] += float(w[z, y, x]) * float(cin[z, y, x])
                sw[z // 2, y // 2, x // 2] += w[z, y, x]
    out = np.where(sw > 0, sv / np.where(sw > 0, sw, 1), 0.0).astype(np.float32)
    return out, sw


def network_on_grid(w, cin):
    """N over 16^3 tiles of the dense grid cin ([z][y][x]), one-cell halo, 0 outside."""
    nz, ny, nx = cin.shape
    tz_n, ty_n, tx_n = (nz + 15) // 16, (ny + 15) // 16, (nx + 15) // 16
    patches = []
    pad = np.zeros((16 * tz_n + 2, 16 * ty_n + 2, 16 * tx_n + 2), np.float32)
    pad[1:nz + 1, 1:ny + 1, 1:nx + 1] = cin
    for tz in range(tz_n):
        for ty in range(ty_n):
            for tx in range(tx_n):
                patches.append(pad[16 * tz:16 * tz + 18, 16 * ty:16 * ty + 18, 16 * tx:16 * tx + 18].reshape(-1))
    out = nno.forward(w, np.array(patches), np.ones((len(patches), 4096), np.uint8), round_bf16=True)
    full = np.zeros((16 * tz_n, 16 * ty_n, 16 * tx_n))
    i = 0
    for tz in range(tz_n):
        for ty in range(ty_n):
            for tx in range(tx_n):
                full[16 * tz:16 * tz + 16, 16 * ty:16 * ty + 16, 16 * tx:16 * tx + 16] = out[i].reshape(16, 16, 16)
                i += 1
    return full[:nz, :ny, :nx]


def coarse_reference(g, cfg, w, R, blocks, prolong, scales):
    bsx = cfg.bsx
    nb = cfg.n // bsx
    bdims = (nb, nb, nb)
    cin = np.zeros((nb, nb, nb), np.float32)
    wgt = np.zeros((nb, nb, nb))
    for b, v in R.items():
        bx, by, bz = b % nb, (b // nb) % nb, b // (nb * nb)
        cin[bz, by, bx] = np.float32(v.astype(np.float64).mean())
        wgt[bz, by, bx] = 1.0
    ref = {b: np.zeros((bsx, bsx, bsx)) for b in blocks}
    unk = {}
    for b in blocks:
        fl = g.get_block(msbgx.CH_FLAGS, 0, b, bsx).astype(np.uint16).reshape(bsx, bsx, bsx)
        D = g.get_block(msbgx.CH_DIAGONAL, 0, b, bsx).reshape(bsx, bsx, bsx)
        unk[b] = ((fl & (msbgx.CELL_SOLID | msbgx.CELL_VOID)) == 0) & (D != 0)
    for k in range(scales):
        f = 1 << k
        if k > 0:
            cin, wgt = coarsen(cin, wgt)
        cout = network_on_grid(w, cin)
        for b in blocks:
            bx, by, bz = b % nb, (b // nb) % nb, b // (nb * nb)
            if prolong == msbgx.PROLONG_LINEAR:
                v = prolong_linear(cout, wgt > 0, b, bsx, bdims, f)
            else:
                v = np.full((bsx, bsx, bsx), cout[bz // f, by // f, bx // f])
            ref[b] += np.where(unk[b], v, 0.0)
    return ref


CASES = [("c4_384_b16", lambda: msbgx.config_c4(n=384)), ("c5_128_b32", lambda: msbgx.config_c5(n=128)),
         ("c5_256_b32", lambda: msbgx.config_c5(n=256))]


@pytest.mark.gpu
@pytest.mark.parametrize("prolong,scales", [(0, 1), (1, 1), (1, 2), (0, 3), (1, 3)],
                         ids=["constant", "linear", "linear_2scales", "constant_3scales", "linear_3scales"])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_coarse_apply_matches_restatement(name, make, prolong, scales):
    cfg = make()
    print(f"coarse apply {name} prolong={prolong} scales={scales}", flush=True)
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    net = msbgx.LearnedNet(seed=13, mode=msbgx.NET_COARSE)
    net.prolongation = prolong
    net.scales = scales
    assert net.prolongation == prolong and net.scales == scales
    w = net.weights.astype(np.float64)
    blocks = [int(b) for b in g.level0_blocks()]
    rng = np.random.default_rng(7)
    R = {}
    bsx3 = cfg.bsx ** 3
    # a smooth field plus noise, so that the block means vary between blocks
    nb = cfg.n // cfg.bsx
    for b in blocks:
        bx, by, bz = b % nb, (b // nb) % nb, b // (nb * nb)
        v = (np.sin(0.7 * bx) + np.cos(0.5 * by) - 0.3 * bz + rng.normal(0, 1, bsx3)).astype(np.float32)
        g.set_block(msbgx.CH_R, 0, b, v)
        R[b] = v
    g.zero_channel(msbgx.CH_Z, 0)
    net.apply(g, msbgx.CH_R, msbgx.CH_Z)
    ref = coarse_reference(g, cfg, w, R, blocks, prolong, scales)
    got = np.array([g.get_block(msbgx.CH_Z, 0, b, cfg.bsx) for b in blocks])
    exp = np.array([ref[b].reshape(-1) for b in blocks])
    r = rel(got, exp)
    print(f"  {len(blocks)} PCG blocks, rel {r:.2e}", flush=True)
    assert np.count_nonzero(exp) > 0 and r <= 2e-3
    # apply accumulates: a second call doubles the correction (up to the
    # rounding of adding the scales' directions one by one)
    net.apply(g, msbgx.CH_R, msbgx.CH_Z)
    got2 = np.array([g.get_block(msbgx.CH_Z, 0, b, cfg.bsx) for b in blocks])
    assert rel(got2, 2 * got) <= 1e-6
    g.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("scales", [1, 3])
@pytest.mark.parametrize("name,make", [CASES[0], CASES[1]], ids=["c4_128", "c5_128"])
def test_linear_prolongation_solve(name, make, scales):
    """Zero network: plain iterations (a zero direction gets no weight); NaN
    network: fallback; a random network with trilinear prolongation still
    converges -- at one scale (Galerkin step) and three (Galerkin
    combination)."""
    cfg = make() if name != "c4_128" else msbgx.config_c4(n=128)
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    it0, _ = g.solve_pressure_flip(max_iter=1000, tol=1e-6)
    for fill, expect_fb in ((0.0, False), (np.nan, True), (None, None)):
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        net = msbgx.LearnedNet(seed=2, mode=msbgx.NET_COARSE)
        net.prolongation = msbgx.PROLONG_LINEAR
        net.scales = scales
        if fill is not None:
            net.weights = np.full(msbgx.NET_PARAMS, fill, np.float32)
        it, rr, fb = net.solve(g, max_iter=1000, tol=1e-6)
        print(f"  {name} linear prolongation, {scales} scales, weights {fill}: {it} iterations (plain {it0}) "
              f"fallback {fb}",
              flush=True)
        assert rr <= 1e-6
        if fill is not None:
            assert (fb >= 0) == expect_fb and it <= it0 + 3
    g.destroy()


@pytest.mark.gpu
def test_two_direction_pass_matches_separate_passes():
    """Two scales on 32^3 blocks: the residual pass that serves both
    directions and accumulates the Galerkin sums gives the iterates of the
    separate residual passes and Galerkin pass (MSBGX_OPT_UNFUSED_COARSE) --
    the residuals are the same cell by cell and the FP64 sums differ only in
    their order, so a wrong or missing Galerkin term would show at once.  The
    fused pass keeps FP32 directions here (MSBGX_OPT_FP32_DIRECTIONS)."""
    cfg = msbgx.config_c5(n=256)
    print(f"two-direction pass {cfg.name} n={cfg.n}", flush=True)
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    g.set_precision(msbgx.PREC_FP16)
    net = msbgx.LearnedNet(seed=3, mode=msbgx.NET_COARSE)
    net.prolongation = msbgx.PROLONG_LINEAR
    net.scales = 2
    net.set_stall(0, 0.0)
    runs = []
    for opts in (msbgx.OPT_FP32_DIRECTIONS, msbgx.OPT_UNFUSED_COARSE):
        g.set_options(opts)
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr, fb = net.solve(g, max_iter=40, tol=0.0)
        x = np.concatenate([g.get_block(msbgx.CH_PRESSURE, 0, b, cfg.bsx).ravel() for b in g.level0_blocks()])
        runs.append((it, fb, np.array(g.residual_history()), x.astype(np.float64)))
        print(f"  options {opts}: {it} iterations rel {rr:.3e} fallback {fb}", flush=True)
    (it_a, fb_a, h_a, x_a), (it_b, fb_b, h_b, x_b) = runs
    assert it_a == it_b == 40 and fb_a == fb_b == -1
    np.testing.assert_allclose(h_a, h_b, rtol=1e-4)
    assert rel(x_a, x_b) <= 1e-4
    g.destroy()


@pytest.mark.gpu
def test_bf16_directions_follow_fp32_directions():
    """Two scales on 32^3 blocks with a briefly trained network: the default
    BF16 storage of the directions and their residuals (FP32 arithmetic, A zc
    on the PCG blocks) follows the FP32-direction solve -- the same early
    residuals to rounding, iteration counts within 10 % -- and both keep the
    network's gain over plain MG-PCG, which a misplaced or mis-rounded
    direction would lose."""
    cfg = msbgx.config_c5(n=256)
    print(f"BF16 directions {cfg.name} n={cfg.n}", flush=True)
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    g.set_precision(msbgx.PREC_FP16)
    vc = {"n_pre": 1, "n_post": 1}
    it_p, _ = g.solve_pressure_flip(max_iter=1000, tol=1e-6, **vc)
    net = msbgx.LearnedNet(seed=3, mode=msbgx.NET_COARSE)
    net.prolongation = msbgx.PROLONG_LINEAR
    net.scales = 2
    net.collect(g, max_pairs=200, seed=5, tol=1e-7, max_iter=1000)
    net.train(600, batch=32, lr=1e-3, seed=1)
    net.set_stall(0, 0.0)
    runs = {}
    for name, opts in (("fp32", msbgx.OPT_FP32_DIRECTIONS), ("bf16", 0)):
        g.set_options(opts)
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr, fb = net.solve(g, max_iter=1000, tol=1e-6, **vc)
        runs[name] = (it, rr, fb, np.array(g.residual_history()))
        print(f"  {name} directions: {it} iterations rel {rr:.2e} fallback {fb} (plain {it_p})", flush=True)
    g.set_options(0)
    (it32, rr32, fb32, h32), (it16, rr16, fb16, h16) = runs["fp32"], runs["bf16"]
    assert rr32 <= 1e-6 and rr16 <= 1e-6 and fb32 == fb16 == -1
    assert it32 < it_p and it16 < it_p
    assert abs(it16 - it32) <= max(2, it32 // 10)
    np.testing.assert_allclose(h16[:4], h32[:4], rtol=2e-2)
    net.destroy()
    g.destroy()


@pytest.mark.gpu
def test_stall_restarts_the_coarse_form_before_falling_back():
    """A stall test no solve can pass (|r| to drop 100x over 2 iterations):
    the fine form falls back at its first test (iteration 3); the coarse form
    restarts its directions there, starts the window afresh and falls back at
    the next test (iteration 6); both then converge on the plain V-cycle."""
    cfg = msbgx.config_c4(n=128)
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    for mode, expect in ((msbgx.NET_FINE, 3), (msbgx.NET_COARSE, 6)):
        net = msbgx.LearnedNet(seed=2, mode=mode)
        net.set_stall(2, 0.99)
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr, fb = net.solve(g, max_iter=1000, tol=1e-6)
        print(f"  mode {mode}: {it} iterations rel {rr:.2e} fallback at {fb}", flush=True)
        assert fb == expect and rr <= 1e-6
        net.destroy()
    g.destroy()
