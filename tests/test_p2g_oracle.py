"""CPU pins of the oracle's particle-to-grid transfer (oracle/flip_oracle.cpp,
orc_p2g) that checks the GPU transfer (src/flip_sim.cpp:579-828).

* One particle by hand: its cell mass, the faces its trilinear stencil
  reaches on each of the three staggered grids, the normalised velocity,
  phi from the face density and beta.
* An independent float64 numpy scatter (np.add.at over all particles at
  once, no shared code) agrees to FP32 rounding on random liquid and air
  particles, including particles in the one-cell halo outside the domain.
* Invariants: a uniform particle velocity is reproduced exactly on every
  face with weight; cell masses add up to the particles inside the level-0
  blocks; two liquid cells force beta = 1/rho_L on their shared face and two
  air cells 1/rho_G; blocks that are not level 0 receive nothing.
"""
import numpy as np

import oracle_lib as orc
from msbgx import PARTICLE_DTYPE


def particles(rng, n, lo, hi, liquid_frac=0.7):
    p = np.zeros(n, PARTICLE_DTYPE)
    p["pos"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    p["vel"] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    p["phase"] = (rng.random(n) >= liquid_frac).astype(np.int32)
    return p


def test_single_particle_by_hand():
    p = np.zeros(1, PARTICLE_DTYPE)
    p["pos"][0] = (1.25, 1.5, 1.75)
    p["vel"][0] = (1.0, 2.0, 3.0)
    out, act = orc.p2g([0], 4, 4, 4, 4, p)
    assert out["mass"][1, 1, 1] == 1.0 and out["mass"].sum() == 1.0 and act[0] == 1
    # u grid: node (x, y+.5, z+.5); the particle sits at (1.25, 1.0, 1.25) in node units:
    # x-weights .75/.25 on faces 1/2, y-weight 1 on row 1, z-weights .75/.25 on 1/2
    u = out["u"]
    assert set(zip(*np.nonzero(u))) == {(1, 1, 1), (1, 1, 2), (2, 1, 1), (2, 1, 2)}
    assert np.all(u[np.nonzero(u)] == 1.0)
    # v grid: (0.75, 1.5, 1.25): x .25/.75 on 0/1, y .5/.5 on 1/2, z .75/.25 on 1/2
    assert set(zip(*np.nonzero(out["v"]))) == {(z, y, x) for z in (1, 2) for y in (1, 2) for x in (0, 1)}
    # w grid: (0.75, 1.0, 1.75): z .25/.75 on 1/2, y 1 on row 1
    assert set(zip(*np.nonzero(out["w"]))) == {(z, 1, x) for z in (1, 2) for x in (0, 1)}
    # beta on the u face (1,1,1): density 1000 * .5625 = 562.5 (a single liquid
    # particle; the neighbour cell is empty, so no interface clamp)
    rt_min = np.float32(np.log(np.float32(1000.0))) * np.float32(8.0)
    phi = np.sqrt((np.float32(562.5) - rt_min) / np.float32(8000.0))
    beta = np.float32(1.0) / (np.float32(1.0) + phi * np.float32(999.0))
    assert abs(out["beta_u"][1, 1, 1] - beta) <= 1e-6 * beta
    # faces the particle does not reach: phi = 0, beta = 1 / rho_G
    assert out["beta_u"][3, 3, 3] == 1.0


def numpy_scatter(nx, ny, nz, bsx, level, p, rho_l=1000.0, rho_g=1.0, eps=1e-5, beta_min=1e-6, alpha=1.0, ppv=8):
    """float64 scatter of the same transfer, vectorised (independent of the oracle)."""
    pos = p["pos"].astype(np.float64)
    vel = p["vel"].astype(np.float64)
    liq = p["phase"] == 0
    mp = np.where(liq, rho_l, rho_g)
    shape = (nz, ny, nx)
    cell = np.floor(pos).astype(int)
    inside = np.all((cell >= 0) & (cell < [nx, ny, nz]), axis=1)
    mass = np.zeros(shape)
    nliq = np.zeros(shape)
    np.add.at(mass, (cell[inside, 2], cell[inside, 1], cell[inside, 0]), 1.0)
    np.add.at(nliq, (cell[inside, 2], cell[inside, 1], cell[inside, 0]), liq[inside].astype(float))
    res = {"mass": mass}
    sums = []
    for d in range(3):
        g = pos - np.where(np.arange(3) == d, 0.0, 0.5)
        i0 = np.floor(g).astype(int)
        fr = g - i0
        sv, sw, sr = np.zeros(shape), np.zeros(shape), np.zeros(shape)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    dd = np.array([dx, dy, dz])
                    w = np.prod(np.where(dd == 1, fr, 1 - fr), axis=1)
                    c = i0 + dd
                    ok = np.all((c >= 0) & (c < [nx, ny, nz]), axis=1)
                    idx = (c[ok, 2], c[ok, 1], c[ok, 0])
                    np.add.at(sv, idx, vel[ok, d] * w[ok])
                    np.add.at(sw, idx, w[ok])
                    np.add.at(sr, idx, mp[ok] * w[ok])
        res["uvw"[d]] = np.where(sw > eps, sv / np.where(sw > eps, sw, 1), 0.0)
        sums.append(sr)
    # phases and beta
    ph = np.where(mass > eps, np.where(nliq * 2 > np.floor(mass + 0.5), 0, 1), -1)
    lvl = np.asarray(level).reshape(nz // bsx, ny // bsx, nx // bsx)
    data = np.kron(lvl == 0, np.ones((bsx, bsx, bsx), bool)).astype(bool)
    ph = np.where(data, ph, -1)
    rt_min = np.log(rho_l / rho_g) * rho_g * ppv
    den = alpha * ppv * rho_l
    for d, name in enumerate(("beta_u", "beta_v", "beta_w")):
        rt = sums[d]
        phi = np.where(rt < rt_min, 0.0, np.minimum(np.sqrt(np.maximum(rt - rt_min, 0) / den), 1.0))
        ax = 2 - d
        prev = np.full(shape, -1)
        sl = [slice(None)] * 3
        sl[ax] = slice(1, None)
        src = [slice(None)] * 3
        src[ax] = slice(None, -1)
        prev[tuple(sl)] = ph[tuple(src)]
        phi = np.where((prev == 0) & (ph == 0), 1.0, np.where((prev == 1) & (ph == 1), 0.0, phi))
        res[name] = np.maximum(1.0 / (rho_g + phi * (rho_l - rho_g)), beta_min)
    for k in res:
        res[k] = np.where(data, res[k], 0.0)
    return res, data


def test_agrees_with_float64_numpy_scatter():
    rng = np.random.default_rng(3)
    nx, ny, nz, bsx = 32, 16, 16, 8
    level = np.zeros((nz // bsx) * (ny // bsx) * (nx // bsx), np.int32)
    level[[1, 5]] = 1                                    # two blocks without level-0 data
    p = particles(rng, 20000, -1.0, [nx + 1.0, ny + 1.0, nz + 1.0])
    out, act = orc.p2g(level, nx, ny, nz, bsx, p)
    ref, data = numpy_scatter(nx, ny, nz, bsx, level, p)
    np.testing.assert_array_equal(out["mass"], ref["mass"])
    for k in ("u", "v", "w"):
        np.testing.assert_allclose(out[k], ref[k], rtol=2e-5, atol=2e-6)
    for k in ("beta_u", "beta_v", "beta_w"):
        # phi enters through a square root near rhoTildeMin: compare beta itself
        np.testing.assert_allclose(out[k], ref[k], rtol=1e-4, atol=1e-7)
    for b in range(len(level)):
        assert act[b] == (level[b] == 0)                 # every level-0 block holds particles here


def test_invariants():
    rng = np.random.default_rng(5)
    n, bsx = 16, 8
    level = np.array([0, 0, 0, 1, 0, 0, 0, 0], np.int32)
    p = particles(rng, 30000, 0.0, float(n), liquid_frac=1.0)
    p["vel"] = np.array([0.75, -1.5, 2.25], np.float32)      # uniform
    out, _ = orc.p2g(level, n, n, n, bsx, p)
    data = np.kron(level.reshape(2, 2, 2) == 0, np.ones((bsx,) * 3, bool))
    assert out["mass"].sum() == np.count_nonzero(data[tuple(np.floor(p["pos"][:, ::-1]).astype(int).T)])
    for k, c in zip("uvw", (0.75, -1.5, 2.25)):
        nz_ = out[k] != 0
        np.testing.assert_allclose(out[k][nz_], c, rtol=1e-6)
        assert nz_.sum() > 0.5 * data.sum()
    assert not np.any(out["mass"][~data]) and not np.any(out["u"][~data]) and not np.any(out["beta_u"][~data])
    # two liquid cells: beta = 1 / rho_L; two air cells: 1 / rho_G
    assert np.isclose(out["beta_u"][4, 4, 4], 1e-3, rtol=1e-6)
    p["phase"] = 1
    out, _ = orc.p2g(level, n, n, n, bsx, p)
    assert out["beta_u"][4, 4, 4] == 1.0
