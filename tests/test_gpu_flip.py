"""FLIP caller steps on the GPU (src/flip_sim.cpp:329-338, 1317-1349).

* Random liquid, face coefficients, velocities (with NaNs) and pressures on
  single-level and multires grids, 16^3 and 32^3 blocks: the device velocity
  round trip, the divergence right-hand side and the pressure-gradient update
  in both sign modes are bit-identical to the oracle; cells of blocks without
  level-0 data read back 0.
* Full size (256^3 sphere scene): the right-hand side from the device
  velocity equals the scene prepare's right-hand side bit for bit, and after
  an FP64-refined solve the corrected update leaves a divergence below 1e-5
  of the original while the literal one doubles it.
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
    for _ in range(4):
        c = rng.uniform(0.2, 0.8, 3) * n
        r = rng.uniform(0.1, 0.3) * n
        mass[(xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r] = 1.0
    mass[: n // 5] = 1.0                                   # a slab touching three walls
    flags = np.where(mass > EPS, 0, msbgx.CELL_VOID).astype(np.uint16)
    beta = [rng.uniform(0.25, 1.5, mass.shape).astype(np.float32) for _ in range(3)]
    g = msbgx.MultiresSparseGrid.create(n, n, n, bsx, levels)
    g.set_domain_boundary_spec(*([msbgx.DBC_SOLID] * 6))
    nb = (n // bsx) ** 3
    lv = rng.choice([0, levels - 1], size=nb, p=[0.5, 0.5]).astype(np.int32) if multires else np.zeros(nb, np.int32)
    g.set_refinement_map(g.regularize_refinement_map(lv))
    g.prepare_pressure_solve_flip(dense=(flags, *beta, np.zeros_like(mass)))
    level, _ = g.block_info(0)
    vel = [rng.normal(0, 1, mass.shape).astype(np.float32) for _ in range(3)]
    liquid = np.argwhere(mass > EPS)
    for k in range(3):
        z, y, x = liquid[rng.integers(len(liquid))]
        vel[k][z, y, x] = np.nan
        for z, y, x in liquid[rng.integers(len(liquid), size=50)]:
            vel[k][z, y, x] = -0.0                         # signed zeros must survive bit for bit
    # pressure: random on level-0 blocks, through the channel
    p = np.zeros_like(mass)
    nbx = n // bsx
    for b in range(nb):
        if level[b] != 0:
            continue
        blk = rng.normal(0, 2, bsx ** 3).astype(np.float32)
        blk[rng.random(bsx ** 3) < 0.05] = 0.0               # zero pressure on some fluid cells
        g.set_block(msbgx.CH_PRESSURE, 0, b, blk)
        bx, by, bz = b % nbx, (b // nbx) % nbx, b // (nbx * nbx)
        p[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = blk.reshape(bsx, bsx, bsx)
    data = np.zeros(mass.shape, bool)
    for b in range(nb):
        if level[b] == 0:
            bx, by, bz = b % nbx, (b // nbx) % nbx, b // (nbx * nbx)
            data[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = True
    return g, level, mass, beta, vel, p, data


def assert_bitwise(a, b):
    """Equal bit patterns (so -0.0 != +0.0); NaNs only need to sit in the same places."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    np.testing.assert_array_equal(na, nb)
    np.testing.assert_array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def dense_channel(g, ch, level, n, bsx):
    out = np.zeros((n, n, n), np.float32)
    nbx = n // bsx
    for b in range(len(level)):
        if level[b] != 0:
            continue
        bx, by, bz = b % nbx, (b // nbx) % nbx, b // (nbx * nbx)
        out[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = \
            g.get_block(ch, 0, b, bsx).reshape(bsx, bsx, bsx)
    return out


@pytest.mark.parametrize("n,bsx,levels,multires", [(64, 16, 3, False), (64, 16, 3, True), (128, 32, 3, True)])
def test_flip_steps_match_oracle(n, bsx, levels, multires):
    print(f"flip steps vs oracle: n={n} bsx={bsx} multires={multires}", flush=True)
    g, level, mass, beta, vel, p, data = dense_case(n, bsx, levels, seed=n + bsx, multires=multires)
    g.flip_set_velocity(*vel)
    back = g.flip_get_velocity()
    for a, a0 in zip(back, vel):
        assert_bitwise(a[data], a0[data])
        assert not np.any(a[~data])
    dt = 0.0625
    g.flip_rhs_from_velocity(dt)
    rhs = dense_channel(g, msbgx.CH_DIVERGENCE, level, n, bsx)
    ref = orc.flip_rhs(level, bsx, mass, EPS, vel, dt)
    assert_bitwise(rhs, ref)
    print(f"  rhs bit-identical ({int(np.count_nonzero(ref))} nonzero cells)", flush=True)
    for corrected in (False, True):
        g.flip_set_velocity(*vel)
        g.flip_update_velocity(dt, corrected=corrected)
        got = g.flip_get_velocity()
        exp = orc.flip_update_velocity(level, bsx, mass, EPS, p, beta, vel, dt, negate_p=corrected)
        for k in range(3):
            assert_bitwise(got[k][data], exp[k][data])
        print(f"  update (corrected={corrected}) bit-identical", flush=True)
    if not multires:
        # the timing entry points' algorithmic byte counts, recounted here
        f = mass > EPS
        N = f.size
        nb_f = []
        for ax in (2, 1, 0):
            sh = np.zeros_like(f)
            sl = [slice(None)] * 3
            sl[ax] = slice(1, None)
            src = [slice(None)] * 3
            src[ax] = slice(None, -1)
            sh[tuple(sl)] = f[tuple(src)]
            first = np.zeros_like(f)
            idx = [slice(None)] * 3
            idx[ax] = 0
            first[tuple(idx)] = True
            nb_f.append((sh, first))
        upd = 14 * N + 4 * int(f.sum()) + 8 * sum(int(((f | sh) & ~first).sum()) for sh, first in nb_f)
        rhs_b = 6 * N + 4 * sum(int((f | sh).sum()) for sh, _ in nb_f)
        _, by3, c3 = g.time_kernel(3, reps=1)
        _, by4, c4 = g.time_kernel(4, reps=1)
        assert c3 == c4 == N and by3 == upd and by4 == rhs_b, (by3, upd, by4, rhs_b)


def test_c1_projection_removes_the_divergence():
    cfg = msbgx.config_c1()
    print(f"flip projection at {cfg.n}^3", flush=True)
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    blocks = g.level0_blocks()
    bsx = cfg.bsx
    b_scene = np.concatenate([g.get_block(msbgx.CH_DIVERGENCE, 0, int(b), bsx) for b in blocks])
    dt = cfg.scene.dt
    g.flip_set_velocity_scene(cfg.scene)
    g.flip_rhs_from_velocity(dt)
    b_dev = np.concatenate([g.get_block(msbgx.CH_DIVERGENCE, 0, int(b), bsx) for b in blocks])
    assert_bitwise(b_dev, b_scene)
    outer, inner, rel, _ = g.solve_pressure_refined(tol=1e-10)
    assert rel <= 1e-10
    n0 = np.linalg.norm(b_scene.astype(np.float64))
    for corrected in (True, False):
        g.flip_set_velocity_scene(cfg.scene)
        g.flip_update_velocity(dt, corrected=corrected)
        g.flip_rhs_from_velocity(dt)
        b1 = np.concatenate([g.get_block(msbgx.CH_DIVERGENCE, 0, int(b), bsx) for b in blocks]).astype(np.float64)
        if corrected:
            r = np.linalg.norm(b1) / n0
            print(f"  corrected: |div v'| / |div v| = {r:.2e}", flush=True)
            assert r <= 1e-5
        else:
            r = np.linalg.norm(b1 - 2 * b_scene) / n0
            print(f"  literal: |div v' - 2 div v| / |div v| = {r:.2e}", flush=True)
            assert r <= 1e-5
