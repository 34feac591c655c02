"""CPU pins of the particle right-hand side (the 1024^3 full-scale workload).

msbgx_scene_p2g (include/msbgx_scene.h) computes the particle-to-grid MAC
velocity as a fixed-order gather per face; the reference scatters each
particle to the 8 nodes of each staggered grid and normalises by the weight
sum (src/flip_sim.cpp:659-690, 771-779).  Here an independent numpy scatter
of the same seeded particles (np.add.at, FP64 accumulation) reproduces the
oracle's divergence channel -- which the GPU matches bit for bit in the GPU
parity tests -- to FP32 summation tolerance.  A second test pins the
particle generator's moments (uniform positions, Irwin-Hall velocities)."""
import numpy as np

import msbgx
from oracle_lib import build_oracle

MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix(x):
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def particles(scene, cells):
    """cells: int array [n, 3] (i, j, k) -> pos [n*ppc, 3] f32, vel [n*ppc, 3] f32."""
    ppc = scene.particles_per_cell
    i, j, k = (cells[:, a].astype(np.uint64) for a in range(3))
    cell = (k * np.uint64(scene.ny) + j) * np.uint64(scene.nx) + i
    with np.errstate(over="ignore"):
        s0 = (np.uint64(scene.seed) ^ np.uint64(0x5851F42D4C957F2D)) * np.uint64(0x2545F4914F6CDD1D)
        p = np.arange(ppc, dtype=np.uint64)
        base = s0 + (cell[:, None] * np.uint64(ppc) + p[None, :]) * np.uint64(16)
    base = base.reshape(-1)
    h = splitmix(base)
    f21 = np.float32(1.0 / 2097152.0)
    c = np.repeat(cells, ppc, axis=0).astype(np.float32)
    pos = np.stack([c[:, 0] + (h & np.uint64(0x1FFFFF)).astype(np.float32) * f21,
                    c[:, 1] + ((h >> np.uint64(21)) & np.uint64(0x1FFFFF)).astype(np.float32) * f21,
                    c[:, 2] + ((h >> np.uint64(42)) & np.uint64(0x1FFFFF)).astype(np.float32) * f21], 1)
    vel = np.zeros_like(pos)
    for d in range(3):
        acc = np.zeros(len(base), np.float32)
        for w in range(3):
            g = splitmix(base + np.uint64(1 + 3 * d + w))
            for q in range(4):
                acc = acc + ((g >> np.uint64(16 * q)) & np.uint64(0xFFFF)).astype(np.float32) * np.float32(1 / 65536)
        vel[:, d] = (acc - np.float32(6)) * np.float32(scene.particle_sigma)
    return pos, vel


def scatter_mac(scene, pos, vel, d):
    """Reference-style scatter of component d (flip_sim.cpp:659-690): nodes offset by 0.5 in the
    two other axes, trilinear weights, normalised by the weight sum (> 1e-5)."""
    n = np.array([scene.nx, scene.ny, scene.nz])
    g = pos.astype(np.float32).copy()
    for a in range(3):
        if a != d:
            g[:, a] = g[:, a] - np.float32(0.5)
    i0 = np.floor(g).astype(np.int64)
    fr = g - i0.astype(np.float32)
    num = np.zeros(tuple(n[::-1]))
    den = np.zeros(tuple(n[::-1]))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                off = np.array([dx, dy, dz])
                w = (np.where(dx, fr[:, 0], 1 - fr[:, 0]) * np.where(dy, fr[:, 1], 1 - fr[:, 1])) * \
                    np.where(dz, fr[:, 2], 1 - fr[:, 2])
                idx = i0 + off
                ok = np.all((idx >= 0) & (idx < n), axis=1)
                np.add.at(num, (idx[ok, 2], idx[ok, 1], idx[ok, 0]), vel[ok, d].astype(np.float64) * w[ok])
                np.add.at(den, (idx[ok, 2], idx[ok, 1], idx[ok, 0]), w[ok].astype(np.float64))
    return np.where(den > 1e-5, num / np.where(den > 0, den, 1), 0.0)


def test_particle_rhs_equals_reference_style_scatter():
    n = 32
    scene = msbgx.make_scene(n, spheres=[(0.45, 0.5, 0.55, 0.3)], seed=77, particles_per_cell=8,
                             particle_sigma=0.5)
    cfg = msbgx.Config("p2g_pin", n, 16, 2, scene, level_mode=-1)
    o = build_oracle(cfg, threads=4)
    o.prepare_scene(scene)
    nb = n // 16
    flags = np.zeros((n, n, n), np.uint16)
    rhs = np.zeros((n, n, n))
    for b in range(nb ** 3):
        bx, by, bz = b % nb, (b // nb) % nb, b // (nb * nb)
        sl = (slice(16 * bz, 16 * bz + 16), slice(16 * by, 16 * by + 16), slice(16 * bx, 16 * bx + 16))
        flags[sl] = o.get_block(msbgx.CH_FLAGS, 0, b, 16).reshape(16, 16, 16)
        rhs[sl] = o.get_block(msbgx.CH_DIVERGENCE, 0, b, 16).reshape(16, 16, 16)
    liquid = (flags & msbgx.CELL_VOID) == 0
    zz, yy, xx = np.nonzero(liquid)
    pos, vel = particles(scene, np.stack([xx, yy, zz], 1))
    mac = [scatter_mac(scene, pos, vel, d) for d in range(3)]
    pad = [np.pad(m, [(0, 1) if a == 2 - d else (0, 0) for a in range(3)]) for d, m in enumerate(mac)]
    u, v, w = pad
    div = (((u[:, :, 1:] - u[:, :, :-1]) + v[:, 1:, :]) - v[:, :-1, :] + w[1:, :, :]) - w[:-1, :, :]
    ref = np.where(liquid, div / scene.dt, 0.0)
    err = np.abs(rhs - ref).max()
    print(f"p2g: {liquid.sum()} liquid cells, max |rhs - scatter| {err:.2e}, max |rhs| {np.abs(ref).max():.2f}")
    assert err <= 2e-5 * max(1.0, np.abs(ref).max())
    assert np.abs(ref).max() > 0.01


def test_particle_generator_moments():
    scene = msbgx.make_scene(64, seed=5, particles_per_cell=8, particle_sigma=0.5)
    cells = np.stack(np.meshgrid(np.arange(20), np.arange(20), np.arange(20), indexing="ij"), -1).reshape(-1, 3)
    pos, vel = particles(scene, cells)
    frac = pos - np.repeat(cells, 8, axis=0)
    assert frac.min() >= 0 and frac.max() < 1 and abs(frac.mean() - 0.5) < 0.01
    assert abs(vel.mean()) < 0.01 and abs(vel.std() - 0.5) < 0.01
    assert np.abs(vel).max() <= 3.0                       # Irwin-Hall: bounded by 6 sigma
