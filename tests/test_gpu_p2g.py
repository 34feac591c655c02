"""Particle-to-grid transfer on the GPU (src/flip_sim.cpp:579-828).

* Random liquid and air particles, including particles in the one-cell halo
  outside the domain and beyond it, on single-level and multires grids with
  16^3 and 32^3 blocks: cell mass, the three face velocities and the three
  face betas are bit-identical to the oracle (same summation order), the
  active blocks equal, and a second transfer on the same handle (fewer
  particles) is again identical.
* Scene particles at 256^3 (dam break, 8 per liquid cell, 3.4e7 particles):
  bit-identical to the oracle; against the scene's gather-form velocity and
  right-hand side equal to rounding (summation order, see the test).
"""
import numpy as np
import pytest

import msbgx
import oracle_lib as orc
from msbgx import PARTICLE_DTYPE

pytestmark = pytest.mark.gpu


def assert_bitwise(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def random_particles(rng, n, shape):
    p = np.zeros(n, PARTICLE_DTYPE)
    hi = np.array(shape[::-1], np.float32)
    p["pos"] = rng.uniform(-1.5, hi + 1.5, (n, 3)).astype(np.float32)     # halo and beyond included
    p["vel"] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    p["phase"] = (rng.random(n) < 0.3).astype(np.int32)
    # a dense liquid blob so that phi reaches 1 and the interface clamp acts
    m = n // 3
    p["pos"][:m] = (rng.uniform(0.25, 0.6, (m, 3)) * hi).astype(np.float32)
    p["phase"][:m] = 0
    return p


@pytest.mark.parametrize("n,bsx,levels,multires", [(64, 16, 3, False), (64, 16, 3, True), (128, 32, 3, True)])
def test_p2g_matches_oracle(n, bsx, levels, multires):
    print(f"p2g vs oracle: n={n} bsx={bsx} multires={multires}", flush=True)
    rng = np.random.default_rng(n + bsx + multires)
    g = msbgx.MultiresSparseGrid.create(n, n, n, bsx, levels)
    nb = (n // bsx) ** 3
    lv = rng.choice([0, levels - 1], size=nb, p=[0.5, 0.5]).astype(np.int32) if multires else np.zeros(nb, np.int32)
    lv = g.regularize_refinement_map(lv)
    g.set_refinement_map(lv)
    t = msbgx.ParticleToGrid(g)
    for count in (40 * n ** 2, 9 * n ** 2):
        p = random_particles(rng, count, (n, n, n))
        t.run(p)
        got = t.get()
        ref, act = orc.p2g(lv, n, n, n, bsx, p)
        for k in ref:
            assert_bitwise(got[k], ref[k])
        np.testing.assert_array_equal(t.active_blocks(), np.nonzero(act)[0])
        ms, msg, npart = t.last_times()
        print(f"  {count} particles: bit-identical; {ms:.3f} ms ({msg:.3f} ms gather), "
              f"{int(act.sum())} active blocks", flush=True)
        assert npart == count


def test_scene_particles_at_256():
    """The scene's particles (8 per liquid cell) transferred on the GPU equal
    the oracle bit for bit.  Against the scene's own gather-form velocity
    (which visits particles by the cell that generated them) they agree to
    rounding: a position (float)i + f with f close to 1 rounds to i + 1, and
    the transfer, like the reference, bins it by floor(pos) -- same terms,
    another order."""
    cfg = msbgx.config_c5(n=256, bsx=16, levels=4)
    print(f"p2g of scene particles at {cfg.n}^3", flush=True)
    g = msbgx.build_grid(cfg)
    lv, _ = g.block_info(0)
    t = msbgx.ParticleToGrid(g)
    npart = t.run_scene(cfg.scene)
    ms, msg, _ = t.last_times()
    print(f"  {npart} particles: {ms:.2f} ms ({msg:.2f} ms gather)", flush=True)
    out = t.get()
    hp = orc.scene_particles(cfg.scene)
    assert len(hp) == npart
    ref, act = orc.p2g(lv, cfg.n, cfg.n, cfg.n, cfg.bsx, hp)
    for k in ref:
        assert_bitwise(out[k], ref[k])
    np.testing.assert_array_equal(t.active_blocks(), np.nonzero(act)[0])
    print("  bit-identical to the oracle", flush=True)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    blocks = g.level0_blocks()
    b_scene = np.concatenate([g.get_block(msbgx.CH_DIVERGENCE, 0, int(b), cfg.bsx) for b in blocks])
    g.flip_set_velocity_scene(cfg.scene)
    vs = g.flip_get_velocity()
    same = 0
    for k, a in zip("uvw", vs):
        np.testing.assert_allclose(out[k], a, rtol=1e-5, atol=1e-6)
        same += int(np.count_nonzero(out[k].view(np.uint32) == a.view(np.uint32)))
    frac = same / (3 * a.size)
    print(f"  vs the scene's gather velocity: {frac:.6f} of faces bit-identical", flush=True)
    assert frac > 0.999
    t.to_velocity(g)
    g.flip_rhs_from_velocity(cfg.scene.dt)
    b_p2g = np.concatenate([g.get_block(msbgx.CH_DIVERGENCE, 0, int(b), cfg.bsx) for b in blocks])
    assert np.max(np.abs(b_p2g - b_scene)) <= 1e-5 * np.max(np.abs(b_scene))


def test_p2g_rejects_bad_input():
    g = msbgx.MultiresSparseGrid.create(32, 32, 32, 16, 2)
    t = msbgx.ParticleToGrid(g)
    with pytest.raises(msbgx.MsbgxError) as e:
        t.get()                                           # nothing transferred yet
    assert e.value.rc == msbgx.ESTATE
    p = np.zeros(4, PARTICLE_DTYPE)
    p["pos"] = 3.5
    p["pos"][2, 1] = np.nan
    with pytest.raises(msbgx.MsbgxError) as e:
        t.run(p)                                          # non-finite position
    assert e.value.rc == msbgx.EINVAL
    with pytest.raises(msbgx.MsbgxError) as e:
        t.run(p[:2], rho_g=0.0)                           # invalid density
    assert e.value.rc == msbgx.EINVAL
    t.run(p[:2])
    other = msbgx.MultiresSparseGrid.create(32, 32, 32, 16, 2)
    other.set_refinement_map(np.array([0] + [1] * 7, np.int32))
    flags = np.zeros((32, 32, 32), np.uint16)
    w = [np.ones((32, 32, 32), np.float32) for _ in range(3)]
    other.prepare_pressure_solve_flip(dense=(flags, *w, np.zeros((32, 32, 32), np.float32)))
    with pytest.raises(msbgx.MsbgxError) as e:
        t.to_velocity(other)                              # level map differs from the transfer's
    assert e.value.rc == msbgx.EINVAL
