"""Failure handling on the GPU, forced from the host (every kernel exits
normally): iteration limits, grids without liquid, repeated solves, and the
learned solve's fallback reporting."""
import numpy as np
import pytest

import msbgx

pytestmark = pytest.mark.gpu


def prepared(cfg):
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    return g


def test_max_iter_one_reports_the_limit():
    print("failure: max_iter=1", flush=True)
    g = prepared(msbgx.config_c1(n=64))
    it, rr = g.solve_pressure_flip(max_iter=1, tol=1e-6)
    assert it == 1 and np.isfinite(rr) and rr > 1e-6
    assert len(g.residual_history()) == 1


def test_zero_tolerance_runs_exactly_max_iter():
    g = prepared(msbgx.config_c1(n=64))
    it, rr = g.solve_pressure_flip(max_iter=7, tol=0.0)
    assert it == 7 and len(g.residual_history()) == 7


def test_grid_without_liquid_returns_immediately():
    print("failure: all air", flush=True)
    cfg = msbgx.Config("air", 64, 16, 3, msbgx.make_scene(64, spheres=[]), level_mode=-1)
    g = prepared(cfg)
    it, rr = g.solve_pressure_flip(max_iter=100, tol=1e-6)
    assert it == 0 and rr == 0.0
    net = msbgx.LearnedNet(seed=1)
    it, rr, fb = net.solve(g, max_iter=100, tol=1e-6)
    assert it == 0 and rr == 0.0


def test_solve_accumulates_into_the_pressure_like_the_reference():
    """The reference PCG starts from r = b and adds to x (src/msbg3.cpp:3107-3195);
    a second solve therefore adds a second correction.  After zeroing the
    pressure, a re-solve reproduces the first result exactly."""
    g = prepared(msbgx.config_c1(n=64))
    it1, rr1 = g.solve_pressure_flip(max_iter=200, tol=1e-6)
    p1 = g.get_block(msbgx.CH_PRESSURE, 0, int(g.level0_blocks()[0]), 16).copy()
    g.zero_channel(msbgx.CH_PRESSURE, 0)
    it2, rr2 = g.solve_pressure_flip(max_iter=200, tol=1e-6)
    p2 = g.get_block(msbgx.CH_PRESSURE, 0, int(g.level0_blocks()[0]), 16)
    assert (it1, rr1) == (it2, rr2)
    np.testing.assert_array_equal(p1, p2)
