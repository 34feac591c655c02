"""CPU tests of the product library: the C ABI loads and exports every
symbol include/msbgx.h declares, argument errors are reported, and the host
topology (refinement regularisation, block levels/flags, level policy,
relaxation lists, smoother weight) is bit-exact with the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import msbgx
from conftest import ROOT, has_gpu
from oracle_lib import Oracle, build_oracle, level_map_from_scene

HEADERS = [os.path.join(ROOT, "include", "msbgx.h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        names |= set(re.findall(r"\b(msbgx_[a-z0-9_]+)\s*\(", txt))
    names -= {"msbgx_scene_is_liquid", "msbgx_scene_velocity", "msbgx_scene_cell_type", "msbgx_scene_face_coeff",
              "msbgx_scene_rhs", "msbgx_splitmix64"}
    return names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", msbgx.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (msbgx_[a-z0-9_]+)", out.stdout))
    missing = declared_symbols() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    assert set(msbgx.EXPORTED) == declared_symbols()


def test_create_rejects_bad_arguments():
    with pytest.raises(msbgx.MsbgxError):
        msbgx.MultiresSparseGrid.create(100, 64, 64, 16, 3)   # not a multiple of the block size
    with pytest.raises(msbgx.MsbgxError):
        msbgx.MultiresSparseGrid.create(64, 64, 64, 12, 3)    # not a power of two
    with pytest.raises(msbgx.MsbgxError):
        msbgx.MultiresSparseGrid.create(64, 64, 64, 4, 3)     # 4 >> 2 < 2
    g = msbgx.MultiresSparseGrid.create(64, 64, 64, 16, 3)
    with pytest.raises(msbgx.MsbgxError):
        g.set_refinement_map(np.array([0, 2] + [2] * 62, np.int32))  # not 2:1 graded
    with pytest.raises(msbgx.MsbgxError):
        g.solve_pressure_flip()                               # solve before prepare


def test_flip_steps_reject_bad_calls():
    g = msbgx.MultiresSparseGrid.create(64, 64, 64, 16, 3)
    z = np.zeros((64, 64, 64), np.float32)
    for call in (lambda: g.flip_update_velocity(0.1), lambda: g.flip_rhs_from_velocity(0.1),
                 lambda: g.flip_set_velocity(z, z, z), lambda: g.flip_get_velocity()):
        with pytest.raises(msbgx.MsbgxError) as e:
            call()                                            # no prepare yet
        assert e.value.rc == msbgx.ESTATE
    with pytest.raises(ValueError):
        g.flip_set_velocity(z[:10], z, z)
    for dt in (float("nan"), float("inf")):
        with pytest.raises(msbgx.MsbgxError) as e:
            g.flip_update_velocity(dt)
        assert e.value.rc == msbgx.EINVAL
    with pytest.raises(msbgx.MsbgxError) as e:
        g.flip_rhs_from_velocity(0.0)
    assert e.value.rc == msbgx.EINVAL
    rc = msbgx.lib.msbgx_flip_update_velocity(g._h, 0.1, 7)
    assert rc == msbgx.EINVAL


@pytest.mark.parametrize("seed", [0, 1])
def test_regularize_and_flags_match_oracle(seed):
    rng = np.random.default_rng(seed)
    g = msbgx.MultiresSparseGrid.create(128, 96, 64, 16, 4)
    o = Oracle(128, 96, 64, 16, 4)
    lv = rng.choice([0, 3], size=g.n_blocks, p=[0.15, 0.85]).astype(np.int32)
    a, b = g.regularize_refinement_map(lv), o.regularize_refinement_map(lv)
    assert np.array_equal(a, b)
    g.set_refinement_map(a)
    o.set_refinement_map(b)
    for mg in range(4):
        la, fa = g.block_info(mg)
        lb, fb = o.block_info(mg)
        assert np.array_equal(la, lb) and np.array_equal(fa, fb)


@pytest.mark.parametrize("mk", [lambda: msbgx.config_c2(n=256, bsx=16), lambda: msbgx.config_c4(n=128)])
def test_level_policy_matches_oracle(mk):
    cfg = mk()
    a = msbgx.level_map(cfg)
    b = level_map_from_scene(cfg.scene, cfg.bsx, cfg.levels, cfg.level_mode, cfg.level_d0)
    assert np.array_equal(a, b)


@pytest.mark.skipif(has_gpu(), reason="exercises the host-only metadata path of a failed device upload")
@pytest.mark.parametrize("mk", [lambda: msbgx.config_c1(n=64), lambda: msbgx.config_c4(n=128),
                                lambda: msbgx.config_c2(n=256, bsx=32)])
def test_host_metadata_matches_oracle_without_device(mk):
    """Without a device, prepare fails at the upload with MSBGX_EHIP after the
    host metadata is built; block flags, relaxation lists and omega are then
    compared with the oracle."""
    cfg = mk()
    g = msbgx.build_grid(cfg)
    with pytest.raises(msbgx.MsbgxError) as e:
        g.prepare_pressure_solve_flip(scene=cfg.scene)
    assert e.value.rc == -3
    o = build_oracle(cfg)
    o.prepare_scene(cfg.scene)
    assert g.omega() == o.omega()
    for mg in range(cfg.levels):
        assert np.array_equal(g.block_info(mg)[1], o.block_info(mg)[1])
        assert np.array_equal(g.relax_list(mg), o.relax_list(mg))


def test_scene_functions_host_device_contract():
    """The scene header is a pure function of integers: spot-check values."""
    s = msbgx.make_scene(32, spheres=[(0.5, 0.5, 0.5, 0.25)])
    o = Oracle(32, 32, 32, 8, 2)
    o.set_refinement_map(np.zeros(o.n_blocks, np.int32))
    o.prepare_scene(s)
    fl = o.get_block(11, 0, 0, 8)
    assert np.all((fl & msbgx.CELL_VOID) != 0)            # corner block is air
    center = (1 + 4 * (1 + 4 * 1))                       # block (1,1,1) holds the sphere centre
    fl = o.get_block(11, 0, center, 8).reshape(8, 8, 8)
    assert (fl[7, 7, 7] & msbgx.CELL_VOID) == 0


@pytest.mark.parametrize("what", ["face", "rhs", "inf_face"])
def test_non_finite_inputs_are_rejected_before_upload(what):
    """NaN / Inf in the face coefficients or the right-hand side is an
    argument error (reported before any device work)."""
    n = 32
    g = msbgx.MultiresSparseGrid.create(n, n, n, 16, 2)
    flags = np.zeros((n, n, n), np.uint16)
    fu, fv, fw = (np.ones((n, n, n), np.float32) for _ in range(3))
    rhs = np.zeros((n, n, n), np.float32)
    rhs[5, 5, 5] = 1.0
    if what == "face":
        fv[3, 4, 5] = np.nan
    elif what == "inf_face":
        fu[7, 7, 7] = np.inf
    else:
        rhs[9, 9, 9] = np.nan
    with pytest.raises(msbgx.MsbgxError) as e:
        g.prepare_pressure_solve_flip(dense=(flags, fu, fv, fw, rhs))
    assert e.value.rc == -1 and "non-finite" in str(e.value)
