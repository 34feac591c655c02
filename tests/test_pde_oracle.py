"""CPU pins of the oracle's mean-curvature smoothing (oracle/flip_oracle.cpp,
orc_pde_mean_curvature; src/msbg4.cpp:60-960 as src/msbg_demo.cpp:716-727
calls it), which checks the GPU kernel.

* Fixed points: a constant field and a ramp along x (its level sets are
  planes: zero mean curvature, also under the Neumann halo at every block
  and domain face, since the field does not vary along y and z) come back
  unchanged.
* Closed form: for a radial cone rho = c - b r the curvature term is
  div(grad rho / |grad rho|) |grad rho| = -2 b / r, so two iterations (one
  rounded up to two) move each voxel by about -4 b dt / r.
* Only listed level-0 blocks change; a block of another level is untouched.
"""
import numpy as np

import oracle_lib as orc


def grid_levels(n, bsx):
    return np.zeros((n // bsx) ** 3, np.int32)


def test_constant_and_ramp_are_fixed_points():
    n, bsx = 32, 8
    lv = grid_levels(n, bsx)
    active = np.arange(len(lv), dtype=np.int32)
    const = np.full((n, n, n), 23456, np.uint16)
    np.testing.assert_array_equal(orc.pde_mean_curvature(lv, bsx, const, active, 5, 0.1), const)
    x = np.arange(n, dtype=np.float64)
    ramp = np.broadcast_to(np.round((0.2 + 0.02 * x) * 65535).astype(np.uint16), (n, n, n)).copy()
    np.testing.assert_array_equal(orc.pde_mean_curvature(lv, bsx, ramp, active, 5, 0.1), ramp)


def test_cone_follows_mean_curvature_closed_form():
    n, bsx = 48, 16
    lv = grid_levels(n, bsx)
    active = np.arange(len(lv), dtype=np.int32)
    c, b, dt = 0.9, 0.02, 0.1
    zz, yy, xx = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    r = np.sqrt((xx - 23.5) ** 2 + (yy - 24.25) ** 2 + (zz - 23.75) ** 2)
    rho = np.round(np.clip(c - b * r, 0, 1) * 65535).astype(np.uint16)
    out = orc.pde_mean_curvature(lv, bsx, rho, active, 1, dt)       # 1 -> 2 iterations
    d = (out.astype(np.float64) - rho) / 65535
    sel = (r > 8) & (r < 16)
    expect = -4 * b * dt / r[sel]
    err = np.abs(d[sel] - expect)
    assert np.median(err / np.abs(expect)) < 0.05 and np.max(err) < 0.25 * np.max(np.abs(expect)) + 3 / 65535


def test_only_listed_level0_blocks_change():
    rng = np.random.default_rng(3)
    n, bsx = 32, 16
    lv = np.array([0, 0, 1, 0, 0, 1, 0, 0], np.int32)               # 2x2x2 blocks, two of level 1
    zz, yy, xx = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    rho = np.round(np.clip(0.5 + 0.3 * np.sin(xx / 5) * np.cos(yy / 7) + 0.05 * rng.random((n, n, n)), 0, 1)
                   * 65535).astype(np.uint16)
    active = np.array([0, 2, 3, 5, 7], np.int32)                      # 1, 4, 6 not listed; 2, 5 not level 0
    out = orc.pde_mean_curvature(lv, bsx, rho, active, 3, 0.1)
    for b in range(8):
        bx, by, bz = b % 2, (b // 2) % 2, b // 4
        sl = (slice(bz * bsx, (bz + 1) * bsx), slice(by * bsx, (by + 1) * bsx), slice(bx * bsx, (bx + 1) * bsx))
        changed = np.any(out[sl] != rho[sl])
        assert changed == (b in (0, 3, 7)), b
