"""The allocation-free entry point msbgx_pde_mean_curvature_noalloc (static
per-GPU work space, no device allocation) against the allocating entry point
msbgx_pde_mean_curvature, through the C ABI: bit-identical uint16 output on
two calls in a row, the second with a smaller grid and block size (the work
space keeps the first call's contents), an unsorted active list with blocks
of other levels, and the size limit rejected before any device work."""
import numpy as np
import pytest

import msbgx

pytestmark = pytest.mark.gpu


def field(rng, shape):
    nz, ny, nx = shape
    zz, yy, xx = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                             np.arange(nx, dtype=np.float64), indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        c = rng.uniform(0.2, 0.8, 3) * (nx, ny, nz)
        r = rng.uniform(0.1, 0.3) * min(shape)
        d = np.sqrt((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2)
        f = np.maximum(f, 0.5 - 0.5 * np.tanh((d - r) / 2.5))
    f += 0.02 * rng.standard_normal(f.shape)
    return np.round(np.clip(f, 0, 1) * 65535).astype(np.uint16)


def grid(rng, shape, bsx, levels, p0):
    sx, sy, sz = shape
    g = msbgx.MultiresSparseGrid.create(sx, sy, sz, bsx, levels)
    nb = (sx // bsx) * (sy // bsx) * (sz // bsx)
    lv = rng.choice([0, levels - 1], size=nb, p=[p0, 1 - p0]).astype(np.int32)
    g.set_refinement_map(g.regularize_refinement_map(lv))
    return g, nb


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_equals_allocating_entry_point(order):
    rng = np.random.default_rng(2024)
    cases = [((128, 96, 160), 16, 3, 0.7, 0.05), ((48, 32, 64), 8, 2, 0.6, 0.1)]
    for shape, bsx, levels, p0, dt in cases:
        g, nb = grid(rng, shape, bsx, levels, p0)
        rho = field(rng, shape[::-1])
        active = rng.permutation(nb)[: int(0.85 * nb)].astype(np.int32)   # unsorted, mixed levels
        print(f"no-alloc case shape={shape} bsx={bsx} blocks={nb} active={len(active)}", flush=True)
        ref, _ = g.pde_mean_curvature(rho, active, num_iter=5, time_step=dt)
        got = g.pde_mean_curvature_noalloc(rho, active, num_iter=5, time_step=dt)
        changed = int(np.count_nonzero(ref != rho))
        print(f"  {changed} voxels changed; equal: {np.array_equal(got, ref)}", flush=True)
        assert changed > 0
        np.testing.assert_array_equal(got, ref)
        g.destroy()


def test_noalloc_rejects_oversize_input():
    # 384^3 uint16 is 113 MB > MSBGX_PDE_NOALLOC_SCRATCH_BYTES (80 MiB): rejected on the host
    g = msbgx.MultiresSparseGrid.create(384, 384, 384, 16, 2)
    rho = np.zeros((384, 384, 384), np.uint16)
    with pytest.raises(msbgx.MsbgxError) as e:
        g.pde_mean_curvature_noalloc(rho, np.arange(8, dtype=np.int32))
    assert e.value.rc == msbgx.EINVAL
    assert "scratch" in msbgx.lib.msbgx_last_error().decode()
    # block ids out of range are rejected as by the allocating entry point
    g2 = msbgx.MultiresSparseGrid.create(32, 32, 32, 16, 2)
    with pytest.raises(msbgx.MsbgxError):
        g2.pde_mean_curvature_noalloc(np.zeros((32, 32, 32), np.uint16), np.array([8], np.int32))
