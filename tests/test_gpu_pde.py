"""msbg_demo's mean-curvature smoothing on the GPU (src/msbg4.cpp:60-960,
src/msbg_demo.cpp:716-727): the smoothed uint16 channel is bit-identical to
the oracle's, on single-level and multires grids with 16^3 and 8^3 blocks,
for random subsets of listed blocks and both the demo's time steps."""
import numpy as np
import pytest

import msbgx
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def smooth_field(rng, n):
    zz, yy, xx = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    f = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(0.2, 0.8, 3) * n
        r = rng.uniform(0.1, 0.3) * n
        d = np.sqrt((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2)
        f = np.maximum(f, 0.5 - 0.5 * np.tanh((d - r) / 2.5))
    f += 0.02 * rng.standard_normal(f.shape)
    return np.round(np.clip(f, 0, 1) * 65535).astype(np.uint16)


@pytest.mark.parametrize("n,bsx,levels,multires,dt", [(64, 16, 3, False, 0.05), (64, 16, 3, True, 0.1),
                                                      (64, 8, 3, True, 0.05)])
def test_pde_matches_oracle(n, bsx, levels, multires, dt):
    rng = np.random.default_rng(n + bsx + int(multires))
    g = msbgx.MultiresSparseGrid.create(n, n, n, bsx, levels)
    nb = (n // bsx) ** 3
    lv = rng.choice([0, levels - 1], size=nb, p=[0.6, 0.4]).astype(np.int32) if multires else np.zeros(nb, np.int32)
    lv = g.regularize_refinement_map(lv)
    g.set_refinement_map(lv)
    rho = smooth_field(rng, n)
    active = np.sort(rng.choice(nb, size=int(0.8 * nb), replace=False)).astype(np.int32)
    got, ms = g.pde_mean_curvature(rho, active, num_iter=5, time_step=dt)
    ref = orc.pde_mean_curvature(lv, bsx, rho, active, 5, dt)
    changed = int(np.count_nonzero(ref != rho))
    print(f"pde n={n} bsx={bsx} multires={multires} dt={dt}: {changed} voxels changed, {ms:.3f} ms", flush=True)
    assert changed > 0
    np.testing.assert_array_equal(got, ref)
