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


C8 = msbgx.PDE_OPT_8_COLOR
# (type, scheme flag, anis_diff_alpha, taubin, block edges)
PDE_CASES = [
    (4, C8, 0.0, 0.0, (32, 8)),       # the demo's type at block edge 32 (dynamic LDS) and 8
    (4, C8, 0.0, 0.5, (16,)),         # Taubin alternation
    (2, C8, 0.0, 0.0, (16, 8, 4)),    # bi-Laplacian: two-voxel halo
    (1, 0, 0.0, 0.0, (16,)),          # red-black, Laplacian times |grad|
    (1, C8, 0.0, 0.3, (16,)),
    (3, 0, 0.0, 0.0, (16,)),
    (17, 0, 0.0, 0.0, (16,)),
    (17, C8, 0.0, 0.0, (16,)),
    (14, 0, 0.15, 0.0, (32, 16, 8)),  # Perona-Malik, rational conductance
    (15, 0, -0.15, 0.0, (32, 16)),    # its Jacobi form, exponential conductance
    (15, C8, 0.15, 0.0, (8,)),
]


@pytest.mark.parametrize("dtype", [np.uint16, np.float32], ids=["u16", "f32"])
@pytest.mark.parametrize("typ,scheme,alpha,taubin,edges", PDE_CASES,
                         ids=[f"t{c[0]}_{'c8' if c[1] else 'rb'}_{i}" for i, c in enumerate(PDE_CASES)])
def test_pde_apply_matches_oracle(typ, scheme, alpha, taubin, edges, dtype):
    """applyChannelPdeFast's types and both in-place schemes, uint16 and
    float channels, multires level maps, a random subset of listed blocks,
    hard constraints on the float runs: bit-identical to the oracle."""
    for bsx in edges:
        n = 64 if bsx < 32 else 96
        rng = np.random.default_rng(typ * 100 + bsx + (7 if dtype == np.float32 else 0))
        levels = 3 if bsx >= 8 else 2
        g = msbgx.MultiresSparseGrid.create(n, n, n, bsx, levels)
        nb = (n // bsx) ** 3
        lv = g.regularize_refinement_map(rng.choice([0, levels - 1], size=nb, p=[0.7, 0.3]).astype(np.int32))
        g.set_refinement_map(lv)
        rho = smooth_field(rng, n)
        data = rho if dtype == np.uint16 else (rho.astype(np.float32) / 65535 * 1.2 - 0.1).astype(np.float32)
        active = np.sort(rng.choice(nb, size=max(1, int(0.8 * nb)), replace=False)).astype(np.int32)
        hard = dtype == np.float32 and typ in (17, 4)
        kw = dict(do_constr_zero_one=1 if dtype == np.uint16 else 0, num_iter=3, time_step=0.08, taubin=taubin,
                  anis_diff_alpha=alpha, dens_thr=0.5 if hard else 0.0, dens_thr_tol1=0.05 if hard else 0.0,
                  dens_thr_tol2=0.03 if hard else 0.0)
        lt = -(typ + scheme)
        print(f"pde_apply type {typ} laplTyp {lt} bsx {bsx} n {n} {dtype.__name__} hard={hard}", flush=True)
        ref = orc.pde_apply(lv, bsx, data, active, lt, **kw)
        assert ref is not None
        got, ms = g.pde_apply(data, active, lt, **kw)
        changed = int(np.count_nonzero(ref != data))
        print(f"  {changed} voxels changed, {ms:.3f} ms", flush=True)
        assert changed > 0
        if dtype == np.float32:
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
        else:
            np.testing.assert_array_equal(got, ref)
        g.destroy()


def test_pde_apply_rejects_undefined_combinations():
    g = msbgx.MultiresSparseGrid.create(64, 64, 64, 8, 3)
    g.set_refinement_map(np.zeros(512, np.int32))
    x = np.full((64, 64, 64), 0.5, np.float32)
    a = np.arange(512, dtype=np.int32)
    for lt, kw in ((-(17 + C8), {}), (-4, {}), (17, {}), (-14, {"anis_diff_alpha": 0.0}),
                   (-(4 + C8), {"grid_spacing": 2.0}), (-(4 + C8), {"time_step": -0.1})):
        with pytest.raises(msbgx.MsbgxError):
            g.pde_apply(x, a, lt, **kw)
    with pytest.raises(msbgx.MsbgxError):
        g.pde_apply(x.astype(np.uint16), a, -(4 + C8), do_constr_zero_one=0)
    g.destroy()
