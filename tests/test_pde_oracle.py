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


# ---------------------------------------------------------------------------
# The other types and the red-black scheme of applyChannelPdeFast
# (orc_pde_apply; src/msbg4.cpp:100-112, 411-434, 555-820).

C8 = 1 << 29   # OPT_8_COLOR_SCHEME


def _block_slices(b, bsx, nb):
    bx, by, bz = b % nb, (b // nb) % nb, b // (nb * nb)
    return (slice(bz * bsx, (bz + 1) * bsx), slice(by * bsx, (by + 1) * bsx), slice(bx * bsx, (bx + 1) * bsx))


def test_constant_is_a_fixed_point_of_every_type_and_scheme():
    n, bsx = 32, 16
    lv = grid_levels(n, bsx)
    active = np.arange(len(lv), dtype=np.int32)
    for dtype, val in ((np.uint16, 40000), (np.float32, 0.375)):
        const = np.full((n, n, n), val, dtype)
        for typ, alpha in ((1, 0), (2, 0), (3, 0), (4, 0), (14, 0.2), (15, 0.2), (14, -0.2), (15, -0.2), (17, 0)):
            schemes = (-(typ + C8),) if typ in (2, 4) else (-(typ + C8), -typ)
            for lt in schemes:
                out = orc.pde_apply(lv, bsx, const, active, lt, num_iter=3, time_step=0.1, anis_diff_alpha=alpha)
                assert out is not None, (typ, lt)
                np.testing.assert_array_equal(out, const, err_msg=f"type {typ} laplTyp {lt} {dtype}")


def test_undefined_combinations_are_declined():
    lv8 = grid_levels(32, 8)
    lv16 = grid_levels(32, 16)
    x = np.full((32, 32, 32), 0.5, np.float32)
    a8, a16 = np.arange(64, dtype=np.int32), np.arange(8, dtype=np.int32)
    assert orc.pde_apply(lv8, 8, x, a8, -(17 + C8)) is None          # fixed 18-voxel rows need 16^3 blocks
    assert orc.pde_apply(lv16, 16, x, a16, -4) is None               # red-black mean curvature
    assert orc.pde_apply(lv16, 16, x, a16, -2) is None               # red-black bi-Laplacian
    assert orc.pde_apply(lv16, 16, x, a16, 17) is None               # unordered scheme
    assert orc.pde_apply(lv16, 16, x, a16, -14, anis_diff_alpha=0.0) is None
    assert orc.pde_apply(lv16, 16, x.astype(np.uint16), a16, -17, do_constr_zero_one=0) is None


def _one_colour(n, bsx):
    """Level map of an n^3 raster and the blocks of red-black colour 1: every
    listed block reads only unlisted, unchanged neighbours."""
    nb = n // bsx
    lv = np.zeros(nb ** 3, np.int32)
    act = [b for b in range(nb ** 3) if ((b % nb) + (b // nb) % nb + b // (nb * nb)) & 1]
    return lv, np.array(act, np.int32), nb


def _interior(n, bsx, b, nb, margin):
    """Voxel mask of block b at least `margin` from its faces."""
    m = np.zeros((n, n, n), bool)
    sl = _block_slices(b, bsx, nb)
    inner = tuple(slice(s.start + margin, s.stop - margin) for s in sl)
    m[inner] = True
    return m


def test_diffusion_types_on_a_quadratic_closed_form():
    """phi = c |r|^2: the 7-point Laplacian is 6c exactly, so on a block whose
    neighbours do not move, its voxels two or more from the faces shift by
    a constant per iteration: Jacobi diffusion (17) by 6 c dt / (1 + 6 dt)
    each time, simple diffusion (3) by c dt."""
    n, bsx = 48, 16
    lv, act, nb = _one_colour(n, bsx)
    c, dt = 1e-4, 0.2
    zz, yy, xx = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    phi = (0.1 + c * ((xx - 20.5) ** 2 + (yy - 23.5) ** 2 + (zz - 25.5) ** 2)).astype(np.float32)
    sel = np.zeros((n, n, n), bool)
    for b in act:
        sel |= _interior(n, bsx, int(b), nb, 2)
    for typ, shift in ((17, 2 * 6 * c * dt / (1 + 6 * dt)), (3, 2 * c * dt)):
        out = orc.pde_apply(lv, bsx, phi, act, -typ, do_constr_zero_one=0, num_iter=2, time_step=dt)
        d = out.astype(np.float64) - phi
        np.testing.assert_allclose(d[sel], shift, rtol=2e-3, err_msg=f"type {typ}")
        unl = np.ones((n, n, n), bool)
        for b in act:
            unl[_block_slices(int(b), bsx, nb)] = False
        np.testing.assert_array_equal(out[unl], phi[unl])


def test_laplacian_times_gradient_on_a_parabola():
    """Type 1, phi = c (x - x0)^2: L = 2c, |grad| = 2c |x - x0| (central
    differences are exact on quadratics).  Iteration 1 adds D1 = 4 c^2 dt
    |x - x0|, linear on each side of x0, so iteration 2 keeps L = 2c and adds
    dt 2c |2c (x - x0) + D1'| with D1' the slope of D1."""
    n, bsx = 48, 16
    lv, act, nb = _one_colour(n, bsx)
    c, dt, x0 = 2e-3, 0.5, 7.5
    x = np.arange(n, dtype=np.float64)
    phi1d = 0.05 + c * (x - x0) ** 2
    phi = np.broadcast_to(phi1d, (n, n, n)).astype(np.float32).copy()
    out = orc.pde_apply(lv, bsx, phi, act, -1, do_constr_zero_one=0, num_iter=2, time_step=dt)
    s = np.sign(x - x0)
    d1 = 4 * c * c * dt * np.abs(x - x0)
    grad2 = np.abs(2 * c * (x - x0) + 4 * c * c * dt * s)
    expect1d = phi1d + d1 + dt * 2 * c * grad2
    sel = np.zeros((n, n, n), bool)
    for b in act:
        sel |= _interior(n, bsx, int(b), nb, 2)
    sel &= np.broadcast_to(np.abs(x - x0) > 2, (n, n, n))
    np.testing.assert_allclose(out.astype(np.float64)[sel], np.broadcast_to(expect1d, (n, n, n))[sel], rtol=2e-6)


def test_bilaplacian_is_the_composed_laplacian():
    """Type 2, phi = s x^4: the 25-point operator is the 7-point Laplacian
    applied twice, which gives 24 s on x^4 (and 0 on the cubic change of
    iteration 1); |grad| by central differences.  Checked on voxels two or
    more from the block faces, the 8-colour scheme with one colour listed."""
    n, bsx = 32, 16
    nb = n // bsx
    lv = np.zeros(nb ** 3, np.int32)
    act = np.array([0], np.int32)
    s, dt = 1e-5, 0.1
    x = np.arange(n, dtype=np.float64)
    phi1d = 0.02 + s * x ** 4

    def lap1d(f):
        return f[2:] + f[:-2] - 2 * f[1:-1]

    def grad1d(f):
        return np.abs(f[2:] - f[:-2]) / 2

    H = np.full(n, np.nan)
    H[2:-2] = lap1d(lap1d(phi1d))
    G = np.full(n, np.nan)
    G[1:-1] = grad1d(phi1d)
    assert np.allclose(H[2:-2], 24 * s)
    phi_1 = phi1d - dt * H * G
    G2 = np.full(n, np.nan)
    G2[1:-1] = grad1d(phi_1)
    H2 = np.full(n, np.nan)
    H2[2:-2] = lap1d(lap1d(phi_1))
    expect = phi_1 - dt * H2 * G2
    phi = np.broadcast_to(phi1d, (n, n, n)).astype(np.float32).copy()
    out = orc.pde_apply(lv, bsx, phi, act, -(2 + C8), do_constr_zero_one=0, num_iter=2, time_step=dt)
    sel = _interior(n, bsx, 0, nb, 2)
    sel &= np.broadcast_to((x >= 4) & (x < bsx - 4), (n, n, n))   # the composed operator's reach, twice
    got = out.astype(np.float64)[sel]
    exp = np.broadcast_to(expect, (n, n, n))[sel]
    assert np.all(np.isfinite(exp))
    np.testing.assert_allclose(got - np.broadcast_to(phi1d, (n, n, n))[sel], exp - np.broadcast_to(phi1d, (n, n, n))[sel],
                               rtol=2e-3, atol=3e-8)


def test_perona_malik_on_a_step():
    """Types 14 and 15 on a step a | b along x (one colour listed, the step
    inside a block): the two voxels at the step move by dt G (b - a) with
    conductance G = 1 / (1 + ((b - a) / alpha)^2) (alpha > 0) or
    exp(-((b - a) / |alpha|)^2) (alpha < 0); a linear ramp is a fixed point."""
    import math
    n, bsx = 32, 16
    nb = n // bsx
    lv = np.zeros(nb ** 3, np.int32)
    act = np.array([0], np.int32)
    a_, b_, dt, x0 = 0.3, 0.6, 0.1, 8
    x = np.arange(n)
    phi = np.broadcast_to(np.where(x < x0, a_, b_), (n, n, n)).astype(np.float32).copy()
    for alpha in (0.2, -0.2):
        def G(g):
            return math.exp(-(g / -alpha) ** 2) if alpha < 0 else 1 / (1 + (g / alpha) ** 2)
        for typ in (14, 15):
            out = orc.pde_apply(lv, bsx, phi, act, -typ, do_constr_zero_one=0, num_iter=2, time_step=dt,
                                anis_diff_alpha=alpha)
            g = b_ - a_
            if typ == 14:
                d1 = dt * G(g) * g
                l1, r1 = a_ + d1, b_ - d1
                # iteration 2 at x0-1: neighbours a (x0-2, unchanged) and r1; at x0: l1 and b
                gl, gr = r1 - l1, l1 - r1
                l2 = l1 + dt * (G(a_ - l1) * (a_ - l1) + G(gl) * gl)
                r2 = r1 + dt * (G(gr) * gr + G(b_ - r1) * (b_ - r1))
            else:
                l1 = (a_ + dt * (G(g) * b_ + 5 * a_)) / (1 + dt * (G(g) + 5))
                r1 = (b_ + dt * (G(g) * a_ + 5 * b_)) / (1 + dt * (G(g) + 5))
                gl, gr = r1 - l1, l1 - r1
                l2 = (l1 + dt * (G(a_ - l1) * a_ + G(gl) * r1 + 4 * l1)) / (1 + dt * (G(a_ - l1) + G(gl) + 4))
                r2 = (r1 + dt * (G(gr) * l1 + G(b_ - r1) * b_ + 4 * r1)) / (1 + dt * (G(gr) + G(b_ - r1) + 4))
            sl = (slice(0, bsx - 1), slice(0, bsx - 1))   # y, z away from the unlisted neighbours
            np.testing.assert_allclose(out[sl + (x0 - 1,)].astype(np.float64), l2, rtol=1e-6)
            np.testing.assert_allclose(out[sl + (x0,)].astype(np.float64), r2, rtol=1e-6)
    ramp = np.broadcast_to((0.2 + 0.01 * x).astype(np.float32), (n, n, n)).copy()
    for typ in (14, 15):
        out = orc.pde_apply(lv, bsx, ramp, act, -typ, do_constr_zero_one=0, num_iter=2, time_step=dt,
                            anis_diff_alpha=0.2)
        inner = (slice(0, bsx), slice(0, bsx), slice(2, bsx))   # x = 0 meets the Neumann wall
        np.testing.assert_allclose(out[inner], ramp[inner], rtol=0, atol=2e-7)


def test_red_black_updates_colour_one_first():
    """Jacobi diffusion on two blocks along x holding a | b (constant in y
    and z, Neumann walls): per iteration the block of colour 1 updates first
    from a snapshot of its halo, then the block of colour 0 reads its
    updated face.  The problem reduces to x alone (each voxel's y and z
    neighbours equal it); the 1-D model below takes the colours in that
    order, and the reversed order gives different face values."""
    n, bsx, dt, a_, b_ = 32, 16, 0.3, 0.25, 0.75
    lv = np.zeros(2, np.int32)                 # a 32 x 16 x 16 domain: two blocks along x
    act = np.array([0, 1], np.int32)          # blocks (0,0,0) colour 0 and (1,0,0) colour 1
    x = np.arange(n)
    f = np.where(x < bsx, a_, b_).astype(np.float64)

    def update(f, lo, hi):
        snap = f.copy()
        for i in range(lo, hi):
            left = snap[i - 1] if i > 0 else snap[i]
            right = snap[i + 1] if i < n - 1 else snap[i]
            f[i] = (snap[i] + dt * (left + right + 4 * snap[i])) / (1 + 6 * dt)

    def model(first, second):
        g = f.copy()
        for _ in range(2):
            update(g, *first)
            update(g, *second)
        return g

    right_first = model((bsx, n), (0, bsx))
    left_first = model((0, bsx), (bsx, n))
    assert np.max(np.abs(right_first - left_first)) > 1e-3
    phi = np.broadcast_to(f.astype(np.float32), (bsx, bsx, n)).copy()
    out = orc.pde_apply(lv, bsx, phi, act, -17, do_constr_zero_one=0, num_iter=2, time_step=dt)
    got = out.astype(np.float64)
    np.testing.assert_allclose(got, np.broadcast_to(right_first, got.shape), rtol=1e-6)


def test_hard_constraints_and_taubin():
    """Hard constraints keep every voxel within [phi0 - tol1, phi0 + tol2]
    and [0, 1]; a Taubin pair with strength 1 on an odd field: iteration 0
    moves by -D, iteration 1 by +D', net near zero (exactly zero for type 1
    on a parabola where D' = D at voxels where |grad| is unchanged)."""
    rng = np.random.default_rng(5)
    n, bsx = 32, 16
    lv = grid_levels(n, bsx)
    active = np.arange(8, dtype=np.int32)
    phi = rng.random((n, n, n)).astype(np.float32)
    tol1, tol2 = 0.01, 0.02
    out = orc.pde_apply(lv, bsx, phi, active, -17, dens_thr=0.5, dens_thr_tol1=tol1, dens_thr_tol2=tol2,
                        num_iter=4, time_step=0.5)
    lo = np.maximum(0.0, phi - np.float32(tol1))
    hi = np.minimum(1.0, phi + np.float32(tol2))
    assert np.all(out >= lo) and np.all(out <= hi)
    assert np.any(out == lo) and np.any(out == hi)
    free = orc.pde_apply(lv, bsx, phi, active, -17, do_constr_zero_one=0, num_iter=4, time_step=0.5)
    assert np.any(free < lo) or np.any(free > hi)
    # Taubin with strength 1 on a constant-gradient field: -D then +D
    x = np.arange(n, dtype=np.float64)
    ramp = np.broadcast_to((0.3 + 0.01 * x).astype(np.float32), (n, n, n)).copy()
    t = orc.pde_apply(lv, bsx, ramp, active, -1, do_constr_zero_one=0, num_iter=2, time_step=0.5, taubin=1.0)
    np.testing.assert_allclose(t[:, :, 2:-2], ramp[:, :, 2:-2], atol=1e-7)
    # at the Neumann walls the ramp is not a fixed point: -D at iteration 0, +D' at iteration 1
    assert not np.array_equal(t[:, :, 0], ramp[:, :, 0])
