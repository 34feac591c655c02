"""Pins of the oracle's precision-form smoother (CPU only): the 16-bit
rounding against numpy's binary16 conversion and against nearest-even bfloat16
computed in float64, and the 16-bit relax equal to the FP32 relax over
operands rounded beforehand."""
import numpy as np
import pytest

import msbgx
import oracle_lib
from oracle_lib import Oracle
from test_oracle_pins import blocks_to_field, field_to_blocks

CH_B, CH_R, CH_Z, CH_DIAG = 1, 2, 3, 10
F32 = np.float32


def sample_floats(rng, n):
    """Random float32 bit patterns over every exponent, plus edge values."""
    bits = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    v = bits.view(F32)
    edge = np.array([0.0, -0.0, 1.0, 65504.0, 65519.996, 65520.0, 6.1035156e-05, 5.9604645e-08, 2.9802322e-08,
                     2.9802326e-08, 1.0009766, 1.0004883, 3.3895314e38, np.inf, -np.inf], F32)
    return np.concatenate([v, edge])


def test_half_rounding_equals_numpy():
    rng = np.random.default_rng(1)
    v = sample_floats(rng, 200000)
    got = oracle_lib.round16(v, 1)
    with np.errstate(over="ignore"):
        exp = v.astype(np.float16).astype(F32)
    nan = np.isnan(exp)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), exp[~nan].view(np.uint32))


def test_bf16_rounding_is_nearest_even():
    rng = np.random.default_rng(2)
    v = sample_floats(rng, 200000)
    v = v[np.isfinite(v) & (np.abs(v) < 3.3e38)]
    got = oracle_lib.round16(v, 2)
    bits = v.view(np.uint32)
    lo = (bits & 0xFFFF0000).view(F32).astype(np.float64)          # truncated toward zero
    hi_bits = (bits & 0xFFFF0000) + 0x10000                         # next bfloat16 away from zero
    hi = hi_bits.astype(np.uint32).view(F32).astype(np.float64)
    x = v.astype(np.float64)
    dlo, dhi = np.abs(x - lo), np.abs(hi - x)
    lo_even = ((bits >> 16) & 1) == 0
    pick_hi = (dhi < dlo) | ((dhi == dlo) & ~lo_even)
    exp = np.where(pick_hi, hi, lo).astype(F32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert np.count_nonzero(got != v) > 0


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("reverse", [False, True])
def test_16bit_relax_is_fp32_relax_over_rounded_operands(prec, reverse):
    """Face weights k/16 are exact in both formats; the inverse diagonal
    1/sum is not.  The 16-bit relax of oracle A must equal the FP32 relax of
    oracle B whose diagonal was replaced by its rounded copy."""
    n, bsx = 32, 8
    rng = np.random.default_rng(10 * prec + reverse)
    void = rng.random((n, n, n)) < 0.1
    flags = np.where(void, msbgx.CELL_VOID, 0).astype(np.uint16)
    w = [(rng.integers(4, 28, (n, n, n)) / 16.0).astype(F32) for _ in range(3)]
    rhs = rng.standard_normal((n, n, n)).astype(F32)
    side = []
    for k in range(2):
        o = Oracle(n, n, n, bsx, 2)
        o.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
        o.set_refinement_map(np.zeros(o.n_blocks, np.int32))
        o.prepare_dense((n, n, n), flags, *w, rhs)
        side.append(o)
    a, b = side
    assert a.omega() == b.omega()
    a.set_smoother_precision(prec)
    d = blocks_to_field(b, n, bsx, CH_DIAG)
    dr = oracle_lib.round16(d, prec)
    assert np.count_nonzero(dr != d) > 0
    field_to_blocks(b, dr, n, bsx, CH_DIAG)
    x0 = rng.standard_normal((n, n, n)).astype(F32)
    for o in (a, b):
        field_to_blocks(o, x0, n, bsx, CH_Z)
        o.copy_channel(CH_B, CH_R)
        o.relax(0, reverse, 2, CH_Z, CH_R)
    ga, gb = blocks_to_field(a, n, bsx, CH_Z), blocks_to_field(b, n, bsx, CH_Z)
    assert np.array_equal(ga, gb)
    # and the rounding changes the result against the FP32 relax of A's operands
    c = Oracle(n, n, n, bsx, 2)
    c.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    c.set_refinement_map(np.zeros(c.n_blocks, np.int32))
    c.prepare_dense((n, n, n), flags, *w, rhs)
    field_to_blocks(c, x0, n, bsx, CH_Z)
    c.copy_channel(CH_B, CH_R)
    c.relax(0, reverse, 2, CH_Z, CH_R)
    assert np.count_nonzero(blocks_to_field(c, n, bsx, CH_Z) != ga) > 0


# ------------------------------------------------ FP64 coarsest-level relax

def _fma64(a, b, c):
    """Correctly rounded a * b + c in float64 (exact product and sum)."""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def brute_coarse64(b, w, D, omega, nb, e, h, relaxed, seg, n_iter, bc_plus_w):
    """The coarsest level's FP64 relax restated over Python floats from x = 0:
    n_iter iterations of colour 1 then colour 0 over the listed blocks of a
    raster of same-level blocks; per block a halo snapshot of h*x (0 outside
    the domain), five local sweeps, forward, reverse, ... (sweep k reversed
    iff k is odd), SIMD segments of `seg` lanes along x Jacobi inside and
    Gauss-Seidel across, the fifth sweep reading only the fourth's values and
    stored divided by h; b, W and D are the FP32 values read exactly."""
    nbx, nby, nbz = nb
    m = (nbz * e, nby * e, nbx * e)
    x = np.zeros(m, np.float64)
    om = float(np.float32(omega))
    omm = 1.0 - om
    for _ in range(n_iter):
        for colour in (1, 0):
            snap = x.copy()
            for bid in relaxed:
                bx, by, bz = bid % nbx, (bid // nbx) % nby, bid // (nbx * nby)
                if (bx + by + bz) % 2 != colour:
                    continue
                z0, y0, x0 = bz * e, by * e, bx * e
                X = np.zeros((e + 2, e + 2, e + 2))
                for k in range(-1, e + 1):
                    for j in range(-1, e + 1):
                        for i in range(-1, e + 1):
                            inner = 0 <= i < e and 0 <= j < e and 0 <= k < e
                            face = sum(not (0 <= t < e) for t in (i, j, k)) == 1
                            gz, gy, gx = z0 + k, y0 + j, x0 + i
                            if (inner or face) and 0 <= gz < m[0] and 0 <= gy < m[1] and 0 <= gx < m[2]:
                                X[k + 1, j + 1, i + 1] = snap[gz, gy, gx] * h
                skip = np.zeros((e, e, e // seg), bool)
                for k in range(e):
                    for j in range(e):
                        for s_ in range(e // seg):
                            skip[k, j, s_] = all(D[z0 + k, y0 + j, x0 + s_ * seg + l] == 0 for l in range(seg))

                def wp(kdir, k, j, i):
                    g = [z0 + k, y0 + j, x0 + i]
                    g[2 - kdir] += 1
                    if g[2 - kdir] >= m[2 - kdir]:
                        return float(bc_plus_w)
                    return float(w[kdir][g[0], g[1], g[2]])

                for sw in range(5):
                    rev = (sw & 1) == 0
                    last = sw == 4
                    out = {}
                    for kk in (range(e - 1, -1, -1) if rev else range(e)):
                        for jj in (range(e - 1, -1, -1) if rev else range(e)):
                            for ss in (range(e // seg - 1, -1, -1) if rev else range(e // seg)):
                                if skip[kk, jj, ss]:
                                    continue
                                ys = []
                                for l in range(seg):
                                    ii = ss * seg + l
                                    c = (z0 + kk, y0 + jj, x0 + ii)
                                    K, J, I = kk + 1, jj + 1, ii + 1
                                    sm = X[K, J, I - 1] * float(w[0][c])
                                    sm = _fma64(X[K, J, I + 1], wp(0, kk, jj, ii), sm)
                                    sm = _fma64(X[K, J - 1, I], float(w[1][c]), sm)
                                    sm = _fma64(X[K, J + 1, I], wp(1, kk, jj, ii), sm)
                                    sm = _fma64(X[K - 1, J, I], float(w[2][c]), sm)
                                    sm = _fma64(X[K + 1, J, I], wp(2, kk, jj, ii), sm)
                                    d = float(D[c])
                                    f0 = X[K, J, I]
                                    ys.append(_fma64(om * (float(b[c]) + sm), d, omm * f0) if d != 0 else f0)
                                for l in range(seg):
                                    if last:
                                        out[(kk, jj, ss * seg + l)] = ys[l] / h
                                    else:
                                        X[kk + 1, jj + 1, ss * seg + l + 1] = ys[l]
                for (kk, jj, ii), v in out.items():
                    x[z0 + kk, y0 + jj, x0 + ii] = v
    return x


@pytest.mark.parametrize("bsx0,nbe,lev,seg", [(8, 4, 1, 4), (16, 4, 1, 8), (4, 4, 1, 2)])
def test_coarsest_relax_fp64_brute_force(bsx0, nbe, lev, seg):
    """Two MG levels, every block at level 1 on a 4x4x4 raster: the coarsest
    level's relax list holds the 2x2x2 blocks whose doubled-coordinate
    samples reach fluid, each with relaxed and unrelaxed neighbours.  The
    oracle's FP64 relax (one iteration from x = 0) equals the float64 brute
    force bit for bit, and differs from the FP32 relax of the same data."""
    from test_oracle_pins_multires import level_field, set_level_field, dyadic
    n = nbe * bsx0
    e = bsx0 >> lev
    rng = np.random.default_rng(90 + bsx0)
    w = [rng.uniform(0.5, 1.5, (n, n, n)).astype(F32) for _ in range(3)]
    rhs = rng.standard_normal((n, n, n)).astype(F32)
    o = Oracle(n, n, n, bsx0, 2)
    o.set_domain_boundary_spec(*([msbgx.DBC_OPEN] * 6))
    o.set_refinement_map(np.full(o.n_blocks, lev, np.int32))
    o.prepare_dense((n, n, n), np.zeros((n, n, n), np.uint16), *w, rhs)
    nb = (nbe, nbe, nbe)
    relaxed = list(o.relax_list(1))
    assert len(relaxed) == 8    # the blocks whose doubled-coordinate samples stay inside the raster
    D = level_field(o, nb, e, CH_DIAG, 1)
    W = [level_field(o, nb, e, k, 1) for k in (7, 8, 9)]
    b0 = rng.standard_normal((nbe * e,) * 3).astype(F32)
    set_level_field(o, b0, nb, e, CH_R, 1)
    o.relax_coarse64(1, 1, CH_Z, CH_R)
    got = level_field(o, nb, e, CH_Z, 1)
    exp = brute_coarse64(b0, W, D, o.omega(), nb, e, 2.0 ** lev, relaxed, seg, 1, 1.0).astype(F32)
    assert np.array_equal(got, exp), f"{np.count_nonzero(got != exp)} cells differ"
    assert np.count_nonzero(got) > 0
    set_level_field(o, np.zeros_like(b0), nb, e, CH_Z, 1)
    o.relax(1, False, 1, CH_Z, CH_R)
    got32 = level_field(o, nb, e, CH_Z, 1)
    assert not np.array_equal(got32, got)
