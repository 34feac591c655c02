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
