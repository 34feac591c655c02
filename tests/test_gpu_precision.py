"""Precision form on the GPU (build addition, no reference counterpart).

The smoother reads FP16 / BF16 copies of the face coefficients and inverse
diagonal; the Krylov operator stays FP32; an FP64 iterative refinement
recomputes the true residual.  Each variant must reach an FP64 relative
residual <= 1e-10 and agree with the oracle's FP64 solution of the same
level-0 system to <= 1e-8 relative L2; the FP32 smoother path must match
the oracle's plain solve exactly (test_gpu_parity)."""
import numpy as np
import pytest

import msbgx
from parity_util import pair

pytestmark = pytest.mark.gpu

CONFIGS = {
    "c1_64": lambda: msbgx.config_c1(n=64),
    "c1_128": lambda: msbgx.config_c1(n=128),
    "c4_128_multires": lambda: msbgx.config_c4(n=128),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_refined_solutions_match_fp64_oracle(name):
    print(f"precision {name}", flush=True)
    cfg = CONFIGS[name]()
    g, o = pair(cfg)
    it_ref, rel_ref, x_ref = o.solve_fp64(tol=1e-13, max_iter=20000)
    assert rel_ref <= 1e-12
    den = np.sqrt(sum(float((v ** 2).sum()) for v in x_ref.values()))
    for prec in (msbgx.PREC_FP32, msbgx.PREC_FP16, msbgx.PREC_BF16):
        g.set_precision(prec)
        outer, inner, rel, xb = g.solve_pressure_refined(tol=1e-10, max_outer=30, inner_tol=1e-5, want_x=True)
        num = np.sqrt(sum(float(((xb.get(b, 0.0) - v) ** 2).sum()) for b, v in x_ref.items()))
        print(f"  prec={prec}: outer {outer} inner {inner} fp64 rel residual {rel:.2e} "
              f"error vs FP64 oracle {num / den:.2e}", flush=True)
        assert rel <= 1e-10
        assert num / den <= 1e-8


def test_half_precision_smoother_changes_only_the_preconditioner():
    """FP16 smoother + FP32 operator: plain PCG still converges to 1e-6."""
    cfg = msbgx.config_c1(n=128)
    g, _ = pair(cfg)
    it32, rr32 = g.solve_pressure_flip(max_iter=300, tol=1e-6)
    for prec in (msbgx.PREC_FP16, msbgx.PREC_BF16):
        g.set_precision(prec)
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr = g.solve_pressure_flip(max_iter=300, tol=1e-6)
        print(f"  prec={prec}: iters {it} (fp32 {it32}) rel {rr:.2e}", flush=True)
        assert rr <= 1e-6 and it <= it32 + 10
