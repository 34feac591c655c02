"""CPU pins of the FP64 network oracle (oracle/nn_oracle.cpp) that checks the
GPU kernels of the learned preconditioner correction.

* The gradient equals central finite differences of the loss (pure FP64).
* In the linear regime (positive inputs and weights, every ReLU active) the
  forward equals an independent composition of scipy.ndimage.correlate
  calls with zero padding on the 18^3 patch -- this pins the tap order, the
  "a tap reads p + (dx, dy, dz)" orientation and the cell window.
* Identity weights give ReLU(x) on the cells, and BF16 rounding is
  round-to-nearest-even (ties to even) on hand-picked values.
"""
import numpy as np
import pytest
from scipy import ndimage

import nn_oracle_lib as nno

T, C = 27, 16
OFF_W1, OFF_W2 = 0, T * C
OFF_W3 = OFF_W2 + T * C * C
OFF_W4 = OFF_W3 + T * C * C


def random_weights(rng, scale=1.0):
    w = np.zeros(nno.NPARAM)
    w[OFF_W1:OFF_W2] = rng.normal(0, np.sqrt(2 / 27), T * C)
    w[OFF_W2:OFF_W3] = rng.normal(0, np.sqrt(2 / 432), T * C * C)
    w[OFF_W3:OFF_W4] = rng.normal(0, np.sqrt(2 / 432), T * C * C)
    w[OFF_W4:] = rng.normal(0, np.sqrt(1 / 432), T * C)
    return w * scale


def samples(rng, n):
    x = rng.normal(0, 1, (n, 18 ** 3)).astype(np.float32)
    t = rng.normal(0, 0.3, (n, 4096)).astype(np.float32)
    m = (rng.random((n, 4096)) < 0.8).astype(np.uint8)
    return x, t, m


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = random_weights(rng)
    x, t, m = samples(rng, 2)
    _, g = nno.loss_grad(w, x, t, m, round_bf16=False)
    picks = np.concatenate([rng.choice(np.arange(a, b), 5, replace=False)
                            for a, b in ((OFF_W1, OFF_W2), (OFF_W2, OFF_W3), (OFF_W3, OFF_W4), (OFF_W4, nno.NPARAM))])
    # the loss is piecewise quadratic (ReLU kinks): a small step keeps almost
    # every perturbation on one piece, where central differences are exact
    for k in picks:
        h = 1e-7 * max(1.0, abs(w[k]))
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        lp, _ = nno.loss_grad(wp, x, t, m, round_bf16=False)
        lm, _ = nno.loss_grad(wm, x, t, m, round_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * abs(g[k]) + 1e-8, (k, fd, g[k])


def corr_layer(a, wl, cin, cout):
    """a: [cin][18][18][18] (z, y, x); wl: [27][cin][cout]; zero padding."""
    out = np.zeros((cout, 18, 18, 18))
    for co in range(cout):
        for ci in range(cin):
            k = wl[:, ci, co].reshape(3, 3, 3)       # [dz][dy][dx]
            out[co] += ndimage.correlate(a[ci], k, mode="constant", cval=0.0)
    return out


def test_linear_regime_equals_scipy_correlation():
    rng = np.random.default_rng(5)
    w = np.abs(random_weights(rng))
    x = rng.random((1, 18 ** 3)).astype(np.float32) + 0.1
    m = np.ones((1, 4096), np.uint8)
    out = nno.forward(w, x, m, round_bf16=False)[0]
    a = x[0].astype(np.float64).reshape(1, 18, 18, 18)
    a = corr_layer(a, w[OFF_W1:OFF_W2].reshape(T, 1, C), 1, C)
    a = corr_layer(a, w[OFF_W2:OFF_W3].reshape(T, C, C), C, C)
    a = corr_layer(a, w[OFF_W3:OFF_W4].reshape(T, C, C), C, C)
    a = corr_layer(a, w[OFF_W4:].reshape(T, C, 1), C, 1)[0]
    ref = a[1:17, 1:17, 1:17].reshape(-1)            # cells, x fastest
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=0)


def identity_weights():
    w = np.zeros(nno.NPARAM)
    c = 13
    w[OFF_W1 + c * C + 0] = 1.0
    w[OFF_W2 + (c * C + 0) * C + 0] = 1.0
    w[OFF_W3 + (c * C + 0) * C + 0] = 1.0
    w[OFF_W4 + c * C + 0] = 1.0
    return w


def test_identity_network_is_relu_on_cells():
    rng = np.random.default_rng(9)
    x = rng.normal(0, 1, (2, 18 ** 3)).astype(np.float32)
    m = np.ones((2, 4096), np.uint8)
    m[1, ::7] = 0
    out = nno.forward(identity_weights(), x, m, round_bf16=False)
    cells = x.reshape(2, 18, 18, 18)[:, 1:17, 1:17, 1:17].reshape(2, -1)
    np.testing.assert_array_equal(out, np.maximum(cells, 0) * m)


@pytest.mark.parametrize("val,expect", [(1.0 + 2.0 ** -8, 1.0), (1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -6),
                                        (1.0 + 2.0 ** -8 + 2.0 ** -12, 1.0 + 2.0 ** -7), (-3.0, 0.0)])
def test_bf16_rounding_is_nearest_even(val, expect):
    x = np.zeros((1, 18 ** 3), np.float32)
    x[0, 1 + 18 + 324] = val                        # cell (0, 0, 0)
    m = np.ones((1, 4096), np.uint8)
    out = nno.forward(identity_weights(), x, m, round_bf16=True)
    assert out[0, 0] == expect
