"""CPU pins of the warm-start network's oracle (oracle/surrogate_oracle.py):
the convolution against scipy's correlate, every backward function and the
whole encoder-decoder against central finite differences in float64, the
pooling/upsampling pair against closed forms, bf16 rounding against known
bit patterns, and Adam's first step against its closed form."""
import os
import sys

import numpy as np
import pytest
from scipy.signal import correlate

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import surrogate_oracle as S  # noqa: E402


def test_bf16_round_known_values():
    assert S.bf16_bits(1.0) == 0x3F80
    assert S.bf16_bits(-2.0) == 0xC000
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1.0).
    assert S.bf16_bits(1.0 + 2.0 ** -8) == 0x3F80
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: to even (1 + 2^-6).
    assert S.bf16_bits(1.0 + 3 * 2.0 ** -8) == 0x3F82
    assert S.bf16_round(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    x = np.random.default_rng(0).normal(size=1000)
    assert np.all(np.abs(S.bf16_round(x) - x) <= np.abs(x) * 2.0 ** -8)
    assert np.array_equal(S.bits_to_f64(S.bf16_bits(x)), S.bf16_round(x))


def test_conv_fwd_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    B, H, W, Ci, Co = 2, 5, 7, 3, 4
    x = rng.normal(size=(B, H, W, Ci))
    w = rng.normal(size=(Co, 3, 3, Ci))
    b = rng.normal(size=Co)
    z = S.conv_fwd(x, w, b)
    for bb in range(B):
        for co in range(Co):
            ref = b[co] + sum(correlate(x[bb, :, :, ci], w[co, :, :, ci], mode="same") for ci in range(Ci))
            np.testing.assert_allclose(z[bb, :, :, co], ref, rtol=1e-12, atol=1e-12)


def _fd(f, x, eps=1e-6):
    g = np.zeros_like(x)
    it = np.nditer(x, flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        old = x[i]
        x[i] = old + eps
        fp = f()
        x[i] = old - eps
        fm = f()
        x[i] = old
        g[i] = (fp - fm) / (2 * eps)
    return g


def test_conv_backward_finite_differences():
    rng = np.random.default_rng(2)
    B, H, W, Ci, Co = 2, 4, 3, 2, 3
    x = rng.normal(size=(B, H, W, Ci))
    w = rng.normal(size=(Co, 3, 3, Ci))
    g = rng.normal(size=(B, H, W, Co))
    loss = lambda: float(np.sum(S.conv_fwd(x, w) * g))
    np.testing.assert_allclose(S.conv_bwd_data(g, w, x.shape), _fd(loss, x), rtol=1e-7, atol=1e-8)
    np.testing.assert_allclose(S.conv_bwd_weight(x, g), _fd(loss, w), rtol=1e-7, atol=1e-8)


def test_pool_up_silu_closed_forms():
    x = np.arange(2 * 4 * 6 * 3, dtype=np.float64).reshape(2, 4, 6, 3)
    p = S.pool_fwd(x)
    assert p.shape == (2, 2, 3, 3)
    assert p[1, 1, 2, 1] == 0.25 * (x[1, 2, 4, 1] + x[1, 2, 5, 1] + x[1, 3, 4, 1] + x[1, 3, 5, 1])
    u = S.up_fwd(p)
    assert u.shape == x.shape and u[1, 3, 5, 2] == p[1, 1, 2, 2]
    rng = np.random.default_rng(3)
    d = rng.normal(size=p.shape)
    # <pool(x), d> = <x, pool_bwd(d)> and <up(p), e> = <p, up_bwd(e)> (adjoints).
    assert np.isclose(np.sum(S.pool_fwd(x) * d), np.sum(x * S.pool_bwd(d)))
    e = rng.normal(size=x.shape)
    assert np.isclose(np.sum(S.up_fwd(p) * e), np.sum(p * S.up_bwd(e)))
    z = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(S.silu(z), z * 0.5 * (1 + np.tanh(z / 2)), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(S.silu_grad(z), (S.silu(z + 1e-6) - S.silu(z - 1e-6)) / 2e-6, rtol=1e-7, atol=1e-9)


def test_adam_first_step_closed_form():
    w, g = np.array([1.0, -2.0]), np.array([0.5, -3.0])
    w1, m, v = S.adam(w, g, np.zeros(2), np.zeros(2), 0.01, 0.9, 0.999, 1e-8, 1)
    # First bias-corrected step moves each weight by lr * sign(g) (up to eps).
    np.testing.assert_allclose(w1, w - 0.01 * np.sign(g), rtol=1e-6)


def test_unet_gradients_finite_differences():
    rng = np.random.default_rng(4)
    levels, base, cin, cout = 3, 2, 3, 2
    layers = S.unet_layers(cin, cout, base, levels)
    params = [(rng.normal(0, 0.5, size=(co, 3, 3, ci)), rng.normal(0, 0.1, size=co)) for ci, co in layers]
    x = rng.normal(size=(2, 8, 4, cin))
    t = rng.normal(size=(2, 8, 4, cout))

    def loss():
        y, _ = S.unet_forward(x, params, levels)
        return S.mse(y, t)[0]

    y, tape = S.unet_forward(x, params, levels)
    _, dy = S.mse(y, t)
    grads = S.unet_backward(dy, tape, levels)
    for k in (0, 3, len(layers) // 2, len(layers) - 2, len(layers) - 1):
        w, b = params[k]
        for got, ref in ((grads[k][0], _fd(loss, w)), (grads[k][1], _fd(loss, b))):
            # Central differences carry O(eps^2) truncation and cancellation
            # error: compare in norm.
            assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref), (k, np.linalg.norm(got - ref))
