"""Pins the CPU oracle of the warm-start network before it checks the GPU kernels."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cnn_oracle as K  # noqa: E402


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)              # bf16 spacing at 1.0
    cases = [(one, one), (one + ulp / 2, one),  # tie -> even (1.0)
             (one + 3 * ulp / 2, one + 2 * ulp),  # tie -> even (1 + 2 ulp)
             (one + ulp * 0.75, one + ulp), (np.float32(-2.5), np.float32(-2.5)),
             (np.float32(3.0e38), np.float32(2.9983e38))]
    for x, want in cases:
        got = K.bf16(np.array([x], np.float32))[0]
        assert got == np.float32(want) or abs(got - want) <= abs(want) * 2 ** -8, (x, got, want)
    assert np.isnan(K.bf16(np.array([np.nan], np.float32))[0])
    r = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    q = K.bf16(r)
    assert np.all((q.view(np.uint32) & 0xFFFF) == 0)
    assert np.all(np.abs(q - r) <= np.abs(r) * 2.0 ** -8)


def test_single_channel_tap_convention_matches_scipy():
    """out[v] = sum_t x[v + off(t)] w[t] is a correlation with zero padding."""
    from scipy import ndimage
    grid = (5, 4, 3)
    nx, ny, nz = grid
    rng = np.random.default_rng(1)
    x = np.zeros((1, nx * ny * nz, K.C))
    x[0, :, 0] = rng.standard_normal(nx * ny * nz)
    w = np.zeros((K.TAPS, K.C, K.C))
    kern = rng.standard_normal((3, 3, 3))      # indexed [dz+1, dy+1, dx+1]
    for t, dz, dy, dx in K.taps():
        w[t, 0, 0] = kern[dz + 1, dy + 1, dx + 1]
    out = K.conv_forward(grid, x, w)[0, :, 0].reshape(nz, ny, nx)
    ref = ndimage.correlate(x[0, :, 0].reshape(nz, ny, nx), kern, mode="constant", cval=0.0)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_conv_backward_matches_finite_differences():
    grid = (4, 3, 3)
    V = 36
    rng = np.random.default_rng(2)
    B, Cs = 2, K.C
    x = rng.standard_normal((B, V, Cs))
    w = rng.standard_normal((K.TAPS, Cs, Cs)) * 0.1
    bias = rng.standard_normal(Cs)
    R = rng.standard_normal((B, V, Cs))

    def f(x_, w_, b_):
        return float(np.sum(K.conv_forward(grid, x_, w_, b_) * R))

    din, dw, db = K.conv_backward(grid, x, w, R)
    h = 1e-6
    for _ in range(12):
        i = tuple(rng.integers(0, s) for s in x.shape)
        e = np.zeros_like(x)
        e[i] = h
        assert abs((f(x + e, w, bias) - f(x - e, w, bias)) / (2 * h) - din[i]) < 1e-6
        j = tuple(rng.integers(0, s) for s in w.shape)
        e = np.zeros_like(w)
        e[j] = h
        assert abs((f(x, w + e, bias) - f(x, w - e, bias)) / (2 * h) - dw[j]) < 1e-6
        k = rng.integers(0, Cs)
        e = np.zeros_like(bias)
        e[k] = h
        assert abs((f(x, w, bias + e) - f(x, w, bias - e)) / (2 * h) - db[k]) < 1e-6


def test_network_gradient_matches_finite_differences_without_rounding():
    """The network's backward pass (ReLU masks, layer chaining, loss scaling)
    in float64 without the bf16 rounding, against central differences."""
    grid = (3, 3, 2)
    V = 18
    rng = np.random.default_rng(3)
    lsz = K.TAPS * K.C * K.C + K.C
    p = rng.standard_normal(4 * lsz) * 0.2
    for l, (w, b) in enumerate(K.unpack(p)):   # padded entries are zero, as in the library
        if l == 0:
            w[:, :, 1:] = 0
        if l == 3:
            w[:, 1:, :] = 0
            b[1:] = 0
    atb = rng.standard_normal((2, V))
    xs = rng.standard_normal((2, V))
    g = K.network(grid, p, atb, xs, exact=True)["grads"]
    h = 1e-6
    for _ in range(16):
        l = int(rng.integers(0, 4))
        t, n, k = (int(v) for v in rng.integers(0, [K.TAPS, K.C, K.C]))
        if l == 0:
            k = 0
        if l == 3:
            n = 0
        for idx in (l * lsz + (t * K.C + n) * K.C + k, l * lsz + K.TAPS * K.C * K.C + n):
            pp, pm = p.copy(), p.copy()
            pp[idx] += h
            pm[idx] -= h
            fd = (K.network(grid, pp, atb, xs, exact=True)["loss"]
                  - K.network(grid, pm, atb, xs, exact=True)["loss"]) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(g[idx])), (l, idx, fd, g[idx])


@pytest.mark.parametrize("relu", [False, True])
def test_conv_forward_is_linear_in_the_weights(relu):
    grid = (3, 4, 2)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 24, K.C))
    w1, w2 = rng.standard_normal((2, K.TAPS, K.C, K.C))
    a = K.conv_forward(grid, x, w1 + w2)
    b = K.conv_forward(grid, x, w1) + K.conv_forward(grid, x, w2)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-10)
    if relu:
        assert np.all(K.conv_forward(grid, x, w1, relu=True) >= 0)
