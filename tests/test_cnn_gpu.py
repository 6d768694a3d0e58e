"""Warm-start network on the GPU against the CPU oracle (oracle/cnn_oracle.py),
and the warm-started LSQR it feeds.

Bars: one layer (bf16 operands, FP32 sums) within 1e-5 of the oracle's exact
sums of the same bf16 operands, relative to the output's norm; the whole
network's loss within 1e-4 and each layer's gradient within 2e-2 (the bf16
rounding of every intermediate activation may round the other way after a
different FP32 summation order).
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cnn_oracle as K  # noqa: E402


def announce(msg):
    print(msg, flush=True)


def test_cnn_rejects_bad_arguments_on_the_host(P):
    with pytest.raises(P.PlsqrError) as e:
        P.WarmStartNet((0, 4, 4))
    assert e.value.code == 1
    with pytest.raises(P.PlsqrError) as e:
        P.conv_forward((0, 3, 3), np.zeros((1, 0, 32), np.float32), np.zeros((27, 32, 32), np.float32))
    assert e.value.code == 1


@pytest.mark.gpu
@pytest.mark.parametrize("relu", [False, True])
def test_conv_forward_matches_oracle(P, gpu_available, relu):
    announce(f"conv forward relu={relu}")
    grid = (9, 7, 5)
    V = 9 * 7 * 5
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, V, 32)).astype(np.float32)
    w = (rng.standard_normal((27, 32, 32)) * 0.1).astype(np.float32)
    b = rng.standard_normal(32).astype(np.float32)
    got = P.conv_forward(grid, x, w, b, relu=relu)
    want = K.conv_forward(grid, K.bf16(x), K.bf16(w), b, relu=relu)
    assert rel(got, want) < 1e-5, rel(got, want)
    assert np.abs(got - want).max() < 1e-4 * np.abs(want).max()


@pytest.mark.gpu
def test_conv_backward_matches_oracle(P, gpu_available):
    announce("conv backward")
    grid = (6, 5, 7)
    V = 6 * 5 * 7
    rng = np.random.default_rng(6)
    x = rng.standard_normal((3, V, 32)).astype(np.float32)
    w = (rng.standard_normal((27, 32, 32)) * 0.1).astype(np.float32)
    d = rng.standard_normal((3, V, 32)).astype(np.float32)
    din, dw, db = P.conv_backward(grid, x, w, d)
    rdin, rdw, rdb = K.conv_backward(grid, K.bf16(x), K.bf16(w), K.bf16(d))
    assert rel(din, rdin) < 1e-5, rel(din, rdin)
    assert rel(dw, rdw) < 1e-5, rel(dw, rdw)
    assert rel(db, rdb) < 1e-5, rel(db, rdb)


@pytest.mark.gpu
def test_network_gradient_matches_oracle(P, gpu_available):
    announce("network gradient")
    grid = (8, 6, 5)
    V = 8 * 6 * 5
    rng = np.random.default_rng(7)
    atb = rng.standard_normal((2, V)) * 3.0
    xs = rng.standard_normal((2, V))
    with P.WarmStartNet(grid, max_batch=2, seed=11) as net:
        params = net.get_params()
        loss, g = net.gradient(atb, xs)
        x0 = net.predict(atb)
    ref = K.network(grid, params, atb, xs)
    assert abs(loss - ref["loss"]) < 1e-4 * ref["loss"], (loss, ref["loss"])
    lsz = 27 * 32 * 32 + 32
    for l in range(4):
        a, b = g[l * lsz:(l + 1) * lsz], ref["grads"][l * lsz:(l + 1) * lsz]
        assert rel(a, b) < 2e-2, (l, rel(a, b))
    # prediction = scale * gain * output channel 0
    s = np.sqrt(np.mean(atb * atb, axis=1))[:, None]
    assert params[-1] == 1.0 and g[-1] == 0.0
    assert rel(x0, s * ref["z"]) < 1e-3
    # padded parameters are zero
    assert np.all(params[:27 * 32 * 32].reshape(27, 32, 32)[:, :, 1:] == 0)
    assert np.all(params[3 * lsz:3 * lsz + 27 * 32 * 32].reshape(27, 32, 32)[:, 1:, :] == 0)


def make_rhs(P, s, grid, seed, sigma, op):
    """b = A_k x_true + noise for the generator's blob field of `seed`."""
    xt = P.gen_field(*grid, seed)
    y = np.zeros(s.m_k + s.m_d)
    op.aprod(1, xt.copy(), y)
    b = y[:s.m_k] + sigma * np.random.default_rng(seed).standard_normal(s.m_k)
    return b, xt


def atb_of(op, s, b):
    x = np.zeros(s.n)
    op.aprod(2, x, np.concatenate([b, np.zeros(s.m_d)]))
    return x


@pytest.mark.gpu
def test_warm_start_falls_back_to_zero_start(P, gpu_available):
    announce("warm start fallback")
    s = P.generate(16, 16, 16, 20000, 0, 2013, 0.01, 0.1)
    with P.Operator(s) as op:
        x_zero, inf_zero = op.lsqr(s.b, 200, tol=1e-6)
        bad = np.full(s.n, 1e3)
        x_bad, inf_bad = op.lsqr_x0(s.b, bad, 200, tol=1e-6)
        x_none, inf_none = op.lsqr_x0(s.b, None, 200, tol=1e-6)
        # x0 = the converged solution: kept, and LSQR stops at once
        x_exact, inf_exact = op.lsqr_x0(s.b, x_zero, 200, tol=1e-6)
        xm, infm = op.lsqr_x0(s.b, x_zero * 0.9, 2000, tol=1e-6, precision=P.PREC_MIXED, refine_every=50)
    assert not inf_bad["used_x0"] and np.array_equal(x_bad, x_zero) and inf_bad["itn"] == inf_zero["itn"]
    assert not inf_none["used_x0"] and np.array_equal(x_none, x_zero)
    # (LSQR's anorm estimate starts small, so test2 needs a few iterations to settle)
    assert inf_exact["used_x0"] and inf_exact["itn"] <= 12 and rel(x_exact, x_zero) < 1e-6
    assert infm["used_x0"] and infm["arnorm"] <= 1e-6 and rel(xm, x_zero) < 1e-3
