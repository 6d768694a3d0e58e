"""CPU oracle of the warm-start network (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU baseline use this file,
as the checker of the GPU kernels in plsqr3.2013.09_amd/csrc/plsqr_cnn.h; the
product never calls it.

The network has no counterpart in the reference (PLSQR3 always starts LSQR
from x = 0, lsqr.c:311-312); what this file restates is the library's own
definition of it:

  * a layer is a 3x3x3 convolution with zero padding on the nx*ny*nz grid,
    32 -> 32 channels, channels-last rows g = sample*V + voxel:
        out[g][n] = bias[n] + sum_t sum_k in[g + off(t)][k] * w[t][n][k],
    tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) reading voxel (x+dx, y+dy, z+dz);
  * the network is 4 such layers with ReLU between them; the input holds
    A^T b / s in channel 0 (s = RMS of A^T b), the output's channel 0 is
    x0 / s;
  * weights and activations are rounded to bf16 (round to nearest even)
    where the GPU rounds them, sums are exact here (float64) and FP32 there.

Pinned by: the finite-difference check of conv_backward against
conv_forward, scipy.ndimage.correlate for the tap orientation and zero
padding, and hand-computed bf16 roundings (tests/test_cnn_oracle.py).
"""
import numpy as np

C = 32
TAPS = 27


def bf16(a):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def taps():
    """(t, dz, dy, dx) for the 27 taps."""
    return [((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1), dz, dy, dx)
            for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _shift(a, grid, dz, dy, dx):
    """s[g] = a[g + off] (zero outside the grid); a is (B, V, ch)."""
    nx, ny, nz = grid
    B = a.shape[0]
    a5 = a.reshape(B, nz, ny, nx, -1)
    p = np.zeros((B, nz + 2, ny + 2, nx + 2, a5.shape[-1]), a.dtype)
    p[:, 1:-1, 1:-1, 1:-1] = a5
    s = p[:, 1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
    return s.reshape(a.shape)


def conv_forward(grid, x, w, bias=None, relu=False):
    """x (B, V, 32), w (27, 32, 32) as [t][n][k]; float64 sums."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    out = np.zeros(x.shape[:2] + (w.shape[1],), np.float64)
    for t, dz, dy, dx in taps():
        out += _shift(x, grid, dz, dy, dx) @ w[t].T
    if bias is not None:
        out += np.asarray(bias, np.float64)
    if relu:
        out = np.maximum(out, 0.0)
    return out


def conv_backward(grid, x, w, dout):
    """(din, dw, db) of out = conv_forward(x, w, bias) for the output gradient dout."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    dout = np.asarray(dout, np.float64)
    din = np.zeros(x.shape, np.float64)
    dw = np.zeros(w.shape, np.float64)
    B = x.shape[0]
    for t, dz, dy, dx in taps():
        xs = _shift(x, grid, dz, dy, dx)
        dw[t] = dout.reshape(-1, dout.shape[-1]).T @ xs.reshape(-1, x.shape[-1])
        din += _shift(dout, grid, -dz, -dy, -dx) @ w[t]
    db = dout.reshape(-1, dout.shape[-1]).sum(axis=0)
    del B
    return din, dw, db


def gain_of(params):
    """The output gain (last parameter; 1 if the vector has none)."""
    p = np.asarray(params)
    return float(p[-1]) if p.size == 4 * (TAPS * C * C + C) + 1 else 1.0


def unpack(params):
    """Library parameter layout -> [(w (27,32,32), bias (32,))] * 4 (views)."""
    p = np.asarray(params)
    lsz = TAPS * C * C + C
    assert p.size in (4 * lsz, 4 * lsz + 1)
    return [(p[l * lsz:l * lsz + TAPS * C * C].reshape(TAPS, C, C), p[l * lsz + TAPS * C * C:(l + 1) * lsz])
            for l in range(4)]


def normalise(atb, xstar=None, gain=1.0):
    """The per-sample scaling of the library: returns (in, t, tn, s) with
    in = atb / s, t = xstar / (s * gain), s = RMS(atb)."""
    atb = np.asarray(atb, np.float64)
    s = np.sqrt(np.mean(atb * atb))
    if not (s > 0) or not np.isfinite(s):
        s = 1.0
    inp = (atb / s).astype(np.float32)
    if xstar is None:
        return inp, None, None, s
    q = np.asarray(xstar, np.float64) / s / gain
    tn = float(np.float32(max(float(np.sum(q * q)), 0.0) or 1.0))
    return inp, q.astype(np.float32), tn, s


def network(grid, params, atb, xstar=None, exact=False):
    """Forward (and, with xstar, loss and parameter gradients) of the network
    with the GPU's bf16 rounding points (exact=True: every value float64, no
    rounding, for derivative checks).  atb, xstar: (B, V).
    Returns dict(z=output channel 0 (B, V), loss, grads (layout of params,
    zero for the gain)); the prediction is x0 = s * gain * z."""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else bf16
    f32 = np.float64 if exact else np.float32
    atb = np.atleast_2d(np.asarray(atb, np.float64))
    B, V = atb.shape
    layers = unpack(params)
    gain = gain_of(params)
    ins, tgts, tns = [], [], []
    for b in range(B):
        i, t, tn, sc = normalise(atb[b], None if xstar is None else np.atleast_2d(xstar)[b], gain)
        if exact:
            i = atb[b] / sc
            if xstar is not None:
                t = np.atleast_2d(xstar)[b] / sc / gain
                tn = float(np.sum(t * t)) or 1.0
        ins.append(i)
        tgts.append(t)
        tns.append(tn)
    a = np.zeros((B, V, C), f32)
    a[:, :, 0] = rnd(np.stack(ins))
    acts = [a]
    wq = [rnd(w) for w, _ in layers]
    z = None
    for l, (w, bias) in enumerate(layers):
        zl = conv_forward(grid, acts[l], wq[l], np.asarray(bias, np.float64))
        if l < 3:
            acts.append(rnd(np.maximum(zl, 0.0).astype(f32)))
        else:
            z = zl.astype(f32)
    res = {"z": z[:, :, 0].astype(np.float64)}
    if xstar is None:
        return res
    tg = np.stack(tgts).astype(np.float64)
    tn = np.array(tns, np.float64)
    d = z[:, :, 0].astype(np.float64) - tg
    res["loss"] = float(np.mean(np.sum(d * d, axis=1) / tn))
    dz = np.zeros((B, V, C), f32)
    dz[:, :, 0] = rnd((2.0 * d / (B * tn[:, None])).astype(f32))
    grads = []
    for l in (3, 2, 1, 0):
        din, dw, db = conv_backward(grid, acts[l], wq[l], dz)
        grads.append((l, dw, db))
        if l > 0:
            dz = rnd(np.where(acts[l] > 0, din, 0.0).astype(f32))
    lsz = TAPS * C * C + C
    out = np.zeros(np.asarray(params).size, np.float64)
    for l, dw, db in grads:
        out[l * lsz:l * lsz + TAPS * C * C] = dw.ravel()
        out[l * lsz + TAPS * C * C:(l + 1) * lsz] = db
    res["grads"] = out
    return res
