"""plsqr_cnn_conv_forward_noalloc: the allocation-free convolution layer.

Called through the C-ABI on device buffers (torch tensors, on torch's current
stream) and compared with plsqr_cnn_conv_forward, the entry point it is a
variant of, on the same inputs: the same kernels run on the same bf16
operands, so the outputs are bitwise equal.  Two calls in a row, the second
on a smaller grid and batch, so the fixed scratch holds the first call's
larger contents when the second runs.  Sizes beyond the scratch are refused
before any device work.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_conv_forward_noalloc_matches_the_allocating_layer(P, gpu_available):
    import torch
    lib = P.load()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    rng = np.random.default_rng(4)
    w = (rng.standard_normal((27, 32, 32)) * 0.1).astype(np.float32)
    bias = (rng.standard_normal(32) * 0.1).astype(np.float32)
    w_d = torch.from_numpy(w).to(dev)
    b_d = torch.from_numpy(bias).to(dev)
    for (nx, ny, nz, batch, relu) in ((12, 10, 9, 3, 1), (6, 5, 4, 1, 0)):
        print(f"noalloc conv forward: grid {nx}x{ny}x{nz}, batch {batch}, relu {relu}", flush=True)
        x = rng.standard_normal((batch, nx * ny * nz, 32)).astype(np.float32)
        want = P.conv_forward((nx, ny, nz), x, w, bias, relu=bool(relu))
        x_d = torch.from_numpy(x).to(dev)
        out_d = torch.full((batch, nx * ny * nz, 32), float("nan"), dtype=torch.float32, device=dev)
        rc = lib.plsqr_cnn_conv_forward_noalloc(nx, ny, nz, batch, C.c_void_p(x_d.data_ptr()),
                                                C.c_void_p(w_d.data_ptr()), C.c_void_p(b_d.data_ptr()), relu,
                                                C.c_void_p(out_d.data_ptr()), C.c_void_p(stream))
        assert rc == 0, lib.plsqr_cnn_last_error(None)
        torch.cuda.synchronize()
        got = out_d.cpu().numpy()
        assert np.array_equal(got, want), float(np.nanmax(np.abs(got - want)))


def test_conv_forward_noalloc_refuses_what_its_scratch_cannot_hold(P, gpu_available):
    import torch
    lib = P.load()
    small = torch.zeros(32 * 27 * 32, dtype=torch.float32, device="cuda:0")
    p = C.c_void_p(small.data_ptr())
    # 64^3 voxels = 262,144 > 131,072: refused on the host before any launch
    rc = lib.plsqr_cnn_conv_forward_noalloc(64, 64, 64, 1, p, p, None, 0, p, None)
    assert rc == 1
    assert b"allocation-free limit" in lib.plsqr_cnn_last_error(None)
