"""Host-side parts of the benchmark and the warm-start pipeline (no GPU)."""
import importlib.util
import os
import subprocess
import sys

from conftest import ROOT


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_of_the_67_cube_system():
    """B1 = (nnz_k + nnz_d)(s+4) + 4(m_k+1) + 4(m_d+1) + s n + 2 s m and
    B2 = (nnz_k + nnz_d)(s+4) + 8(n+1) + s m + 2 s n at the measured sizes of
    the 67^3 / 2M-ray system (FP64: 2.352 / 2.319 GB; FP32: 1.563 / 1.542 GB)."""
    B = load_bench()
    m_k, m_d, n, nnz_k, nnz_d = 2_000_000, 1_176_118, 300_763, 187_571_931, 2_926_828
    m = m_k + m_d
    b1, b2 = B.aprod_bytes(m_k, m_d, n, nnz_k, nnz_d)
    assert b1 == (nnz_k + nnz_d) * 12 + 4 * (m_k + 1) + 4 * (m_d + 1) + 8 * n + 16 * m
    assert b2 == (nnz_k + nnz_d) * 12 + 8 * (n + 1) + 8 * m + 16 * n
    assert round(b1 / 1e9, 3) == 2.352 and round(b2 / 1e9, 3) == 2.319
    f1, f2 = B.aprod_bytes(m_k, m_d, n, nnz_k, nnz_d, 4, 4, 4)
    assert round(f1 / 1e9, 3) == 1.563 and round(f2 / 1e9, 3) == 1.542


def test_warmstart_summary(P):
    import importlib
    W = importlib.import_module(P.__name__ + ".warmstart")
    rows = []
    for j in range(4):
        rows.append({
            "fp64_zero": dict(itn=100 + j, ms=10.0 + j),
            "fp64_warm": dict(itn=80 + j, ms=8.0 + j, used_x0=True, xerr=1e-3 * (j + 1)),
            "mixed_zero": dict(itn=110 + j, ms=9.0, xerr=2e-3, test2=1e-7),
            "mixed_warm": dict(itn=90 + j, ms=7.0, used_x0=j != 3, xerr=3e-3, test2=1e-7),
            "x0_err": 0.1 * (j + 1),
        })
    S = W.summarize(rows)
    assert set(k for k in S if isinstance(S[k], dict)) == {"fp64_zero", "fp64_warm", "mixed_zero", "mixed_warm"}
    assert S["fp64_zero"]["itn_mean"] == 101.5 and S["fp64_zero"]["ms_mean"] == 11.5
    assert S["fp64_warm"]["xerr_max"] == 4e-3 and S["fp64_warm"]["used_x0"] == 4
    assert S["mixed_warm"]["used_x0"] == 3
    assert abs(S["x0_err_median"] - 0.25) < 1e-12
    assert abs(S["mixed_warm"]["speedup_vs_fp64_zero"] - 11.5 / 7.0) < 1e-12
