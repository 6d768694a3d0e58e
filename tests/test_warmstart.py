"""What the trained warm-start network gains: the full pipeline (pairs,
FP64 targets, GPU training, warm-started LSQR) on the 16^3 / 20,000-ray
system, against the same LSQR started from zero."""
import numpy as np
import pytest


@pytest.mark.gpu
def test_trained_network_reduces_lsqr_iterations(P, gpu_available):
    print("warm start pipeline 16^3", flush=True)
    import importlib
    W = importlib.import_module(P.__name__ + ".warmstart")
    s = P.generate(16, 16, 16, 20000, 0, 2013, 0.01, 0.1)
    res, net = W.run(s, (16, 16, 16), n_train=96, n_eval=16, sigma=0.01, seed0=1000, target_tol=1e-8,
                     epochs=30, tol=1e-6, refine_every=25, log=lambda m: print(m, flush=True))
    S, rows = res["summary"], res["rows"]
    print(S, flush=True)
    net.close()
    assert res["loss"]["last"] < 0.5 * res["loss"]["first"]          # it learned
    assert S["x0_err_median"] < 0.6                                   # x0 beats the zero start (error 1)
    assert S["fp64_warm"]["used_x0"] == 16 and S["mixed_warm"]["used_x0"] == 16
    # fewer iterations from x0 on average, and never more than 2 extra on a case
    assert S["fp64_warm"]["itn_mean"] < S["fp64_zero"]["itn_mean"] - 0.5
    assert all(r["fp64_warm"]["itn"] <= r["fp64_zero"]["itn"] + 2 for r in rows)
    # every solve meets the tolerance and lands on the FP64 solution
    for r in rows:
        assert r["mixed_warm"]["test2"] <= 1e-6 and r["mixed_zero"]["test2"] <= 1e-6
        assert r["fp64_warm"]["xerr"] < 1e-4 and r["mixed_warm"]["xerr"] < 1e-3


def test_network_input_is_the_jacobi_step(P):
    """Host side, no GPU: the network's input is A^T b divided by the column
    sums of squares of [A_k; D] (the diagonal of A^T A), checked against
    scipy on a small system; "atb" passes A^T b through; unknown kinds fail."""
    import importlib
    import scipy.sparse as sp
    W = importlib.import_module(P.__name__ + ".warmstart")
    s = P.generate(9, 8, 7, 500, 1, 5, 0.01, 0.05)
    A = sp.vstack([sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)),
                   sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))]).tocsr()
    ssq = W.column_ssq(s)
    np.testing.assert_allclose(ssq, np.asarray(A.multiply(A).sum(axis=0)).ravel(), rtol=1e-13)
    atb = (A.T @ np.concatenate([s.b, np.zeros(s.m_d)]))[None, :]
    np.testing.assert_array_equal(W.net_input(atb, ssq, "atb"), atb)
    np.testing.assert_allclose(W.net_input(atb, ssq, "jacobi"), atb / ssq, rtol=1e-15)
    with pytest.raises(ValueError):
        W.net_input(atb, ssq, "other")
