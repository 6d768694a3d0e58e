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
