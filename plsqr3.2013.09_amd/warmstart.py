"""Learned warm start: training pairs, training and evaluation.

Every step runs on the GPU through libplsqr_hip.so:
  * pairs: x_true from the generator's blob field of a seed, b = A_k x_true
    + sigma N(0, 1) (noise from numpy's seeded generator), A^T b, all through
    the batched operator (4 right-hand sides per launch); the network's input
    is A^T b divided by the diagonal of A^T A (input_kind "jacobi", the
    default) or A^T b itself ("atb");
  * targets x*: plain FP64 LSQR (the batched solver, 4 right-hand sides at a
    time) run to test2 <= target_tol;
  * training: WarmStartNet.train (bf16 MFMA kernels, Adam);
  * evaluation: plain FP64 LSQR from zero against LSQR from the network's x0
    (FP64 and mixed precision), each run to test2 <= tol on the single-system
    operator.
The matrix is fixed (one ray geometry, the deployed scanner); the pairs
differ in the blob field and the noise.
"""
import time

import numpy as np

from . import Batch, Operator, PREC_FP64, PREC_MIXED, PRECOND_COLSCALE, WarmStartNet, gen_field


def make_pairs(batch, system, grid, seeds, sigma):
    """(B (count, m_k), ATB (count, n)) for the seeds; batch is a 1-system Batch."""
    K = batch.K
    m, n = system.m_k + system.m_d, system.n
    Bs, ATB = [], []
    for i in range(0, len(seeds), K):
        chunk = list(seeds[i:i + K])
        xt = np.zeros((K, n))
        for k, sd in enumerate(chunk):
            xt[k] = gen_field(*grid, sd)
        _, Y = batch.aprod(1, [xt], [np.zeros((K, m))])
        b = Y[0][:, :system.m_k].copy()
        for k, sd in enumerate(chunk):
            b[k] += sigma * np.random.default_rng(int(sd)).standard_normal(system.m_k)
        y = np.zeros((K, m))
        y[:, :system.m_k] = b
        X, _ = batch.aprod(2, [np.zeros((K, n))], [y])
        for k in range(len(chunk)):
            Bs.append(b[k])
            ATB.append(X[0][k])
    return np.array(Bs), np.array(ATB)


def solve_targets(batch, B, tol, max_iters=5000, log=None):
    """Plain FP64 LSQR solutions of the rows of B, 4 at a time; (X, iterations).
    log (optional) receives a progress line about every 30 s."""
    K = batch.K
    X, its = [], []
    t_last = time.perf_counter()
    for i in range(0, len(B), K):
        if log is not None and time.perf_counter() - t_last > 30.0:
            log(f"[warm] targets: {i} of {len(B)} solved")
            t_last = time.perf_counter()
        chunk = B[i:i + K]
        pad = np.zeros((K, B.shape[1]))
        pad[:len(chunk)] = chunk
        Xc, infos, _ = batch.solve([pad], max_iters, tol=tol)
        for k in range(len(chunk)):
            X.append(Xc[0][k])
            its.append(infos[0][k]["itn"])
    return np.array(X), np.array(its)


def column_ssq(system):
    """Sums of squares of the columns of [A_k; D] (the diagonal of A^T A), from
    the kernel CSC and the damping CSC on the host."""
    n = system.n
    kc = np.diff(system.k_col_ptr.astype(np.int64))
    dc = np.diff(system.dc_col_ptr.astype(np.int64))
    ssq = np.bincount(np.repeat(np.arange(n), kc), weights=system.k_val ** 2, minlength=n)
    ssq += np.bincount(np.repeat(np.arange(n), dc), weights=system.dc_val ** 2, minlength=n)
    return ssq


def net_input(ATB, colss, kind):
    """The network's input: the back-projection A^T b ("atb"), or A^T b
    divided by the diagonal of A^T A ("jacobi", one Jacobi step of the normal
    equations from zero), which removes the ray-density weighting."""
    if kind == "atb":
        return ATB
    if kind == "jacobi":
        return ATB / np.where(colss > 0, colss, 1.0)
    raise ValueError(f"unknown network input {kind!r}")


def train(net, ATB, X, epochs, batch=4, lr=2e-3, lr_final=5e-4, seed=5):
    """Adam for `epochs` passes over the pairs (a first phase at lr, the last
    third at lr_final); returns the loss per step."""
    net.set_data(ATB, X)
    steps = max(1, int(np.ceil(epochs * len(ATB) / batch)))
    s1 = max(1, (2 * steps) // 3)
    l1 = net.train(s1, batch, lr=lr, seed=seed)
    l2 = net.train(steps - s1, batch, lr=lr_final, seed=seed + 1) if steps > s1 else np.zeros(0)
    return np.concatenate([l1, l2])


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / nb) if nb > 0 else float(np.linalg.norm(a))


def evaluate(op, B, X0, tol, refine_every, max_iters=20000, x_ref=None, colscale=False, refine_adaptive=False,
             precond_power=1.0, log=None):
    """Per held-out right-hand side: FP64 from zero (the reference method),
    FP64 from x0, mixed precision from zero and from x0; with colscale also
    the mixed solves on A D (column scaling), their FP64 test2 taking the
    anorm the plain FP64 solve of the same b reported, so that they stop as
    strictly as it does.  Times are the device loop times of each solve."""
    rows = []
    for j, b in enumerate(B):
        r = {}
        x64, i64 = op.lsqr(b, max_iters, tol=tol)
        r["fp64_zero"] = dict(itn=i64["itn"], ms=i64["elapsed_ms"], test2=i64["arnorm"])
        xr = x64 if x_ref is None else x_ref[j]
        xw, iw = op.lsqr_x0(b, X0[j], max_iters, tol=tol)
        r["fp64_warm"] = dict(itn=iw["itn"], ms=iw["elapsed_ms"], used_x0=iw["used_x0"], xerr=rel(xw, xr))
        t0 = time.perf_counter()
        mk = dict(precision=PREC_MIXED, refine_every=refine_every, refine_adaptive=refine_adaptive)
        xm, im = op.lsqr_x0(b, None, max_iters, tol=tol, **mk)
        r["mixed_zero"] = dict(itn=im["itn"], ms=im["elapsed_ms"], xerr=rel(xm, xr), test2=im["arnorm"])
        xmw, imw = op.lsqr_x0(b, X0[j], max_iters, tol=tol, **mk)
        r["mixed_warm"] = dict(itn=imw["itn"], ms=imw["elapsed_ms"], used_x0=imw["used_x0"],
                               xerr=rel(xmw, xr), test2=imw["arnorm"])
        if colscale:
            kw = dict(mk, precond=PRECOND_COLSCALE, anorm_ref=i64["anorm"], precond_power=precond_power)
            xc, ic = op.lsqr_x0(b, None, max_iters, tol=tol, **kw)
            r["mixed_zero_cs"] = dict(itn=ic["itn"], ms=ic["elapsed_ms"], xerr=rel(xc, xr), test2=ic["arnorm"])
            xcw, icw = op.lsqr_x0(b, X0[j], max_iters, tol=tol, **kw)
            r["mixed_warm_cs"] = dict(itn=icw["itn"], ms=icw["elapsed_ms"], used_x0=icw["used_x0"],
                                      xerr=rel(xcw, xr), test2=icw["arnorm"])
        r["x0_err"] = rel(X0[j], xr)
        del t0
        rows.append(r)
        if log is not None:
            log(f"[warm] evaluated {j + 1} of {len(B)}")
    return rows


def summarize(rows):
    out = {}
    keys = [k for k in ("fp64_zero", "fp64_warm", "mixed_zero", "mixed_warm", "mixed_zero_cs", "mixed_warm_cs")
            if k in rows[0]]
    for key in keys:
        its = np.array([r[key]["itn"] for r in rows], float)
        ms = np.array([r[key]["ms"] for r in rows], float)
        out[key] = {"itn_median": float(np.median(its)), "itn_mean": float(its.mean()),
                    "ms_median": float(np.median(ms)), "ms_mean": float(ms.mean())}
        if "xerr" in rows[0][key]:
            out[key]["xerr_max"] = float(max(r[key]["xerr"] for r in rows))
        if "used_x0" in rows[0][key]:
            out[key]["used_x0"] = int(sum(bool(r[key]["used_x0"]) for r in rows))
    out["x0_err_median"] = float(np.median([r["x0_err"] for r in rows]))
    base = out["fp64_zero"]["ms_mean"]
    for key in keys[1:]:
        out[key]["speedup_vs_fp64_zero"] = base / out[key]["ms_mean"] if out[key]["ms_mean"] > 0 else None
    return out


def run(system, grid, n_train, n_eval, sigma, seed0, target_tol, epochs, tol, refine_every, log=print,
        net_seed=3, lr=2e-3, colscale=False, refine_adaptive=False, precond_power=1.0, input_kind="jacobi"):
    """The whole pipeline on one system; returns a result dict and the trained net."""
    t = {}
    t0 = time.perf_counter()
    bt = Batch([system], 4, grids=[grid])
    op = Operator(system, grid=grid)
    t["setup_s"] = time.perf_counter() - t0
    seeds = [seed0 + i for i in range(n_train + n_eval)]
    t0 = time.perf_counter()
    B, ATB = make_pairs(bt, system, grid, seeds, sigma)
    ATB = net_input(ATB, column_ssq(system), input_kind)
    t["pairs_s"] = time.perf_counter() - t0
    log(f"[warm] {len(seeds)} pairs ({t['pairs_s']:.1f} s)")
    t0 = time.perf_counter()
    X, its = solve_targets(bt, B[:n_train], target_tol, log=log)
    t["targets_s"] = time.perf_counter() - t0
    log(f"[warm] {n_train} FP64 targets to test2 <= {target_tol:g}: iterations median {np.median(its):.0f} "
        f"({t['targets_s']:.1f} s)")
    bt.close()
    net = WarmStartNet(grid, max_batch=4, seed=net_seed)
    t0 = time.perf_counter()
    loss = train(net, ATB[:n_train], X, epochs, lr=lr)
    t["train_s"] = time.perf_counter() - t0
    log(f"[warm] trained {len(loss)} steps, loss {loss[:5].mean():.3f} -> {loss[-20:].mean():.3f} "
        f"({t['train_s']:.1f} s)")
    t0 = time.perf_counter()
    X0 = net.predict(ATB[n_train:])
    t["predict_s"] = time.perf_counter() - t0
    t["predict_ms_per_sample"] = 1000 * t["predict_s"] / max(n_eval, 1)
    t0 = time.perf_counter()
    rows = evaluate(op, B[n_train:], X0, tol, refine_every, colscale=colscale, refine_adaptive=refine_adaptive,
                    precond_power=precond_power, log=log)
    t["eval_s"] = time.perf_counter() - t0
    op.close()
    res = {"summary": summarize(rows), "rows": rows, "times": t,
           "loss": {"first": float(loss[:5].mean()), "last": float(loss[-20:].mean()), "steps": int(len(loss))},
           "targets": {"tol": target_tol, "itn_median": float(np.median(its))}}
    return res, net
