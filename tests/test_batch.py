"""Batched solves (plsqr_batch_*): many independent systems, up to 4
right-hand sides each, against the oracle run on each (system, right-hand
side) pair alone.

Bars: the FP64 batch is held to the same tolerances as the single-system
default operator (1e-13 relative per aprod, 1e-10 relative in x after the same
iteration count); the FP32 batch kernels to the FP32 rounding of their operands
(2e-6 relative per aprod); mixed-precision solves to the FP64 test2 of their
true residual.
"""
import numpy as np
import pytest

from conftest import rel

# uneven grids, both ray geometries, a system with a single kernel row chunk
SPECS = [
    dict(grid=(6, 5, 4), n_rays=200, geometry=0, seed=31, lam=0.2),
    dict(grid=(9, 9, 9), n_rays=3000, geometry=0, seed=32, lam=0.1),
    dict(grid=(12, 8, 10), n_rays=4000, geometry=1, seed=33, lam=0.05),
    dict(grid=(16, 16, 16), n_rays=20000, geometry=0, seed=2013, lam=0.1),
]


def make_batch_inputs(P, specs, n_rhs, zero=()):
    systems, B = [], []
    for i, sp in enumerate(specs):
        nx, ny, nz = sp["grid"]
        s = P.generate(nx, ny, nz, sp["n_rays"], sp["geometry"], sp["seed"], 0.01, sp["lam"])
        bs = []
        for k in range(n_rhs):
            b, _ = P.gen_rhs(s, nx, ny, nz, sp["seed"] * 16 + k + 1, 0.01 + 0.01 * k)
            if (i, k) in zero:
                b[:] = 0.0
            bs.append(b)
        systems.append(s)
        B.append(np.stack(bs))
    return systems, B


# -- host side (no GPU needed) -------------------------------------------------

def test_batch_config_is_deterministic(P):
    a = P.batch_specs("batch24_grid20to48_rhs4")
    b = P.batch_specs("batch24_grid20to48_rhs4")
    assert a == b and len(a) == 24
    for sp in a:
        e = sp["grid"][0]
        assert 20 <= e <= 48 and sp["grid"] == (e, e, e) and sp["n_rays"] == 8 * e ** 3
        assert len(sp["rhs"]) == 4 and all(0.01 <= s <= 0.05 for _, s in sp["rhs"])
    assert len({sp["grid"] for sp in a}) > 5          # uneven sizes


def test_batch_generator_shapes(P):
    systems, B, XT = P.generate_batch(specs=P.batch_specs(
        n_problems=3, edge_lo=5, edge_hi=7, rays_per_col=8, n_rhs=2, geometry=0, seed=5,
        sigma_lo=0.01, sigma_hi=0.05, lam=0.1))
    for s, b, xt in zip(systems, B, XT):
        assert b.shape == (2, s.m_k) and xt.shape == (2, s.n) and s.m_k == 8 * s.n
        assert not np.array_equal(b[0], b[1])


def test_batch_create_validates_on_the_host(P):
    """Argument and system checks run before any device call."""
    s = P.generate(6, 5, 4, 200, 0, 1, 0.01, 0.1)
    with pytest.raises(P.PlsqrError) as e:
        P.Batch([s], 5)
    assert e.value.code == 1
    ri = s.k_row_idx.copy()
    ri[3] = s.m_k + 1
    bad = P.System.from_arrays(s.m_k, s.m_d, s.n, s.k_col_ptr, ri, s.k_val, s.d_row_ptr, s.d_col_idx,
                               s.d_val, s.dc_col_ptr, s.dc_row_idx, s.dc_val, s.b)
    with pytest.raises(P.PlsqrError) as e:
        P.Batch([s, bad], 2, grids=[(6, 5, 4), None])
    assert e.value.code == 2 and "system 1" in str(e.value)
    with pytest.raises(P.PlsqrError) as e:
        P.Batch([s], 1, grids=[(6, 5, 5)])
    assert e.value.code == 1 and "grid" in str(e.value)
    try:  # a valid batch gets as far as the device
        P.Batch([s, s], 3).close()
    except P.PlsqrError as e:
        assert e.code == 4, str(e)


def test_batch_layout_self_check_on_host(P, monkeypatch):
    """With PLSQR_TILE_CHECK=1 every problem's brick layout (one row slice,
    so bricks crossed by more than 1024 rays are staged in rounds) is
    evaluated round by round through its cuts on the host and compared with
    the CSC before any device call; without a GPU creation then stops at
    hipGetDevice (code 4), which means the layout passed."""
    monkeypatch.setenv("PLSQR_TILE_CHECK", "1")
    systems, _ = make_batch_inputs(P, SPECS, 1)
    big = P.generate(16, 16, 16, 60000, 0, 7, 0.01, 0.1)   # ~3000 segments per brick
    try:
        P.Batch(systems + [big], 1).close()
    except P.PlsqrError as e:
        assert e.code == 4 and "self-check" not in str(e), str(e)


# -- GPU -------------------------------------------------------------------------

def announce(msg):
    print(msg, flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("n_rhs", [1, 2, 3, 4])
def test_batch_aprod_matches_oracle(P, O, gpu_available, n_rhs):
    announce(f"batch aprod K={n_rhs}")
    systems, _ = make_batch_inputs(P, SPECS, n_rhs)
    rng = np.random.default_rng(n_rhs)
    X = [rng.standard_normal((n_rhs, s.n)) for s in systems]
    Y = [rng.standard_normal((n_rhs, s.m_k + s.m_d)) for s in systems]
    with P.Batch(systems, n_rhs, flags=P.BATCH_NO_TUNE) as bt:
        for mode in (1, 2):
            for prec, tol in ((P.PREC_FP64, 1e-13), (P.PREC_MIXED, 2e-6)):
                Xg, Yg = bt.aprod(mode, [x.copy() for x in X], [y.copy() for y in Y], precision=prec)
                for p, s in enumerate(systems):
                    orc = O.Oracle(s)
                    for k in range(n_rhs):
                        xo, yo = X[p][k].copy(), Y[p][k].copy()
                        orc.aprod(mode, xo, yo)
                        got, want = (Yg[p][k], yo) if mode == 1 else (Xg[p][k], xo)
                        assert rel(got, want) < tol, (mode, prec, p, k, rel(got, want))
                        other = Xg[p][k] if mode == 1 else Yg[p][k]
                        assert np.array_equal(other, X[p][k] if mode == 1 else Y[p][k])


@pytest.mark.gpu
def test_batch_every_launch_shape_matches_oracle(P, O, gpu_available):
    """Each shape the autotuner may pick computes the same operator."""
    announce("batch shapes")
    systems, _ = make_batch_inputs(P, SPECS[1:3], 4)
    rng = np.random.default_rng(9)
    X = [rng.standard_normal((4, s.n)) for s in systems]
    Y = [rng.standard_normal((4, s.m_k + s.m_d)) for s in systems]
    want1, want2 = [], []
    for p, s in enumerate(systems):
        orc = O.Oracle(s)
        w1, w2 = [], []
        for k in range(4):
            xo, yo = X[p][k].copy(), Y[p][k].copy()
            orc.aprod(1, xo, yo)
            w1.append(yo)
            xo, yo = X[p][k].copy(), Y[p][k].copy()
            orc.aprod(2, xo, yo)
            w2.append(xo)
        want1.append(w1)
        want2.append(w2)
    with P.Batch(systems, 4, flags=P.BATCH_NO_TUNE) as bt:
        for shape in P.BATCH_SHAPES:
            for prec, tol in ((P.PREC_FP64, 1e-13), (P.PREC_MIXED, 2e-6)):
                bt.set_tuning(prec, **shape)
                _, Yg = bt.aprod(1, X, [y.copy() for y in Y], precision=prec)
                Xg, _ = bt.aprod(2, [x.copy() for x in X], Y, precision=prec)
                for p in range(len(systems)):
                    for k in range(4):
                        assert rel(Yg[p][k], want1[p][k]) < tol, (shape, prec, p, k)
                        assert rel(Xg[p][k], want2[p][k]) < tol, (shape, prec, p, k)


def oracle_sensitivity(orc, b, itn):
    """Relative change of the oracle's x after itn iterations when b is
    perturbed by 1e-15 relative: how far rounding differences alone can move
    the iterate at that point of the run."""
    x, _, _ = orc.lsqr(b, itn, trace=False)
    b2 = b * (1 + 1e-15 * np.random.default_rng(1).standard_normal(b.size))
    x2, _, _ = orc.lsqr(b2, itn, trace=False)
    return rel(x2, x)


# Mid-run LSQR iterates of these systems are hypersensitive: a 1e-15 relative
# change of b moves x after 20 iterations by up to 6.5e-4 (12x8x10 grid), and
# any change of summation order does the same.  Once a pair has converged
# (test2 < 1e-12) the same perturbation moves x by 2e-15.  The fixed-count
# comparison is therefore made after 250 iterations (converged for every
# pair), after 5 (before the sensitivity builds up), and the tol-stopped runs
# are held to max(1e-10, 100 x the oracle's measured sensitivity at the pair's
# own stopping iteration): the single perturbation of b underestimates a
# summation order that differs in every operation of every iteration (20 x
# was exceeded once, 5.8e-10 against 4.7e-10, at 53 iterations).
@pytest.mark.gpu
@pytest.mark.parametrize("n_rhs", [1, 3, 4])
def test_batch_fp64_lsqr_matches_oracle(P, O, gpu_available, n_rhs):
    """Fixed iteration counts: every pair within 1e-10 of the oracle run alone."""
    announce(f"batch lsqr fp64 K={n_rhs}")
    systems, B = make_batch_inputs(P, SPECS, n_rhs, zero={(0, n_rhs - 1)})
    with P.Batch(systems, n_rhs) as bt:
        runs = {itn: bt.solve(B, itn) for itn in (5, 250)}
        X2, _, _ = bt.solve(B, 250)
    for itn, (X, infos, tot) in runs.items():
        assert tot > 0
        for p, s in enumerate(systems):
            orc = O.Oracle(s)
            for k in range(n_rhs):
                if itn == 250:
                    assert np.array_equal(X[p][k], X2[p][k]), "not run-to-run reproducible"
                if not B[p][k].any():
                    assert not X[p][k].any() and infos[p][k]["itn"] == 0
                    continue
                xo, res, _ = orc.lsqr(B[p][k], itn)
                assert infos[p][k]["itn"] == itn
                assert rel(X[p][k], xo) < 1e-10, (itn, p, k, rel(X[p][k], xo))
                np.testing.assert_allclose(infos[p][k]["bnorm"], res["bnorm"], rtol=1e-13)
                if itn == 5:
                    np.testing.assert_allclose(infos[p][k]["rnorm"], res["rnorm"], rtol=1e-10)


@pytest.mark.gpu
def test_batch_convergence_masking(P, O, gpu_available):
    """With tol, each pair stops on its own test2; its x is the oracle's x at
    that pair's own iteration count while the other pairs keep going."""
    announce("batch masking")
    systems, B = make_batch_inputs(P, SPECS, 4, zero={(2, 1)})
    tol = 1e-9
    with P.Batch(systems, 4) as bt:
        X, infos, tot = bt.solve(B, 400, tol=tol)
    itns = []
    for p, s in enumerate(systems):
        orc = O.Oracle(s)
        for k in range(4):
            inf = infos[p][k]
            if not B[p][k].any():
                assert inf["itn"] == 0 and not X[p][k].any() and inf["elapsed_ms"] <= tot
                continue
            assert inf["arnorm"] <= tol and inf["itn"] < 400
            xo, res, tro = orc.lsqr(B[p][k], inf["itn"])
            bar = max(1e-10, 100 * oracle_sensitivity(orc, B[p][k], inf["itn"]))
            assert rel(X[p][k], xo) < bar, (p, k, inf["itn"], rel(X[p][k], xo), bar)
            # the oracle's own trajectory crosses tol within a few iterations of it
            _, _, tro = orc.lsqr(B[p][k], 400)
            k_o = int(np.argmax(tro[1:, 4] <= tol)) + 1
            assert abs(k_o - inf["itn"]) <= 3, (k_o, inf["itn"])
            assert 0 < inf["elapsed_ms"] <= tot * (1 + 1e-6)
            itns.append(inf["itn"])
    assert len(set(itns)) > 3, itns                    # the pairs did stop at different points


@pytest.mark.gpu
def test_batch_mixed_refinement(P, O, gpu_available):
    announce("batch mixed")
    import scipy.sparse as sp
    systems, B = make_batch_inputs(P, SPECS[1:], 2)
    with P.Batch(systems, 2, flags=P.BATCH_TUNE_MIXED) as bt:
        X, infos, tot = bt.solve(B, 3000, tol=1e-6, precision=P.PREC_MIXED, refine_every=25)
        X64, inf64, _ = bt.solve(B, 3000, tol=1e-6)
    for p, s in enumerate(systems):
        A = sp.vstack([sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)),
                       sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))]).tocsr()
        x_tight, _, _ = O.Oracle(s).lsqr(B[p][0], 600)
        for k in range(2):
            inf = infos[p][k]
            assert inf["refinements"] >= 1 and inf["arnorm"] <= 1e-6, inf
            r = np.concatenate([B[p][k], np.zeros(s.m_d)]) - A @ X[p][k]
            assert np.linalg.norm(A.T @ r) / (inf["anorm"] * np.linalg.norm(r)) <= 1e-6 * 1.001
            if k == 0:
                e_m, e_64 = rel(X[p][k], x_tight), rel(X64[p][k], x_tight)
                assert e_m < 3 * e_64 + 1e-7, (p, e_m, e_64)


@pytest.mark.gpu
def test_batch_autotune_reports_a_valid_shape(P, gpu_available):
    announce("batch autotune")
    systems, B = make_batch_inputs(P, SPECS[1:3], 4)
    with P.Batch(systems, 4, flags=P.BATCH_NO_TUNE) as bt:
        assert bt.get_tuning()["candidates"] == 0
        X0, _, _ = bt.solve(B, 250)
        t = bt.autotune(P.PREC_FP64, reps=3)
        assert t["candidates"] == 22 and t["ms_rows"] > 0 and t["ms_cols"] > 0
        assert t["rows_per_wave"] in (2, 4, 8, 16) and t["block_cols"] in (128, 256, 512)
        assert t["gather_kib"] in (32, 64)
        assert bt.get_tuning() == t
        X1, _, _ = bt.solve(B, 250)
    for a, b in zip(X0, X1):     # converged pairs: the launch shape changes rounding only
        assert rel(a, b) < 1e-12


@pytest.mark.gpu
def test_batch_mixed_column_scaling(P, O, gpu_available):
    """Column-scaled mixed batch (each problem on A_p D_p, D_p =
    diag(||A_p e_j||^-1.5)): the first inner iterations follow textbook LSQR
    on A_p D_p; the stop holds on the FP64 test2 of A_p with the FP64 batch's
    anorm per pair and with ||A_p||_F; the unscaled FP32 operator is restored
    afterwards (bitwise the same unscaled solve)."""
    announce("batch mixed, column scaling")
    import scipy.sparse as sp
    from test_gpu_parity import lsqr_scaled_np
    systems, B = make_batch_inputs(P, SPECS[1:], 2)
    k_it = 8
    with P.Batch(systems, 2, flags=P.BATCH_TUNE_MIXED) as bt:
        Xm0, _, _ = bt.solve(B, 3000, tol=1e-6, precision=P.PREC_MIXED, refine_every=25)
        X64, inf64, _ = bt.solve(B, 3000, tol=1e-6)
        Xk, _, _ = bt.solve(B, k_it, tol=0.0, precision=P.PREC_MIXED, refine_every=k_it,
                            precond=P.PRECOND_COLSCALE, precond_power=1.5)
        bt.set_anorm_refs([[inf64[p][k]["anorm"] for k in range(2)] for p in range(len(systems))])
        Xs, infs, _ = bt.solve(B, 3000, tol=1e-6, precision=P.PREC_MIXED, refine_every=25,
                               precond=P.PRECOND_COLSCALE, precond_power=1.5)
        bt.set_anorm_refs(None)
        Xf, inff, _ = bt.solve(B, 3000, tol=1e-6, precision=P.PREC_MIXED, refine_every=25,
                               precond=P.PRECOND_COLSCALE, precond_power=1.5)
        Xm1, _, _ = bt.solve(B, 3000, tol=1e-6, precision=P.PREC_MIXED, refine_every=25)
        with pytest.raises(P.PlsqrError):
            bt.solve(B, 10, precond=P.PRECOND_COLSCALE)  # FP64 has no preconditioner
    for p in range(len(systems)):
        for k in range(2):
            assert np.array_equal(Xm0[p][k], Xm1[p][k]), (p, k)
    for p, s in enumerate(systems):
        A = sp.vstack([sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)),
                       sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))]).tocsr()
        cn = np.sqrt(np.asarray(A.multiply(A).sum(axis=0)).ravel())
        d = np.where(cn > 0, np.where(cn > 0, cn, 1.0) ** -1.5, 1.0)
        fro = sp.linalg.norm(A)
        x_tight, _, _ = O.Oracle(s).lsqr(B[p][0], 600)
        for k in range(2):
            bb = np.concatenate([B[p][k], np.zeros(s.m_d)])
            e_same = rel(Xk[p][k], lsqr_scaled_np(A, bb, k_it, d))
            e_other = rel(Xk[p][k], lsqr_scaled_np(A, bb, k_it, np.ones(s.n)))
            assert e_same < 1e-3 and e_same < 0.05 * e_other, (p, k, e_same, e_other)
            for X, inf, an in ((Xs, infs, inf64[p][k]["anorm"]), (Xf, inff, fro)):
                i = inf[p][k]
                assert i["arnorm"] <= 1e-6, (p, k, i)
                np.testing.assert_allclose(i["anorm"], an, rtol=1e-12)
                r = bb - A @ X[p][k]
                assert np.linalg.norm(A.T @ r) / (an * np.linalg.norm(r)) <= 1e-6 * 1.001
                if k == 0:
                    e_s, e_64 = rel(X[p][k], x_tight), rel(X64[p][k], x_tight)
                    assert e_s < 3 * e_64 + 1e-7, (p, e_s, e_64)
        print(f"problem {p}: iterations FP64 {[inf64[p][k]['itn'] for k in range(2)]}, scaled mixed "
              f"{[infs[p][k]['itn'] for k in range(2)]}", flush=True)


@pytest.mark.gpu
def test_batch_mixed_adaptive_checks(P, O, gpu_available):
    """Adaptive FP64 checks in the batched mixed solve: every pair still stops
    on its true FP64 test2 <= tol (checked here with scipy), with no more
    checks than the fixed schedule and no more iterations."""
    announce("batch mixed, adaptive checks")
    import scipy.sparse as sp
    systems, B = make_batch_inputs(P, SPECS[1:], 2)
    with P.Batch(systems, 2, flags=P.BATCH_TUNE_MIXED) as bt:
        X64, inf64, _ = bt.solve(B, 3000, tol=1e-6)
        bt.set_anorm_refs([[inf64[p][k]["anorm"] for k in range(2)] for p in range(len(systems))])
        kw = dict(tol=1e-6, precision=P.PREC_MIXED, refine_every=25, precond=P.PRECOND_COLSCALE,
                  precond_power=1.5)
        Xf, inff, tf = bt.solve(B, 3000, **kw)
        Xa, infa, ta = bt.solve(B, 3000, refine_adaptive=True, **kw)
    for p, s in enumerate(systems):
        A = sp.vstack([sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)),
                       sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))]).tocsr()
        for k in range(2):
            r = np.concatenate([B[p][k], np.zeros(s.m_d)]) - A @ Xa[p][k]
            assert infa[p][k]["arnorm"] <= 1e-6
            assert np.linalg.norm(A.T @ r) / (inf64[p][k]["anorm"] * np.linalg.norm(r)) <= 1e-6 * 1.001
            assert infa[p][k]["refinements"] <= inff[p][k]["refinements"] + 1, (p, k)
    it_f = max(i["itn"] for row in inff for i in row)
    it_a = max(i["itn"] for row in infa for i in row)
    print(f"slowest pair: {it_f} iterations with checks every 25, {it_a} adaptive; batch {tf:.1f} / {ta:.1f} ms",
          flush=True)
    assert it_a <= it_f + 25  # a check may come up to one period later
    # a check costs the whole batch: adaptive chunks never end before a full
    # period, so the batch runs no more checks than the fixed schedule
    ev_f = max(i["refinements"] for row in inff for i in row)
    ev_a = max(i["refinements"] for row in infa for i in row)
    assert ev_a <= ev_f, (ev_a, ev_f)
