"""The CPU oracle, pinned before it is trusted.

  * against the compiled FP64 reference (golden fixtures, bitwise x),
  * against closed forms (dnrm2 = sqrt(sum x^2), d2norm = hypot),
  * against dense brute force and the adjoint identity for the SpMVs,
  * against scipy.sparse.linalg.lsqr for the LSQR recurrences,
  * the standard-error branch (wantse) against a textbook Golub-Kahan
    bidiagonalisation with full reorthogonalisation,
  * and with planted mistakes, to show each check can fail.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import load_golden, rel

CASES = ["grid16_rays20k_itn100", "grid16_rays20k_itn1", "grid4_rays12_itn10",
         "grid7x5x3_rays40_itn20", "grid4_zero_rhs_itn5"]


def dense(s):
    Ak = sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)).toarray()
    D = sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n)).toarray()
    return np.vstack([Ak, D])


@pytest.mark.parametrize("name", CASES)
def test_generator_reproduces_fixture_system(P, MG, name):
    g = load_golden(name)
    s = MG.make_system(P, MG.CASES[name])
    assert MG.system_digest(s) == str(g["sha256"])
    assert [s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d] == list(g["dims"])


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_bitwise(P, O, MG, name):
    """x after itn iterations equals the compiled reference's x bit for bit."""
    g = load_golden(name)
    s = MG.make_system(P, MG.CASES[name])
    x, res, tr = O.Oracle(s).lsqr(s.b, int(g["itn"]))
    assert np.array_equal(x, g["x"]), f"max diff {np.abs(x - g['x']).max()}"


@pytest.mark.parametrize("name", [c for c in CASES if "zero" not in c])
def test_oracle_trace_matches_reference_print(P, O, MG, name):
    """alpha and beta as printed by the reference ("%e", lsqr.c:352,383,521,579)
    and the result.txt iteration lines (lsqr.c:443,809)."""
    g = load_golden(name)
    s = MG.make_system(P, MG.CASES[name])
    itn = int(g["itn"])
    x, res, tr = O.Oracle(s).lsqr(s.b, itn)
    assert len(g["alpha"]) == itn + 1 and len(g["beta"]) == itn + 1
    np.testing.assert_allclose(tr[:, 0], g["alpha"], rtol=6e-7)
    np.testing.assert_allclose(tr[:, 1], g["beta"], rtol=6e-7)
    for row in g["log"]:
        k = int(row[0])
        np.testing.assert_allclose(tr[k, 2], row[2], rtol=2e-9)          # rnorm %16.9e
        np.testing.assert_allclose(tr[k, 3], row[3], rtol=6e-3)          # test1 %9.2e
        np.testing.assert_allclose(tr[k, 4], row[4], rtol=6e-3)          # test2 %9.2e
        if not np.isnan(row[5]):
            np.testing.assert_allclose(tr[k, 5], row[5], rtol=6e-2)      # anorm %8.1e
            np.testing.assert_allclose(tr[k, 6], row[6], rtol=6e-2)      # acond %8.1e
    fin = dict(zip([str(k) for k in g["final_keys"]], g["final_vals"]))
    assert res["itn"] == fin["itn"] and res["istop"] == fin["istop"]
    for key in ("anorm", "acond", "xnorm", "rnorm"):
        np.testing.assert_allclose(res[key], fin[key], rtol=6e-6)
    np.testing.assert_allclose(res["bnorm"], fin["vnorm"], rtol=6e-6)     # printed as vnorm


def test_oracle_zero_rhs(P, O, MG):
    s = MG.make_system(P, MG.CASES["grid4_zero_rhs_itn5"])
    x, res, tr = O.Oracle(s).lsqr(s.b, 5)
    assert np.all(x == 0.0) and res["itn"] == 0 and res["istop"] == 0


def test_spmv_dense_bruteforce_and_adjoint(P, O, MG):
    for name in ("grid4_rays12_itn10", "grid7x5x3_rays40_itn20"):
        s = MG.make_system(P, MG.CASES[name])
        A = dense(s)
        rng = np.random.default_rng(5)
        o = O.Oracle(s)
        x = rng.standard_normal(s.n)
        y = rng.standard_normal(s.m_k + s.m_d)
        y1 = y.copy()
        o.aprod(1, x.copy(), y1)
        np.testing.assert_allclose(y1, y + A @ x, rtol=1e-13, atol=1e-13)
        x2 = x.copy()
        o.aprod(2, x2, y.copy())
        np.testing.assert_allclose(x2, x + A.T @ y, rtol=1e-13, atol=1e-13)
        # adjoint identity <A x, y> = <x, A^T y>
        ax = np.zeros_like(y)
        o.aprod(1, x.copy(), ax)
        aty = np.zeros_like(x)
        o.aprod(2, aty, y.copy())
        assert abs(ax @ y - x @ aty) <= 1e-12 * (abs(ax @ y) + 1)


def test_planted_transpose_mistake_is_caught(P, O, MG):
    """A transposed operand (using A instead of A^T in mode 2) fails the
    brute-force check, so that check is able to pin mode 2."""
    s = MG.make_system(P, MG.CASES["grid4_rays12_itn10"])
    A = dense(s)
    rng = np.random.default_rng(1)
    y = rng.standard_normal(A.shape[0])
    good = A.T @ y
    # the planted variant: A[:n,:n]^T y[:n] style confusion of operands
    wrong = (A[:, :]).T @ np.concatenate([y[s.m_k:], y[:s.m_k]])
    x = np.zeros(s.n)
    O.Oracle(s).aprod(2, x, y.copy())
    np.testing.assert_allclose(x, good, rtol=1e-13, atol=1e-13)
    assert rel(x, wrong) > 1e-3


def test_dnrm2_closed_form_and_special_cases(O):
    o = O.load()
    orc = O.Oracle.__new__(O.Oracle)
    orc.lib = o
    rng = np.random.default_rng(3)
    for n in (2, 3, 17, 1000):
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5)
        ref = math.sqrt(math.fsum(float(t) * float(t) for t in v))
        assert abs(orc.dnrm2(v) - ref) <= 4e-16 * n ** 0.5 * ref + 1e-300
    assert orc.dnrm2(np.zeros(0)) == 0.0
    assert orc.dnrm2(np.array([-3.5])) == 3.5
    assert orc.dnrm2(np.zeros(7)) == 0.0
    big = np.array([1e200, 1e200, 1e200])   # squares overflow; the scaled form does not
    assert abs(orc.dnrm2(big) - math.sqrt(3) * 1e200) <= 1e-15 * 1e200 * 2
    tiny = np.array([1e-200, 1e-200])
    assert abs(orc.dnrm2(tiny) - math.sqrt(2) * 1e-200) <= 1e-15 * 1e-200 * 2
    for a, b in ((3.0, 4.0), (0.0, 0.0), (-1e300, 1e300), (1e-310, 0.0), (2.5, -7.25)):
        assert abs(orc.d2norm(a, b) - math.hypot(a, b)) <= 4e-16 * math.hypot(a, b)


@pytest.mark.parametrize("name,itn,tol", [("grid4_rays12_itn10", 10, 1e-10),
                                          ("grid7x5x3_rays40_itn20", 20, 1e-10),
                                          ("grid16_rays20k_itn100", 10, 1e-8)])
def test_oracle_lsqr_matches_scipy(P, O, MG, name, itn, tol):
    """Same Golub-Kahan recurrences as scipy's LSQR (atol=btol=0, no conlim stop).
    scipy orders its norms and products differently, so the two drift apart by
    rounding (about 2e-10 after 10 iterations on the 16^3 system); a dropped
    term or sign would differ at O(1)."""
    s = MG.make_system(P, MG.CASES[name])
    A = sp.vstack([sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)),
                   sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))]).tocsr()
    rhs = np.concatenate([s.b, np.zeros(s.m_d)])
    ref = spla.lsqr(A, rhs, atol=0.0, btol=0.0, conlim=0.0, iter_lim=itn)
    x, res, tr = O.Oracle(s).lsqr(s.b, itn)
    assert ref[2] == itn
    assert rel(x, ref[0]) < tol
    np.testing.assert_allclose(res["rnorm"], ref[3], rtol=tol)
    np.testing.assert_allclose(res["anorm"], ref[5], rtol=tol)
    np.testing.assert_allclose(res["xnorm"], ref[8], rtol=tol)


def test_oracle_lsqr_invariants(P, O, MG):
    s = MG.make_system(P, MG.CASES["grid16_rays20k_itn100"])
    x, res, tr = O.Oracle(s).lsqr(s.b, 60)
    rn = tr[:, 2]
    assert np.all(np.diff(rn) <= 1e-12 * rn[0])             # rnorm never increases
    xn = tr[1:, 7]
    assert np.all(np.diff(xn) >= -1e-12 * xn[-1])           # xnorm never decreases
    # u, v orthonormal early in the bidiagonalisation
    A = sp.vstack([sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)),
                   sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))]).tocsr()
    r = np.concatenate([s.b, np.zeros(s.m_d)]) - A @ x
    np.testing.assert_allclose(np.linalg.norm(r), res["rnorm"], rtol=1e-8)
    # test2 as ||A^T r|| / (||A|| ||r||) with the estimate ||A|| <= ||A||_F
    assert np.linalg.norm(A.T @ r) / (res["anorm"] * res["rnorm"]) == pytest.approx(res["arnorm"], rel=1e-3)


def test_oracle_damped_lsqr_matches_scipy(P, O, MG):
    """The damp > 0 branch (lsqr.c:609-616), which the reference never takes
    (it passes damp = 0), against scipy's damped LSQR."""
    s = MG.make_system(P, MG.CASES["grid7x5x3_rays40_itn20"])
    A = sp.vstack([sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)),
                   sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))]).tocsr()
    rhs = np.concatenate([s.b, np.zeros(s.m_d)])
    ref = spla.lsqr(A, rhs, damp=0.7, atol=0.0, btol=0.0, conlim=0.0, iter_lim=15)
    x, res, tr = O.Oracle(s).lsqr(s.b, 15, damp=0.7)
    assert rel(x, ref[0]) < 1e-10
    # with damping, rnorm is the norm of the damped residual (scipy's r2norm)
    np.testing.assert_allclose(res["rnorm"], ref[4], rtol=1e-9)


def test_norm_rounding_alone_moves_x(P, O, MG):
    """LSQR iterates are sensitive to the rounding of ||u|| and ||v||: with the
    oracle's exact aprod and only the norms computed as a correctly rounded
    sum of squares instead of the reference's scaled dnrm2 (mycblas.c:194-219),
    x after 100 iterations on the 16^3 system moves by ~1e-10.  This bounds how
    close any implementation that does not reproduce dnrm2 bit for bit can be
    to the reference, and sets the default operator's tolerance."""
    import math
    g = load_golden("grid16_rays20k_itn100")
    s = MG.make_system(P, MG.CASES["grid16_rays20k_itn100"])
    orc = O.Oracle(s)
    mk, md, n = s.m_k, s.m_d, s.n

    def run(norm_u, norm_v, itn=100):
        u = np.concatenate([s.b, np.zeros(md)])
        v = np.zeros(n)
        x = np.zeros(n)
        beta = norm_u(u)
        u *= 1.0 / beta
        orc.aprod(2, v, u.copy())
        alpha = norm_v(v)
        v *= 1.0 / alpha
        w = v.copy()
        rhobar, phibar = alpha, beta
        for _ in range(itn):
            u *= -alpha
            orc.aprod(1, v.copy(), u)
            beta = norm_u(u)
            u *= 1.0 / beta
            v *= -beta
            orc.aprod(2, v, u.copy())
            alpha = norm_v(v)
            v *= 1.0 / alpha
            rho = orc.d2norm(rhobar, beta)
            cs, sn = rhobar / rho, beta / rho
            theta, rhobar = sn * alpha, -cs * alpha
            phi, phibar = cs * phibar, sn * phibar
            t = w.copy()
            x[:] = (phi / rho) * t + x
            w[:] = (-theta / rho) * t + v
        return x

    ref_u = lambda u: orc.lib.orc_norm_u(mk, O._d(np.ascontiguousarray(u[:mk])), md,
                                         O._d(np.ascontiguousarray(u[mk:])))
    ref_v = lambda v: orc.lib.orc_norm_v(n, O._d(v))
    x_ref = run(ref_u, ref_v)
    assert np.array_equal(x_ref, g["x"])                  # the mirror is the reference
    fs = lambda a: math.sqrt(math.fsum(a * a))
    d = rel(run(fs, fs), g["x"])
    assert 2e-11 < d < 1e-9, d


def golub_kahan_se(A, bf, k):
    """Textbook Golub-Kahan bidiagonalisation of A from bf, k steps with full
    reorthogonalisation, and the LSQR quantities it fixes: x_k = V_k y_k with
    y_k = argmin ||B_k y - beta1 e1||, rnorm_k = that minimum, and the
    variance estimate var_k = diag(V_k (B_k^T B_k)^-1 V_k^T) that LSQR
    accumulates as sum_j (w_j / rho_j)^2 (Paige and Saunders 1982)."""
    beta1 = np.linalg.norm(bf)
    U = [bf / beta1]
    v = A.T @ U[0]
    al = [np.linalg.norm(v)]
    V = [v / al[0]]
    be = [beta1]
    for _ in range(k):
        u = A @ V[-1] - al[-1] * U[-1]
        for q in U:
            u -= (q @ u) * q
        be.append(np.linalg.norm(u))
        U.append(u / be[-1])
        v = A.T @ U[-1] - be[-1] * V[-1]
        for q in V:
            v -= (q @ v) * q
        al.append(np.linalg.norm(v))
        V.append(v / al[-1])
    B = np.zeros((k + 1, k))
    for i in range(k):
        B[i, i] = al[i]
        B[i + 1, i] = be[i + 1]
    Vk = np.array(V[:k]).T
    rhs = np.zeros(k + 1)
    rhs[0] = beta1
    y = np.linalg.lstsq(B, rhs, rcond=None)[0]
    var = np.einsum("ij,jk,ik->i", Vk, np.linalg.inv(B.T @ B), Vk)
    return Vk @ y, np.linalg.norm(B @ y - rhs), var


@pytest.mark.parametrize("args", [(3, 3, 3, 60, 0, 7, 0.01, 0.3), (4, 3, 3, 80, 1, 11, 0.02, 0.2),
                                  (5, 4, 3, 200, 0, 13, 0.01, 0.05)])
def test_standard_errors_match_reorthogonalised_bidiagonalisation(P, O, args):
    """se = rnorm / sqrt(m - n) * sqrt(var) (lsqr.c:644-670, 878-891).  For
    the first few iterations the plain recurrences keep orthogonality to
    rounding, so the oracle must match the reorthogonalised textbook values
    to ~1e-14; the bar is 1e-12.  Scaling by sqrt(m) instead of sqrt(m - n),
    or dropping the 1/rho factor, misses it by 10% or more."""
    s = P.generate(*args)
    A = dense(s)
    m, n = A.shape
    bf = np.concatenate([s.b, np.zeros(s.m_d)])
    orc = O.Oracle(s)
    for k in (1, 2, 5):
        x, res, _ = orc.lsqr(s.b, k, want_se=True)
        xg, rn, var = golub_kahan_se(A, bf, k)
        se_g = rn / math.sqrt(m - n) * np.sqrt(var)
        assert rel(x, xg) < 1e-12, (k, rel(x, xg))
        assert abs(res["rnorm"] - rn) <= 1e-12 * rn
        assert np.abs(res["se"] - se_g).max() <= 1e-12 * np.abs(se_g).max(), k
        wrong = rn / math.sqrt(m) * np.sqrt(var)
        assert np.abs(wrong - se_g).max() > 0.05 * np.abs(se_g).max()
