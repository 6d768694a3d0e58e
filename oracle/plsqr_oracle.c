/*
 * plsqr_oracle.c -- CPU restatement of PLSQR3's LSQR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP implementation
 * is compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * It restates, in the reference's arithmetic order (single rank, numProc==1,
 * MY_FLOAT_TYPE = double, no FMA contraction), the functions of:
 *   source/lsqr.c:195-917          lsqr()                       -> orc_lsqr
 *   source/aprod_impl.c:22-267     aprod_mode_1_tridiagonal     -> orc_aprod (mode 1)
 *   source/aprod_impl.c:276-524    aprod_mode_2_tridiagonal     -> orc_aprod (mode 2)
 *   source/CPUSpMV.c:162-183       cpuSpMV_CSC                  -> orc_spmv_csc
 *   source/CPUSpMV.c:50-89         cpuSpMV_CSR_v2               -> orc_spmv_csr_v2
 *   source/CPUSpMV.c:203-245,255-296 cpuSpMVTranspose_CSC(_v2)  -> orc_spmvT_csc_v2
 *   source/mycblas.c:182-222       cblas_dnrm2 (scaled ssq)     -> orc_dnrm2
 *   source/mycblas.c:99-133,158-180 split / parallel norms      -> orc_norm_u, orc_norm_v
 *   source/mycblas.c:286-309       cblas_dscal                  -> orc_dscal
 *   source/cblas.h:53-63           d2norm                       -> orc_d2norm
 *
 * Pinning (what checks each function; see tests/):
 *   orc_spmv_* / orc_aprod : dense brute force on small systems, the adjoint
 *                            identity <Ax,y> = <x,A^T y>, and the compiled
 *                            reference run (golden x of the 16^3 system).
 *   orc_dnrm2, orc_d2norm  : closed forms (sqrt of the sum of squares, hypot)
 *                            within a few ulps, special cases n<=1, zeros.
 *   orc_lsqr               : scipy.sparse.linalg.lsqr on small systems; the
 *                            golden x (bitwise) and alpha/beta trace written
 *                            by the compiled FP64 reference.
 *   orc_lsqr standard errors (se) : the reference always passes se = NULL
 *                            (lsqr.c:80), so no reference output exists; pinned
 *                            against a textbook Golub-Kahan bidiagonalisation
 *                            with full reorthogonalisation (numpy), whose
 *                            rnorm / sqrt(m - n) * sqrt(diag(V (B^T B)^-1 V^T))
 *                            the first iterations must reproduce to rounding.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "plsqr_oracle.h"

/* CPUSpMV.c:162-183: for each column (ascending), Ax[row_idx[e]] += val[e]*x[c] */
void orc_spmv_csc(int n, const int *col_ptr, const int *row_idx, const double *val,
                  const double *x, double *Ax)
{
    int c, e;
    for (c = 0; c < n; c++) {
        const double xc = x[c];
        for (e = col_ptr[c]; e < col_ptr[c + 1]; e++)
            Ax[row_idx[e]] += val[e] * xc;
    }
}

/* CPUSpMV.c:50-89 (serial branch): row-sequential dot starting from Ax[row] */
void orc_spmv_csr_v2(int m, const int *row_ptr, const int *col_idx, const double *val,
                     const double *x, int start, double *Ax)
{
    int r, e;
    for (r = 0; r < m; r++) {
        double t = Ax[r];
        for (e = row_ptr[r]; e < row_ptr[r + 1]; e++)
            t += val[e] * x[col_idx[e] - start];
        Ax[r] = t;
    }
}

/* CPUSpMV.c:203-245 (start = 0) and 255-296: column gather-dot, rows ascending */
void orc_spmvT_csc_v2(int n, const int *col_ptr, const int *row_idx, const double *val,
                      const double *y, int start, double *tAy)
{
    int c, e;
    for (c = 0; c < n; c++) {
        double t = tAy[c];
        for (e = col_ptr[c]; e < col_ptr[c + 1]; e++)
            t += val[e] * y[row_idx[e] - start];
        tAy[c] = t;
    }
}

/* aprod_impl.c:22-267 (mode 1) and 276-524 (mode 2) with numProc == 1:
 * the Allreduce is a copy and the extended vectors equal the local ones
 * (tridiagonal.c:80-85). */
int orc_aprod(const orc_sys *s, int mode, double *x, double *y_k, double *y_d)
{
    int i;
    if (mode == 1) {
        double *tk = (double *)calloc((size_t)s->m_k + 1, sizeof(double));
        double *td = (double *)calloc((size_t)s->m_d + 1, sizeof(double));
        if (!tk || !td) { free(tk); free(td); return -1; }
        orc_spmv_csc(s->n, s->k_col_ptr, s->k_row_idx, s->k_val, x, tk);   /* :86 */
        for (i = 0; i < s->m_k; i++) y_k[i] += tk[i];                       /* :136-140 */
        orc_spmv_csr_v2(s->m_d, s->d_row_ptr, s->d_col_idx, s->d_val, x, 0, td); /* :221 */
        for (i = 0; i < s->m_d; i++) y_d[i] += td[i];                       /* :222-224 */
        free(tk);
        free(td);
        return 0;
    }
    if (mode == 2) {
        double *tk = (double *)calloc((size_t)s->n + 1, sizeof(double));
        double *td = (double *)calloc((size_t)s->n + 1, sizeof(double));
        if (!tk || !td) { free(tk); free(td); return -1; }
        orc_spmvT_csc_v2(s->n, s->k_col_ptr, s->k_row_idx, s->k_val, y_k, 0, tk);    /* :344/354 */
        orc_spmvT_csc_v2(s->n, s->dc_col_ptr, s->dc_row_idx, s->dc_val, y_d, 0, td); /* :462/484 */
        for (i = 0; i < s->n; i++) x[i] += (tk[i] + td[i]);                          /* :465-467 */
        free(tk);
        free(td);
        return 0;
    }
    return -2;
}

/* mycblas.c:182-222: scaled sum of squares, one division per nonzero */
double orc_dnrm2(int n, const double *v)
{
    double scale = 0.0, ssq = 1.0;
    int i;
    if (n <= 0) return 0.0;
    if (n == 1) return fabs(v[0]);
    for (i = 0; i < n; i++) {
        const double x = v[i];
        if (x != 0.0) {
            const double ax = fabs(x);
            if (scale < ax) {
                ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
                scale = ax;
            } else {
                ssq += (ax / scale) * (ax / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

/* mycblas.c:158-180 at one rank: sqrt of the square of the local norm */
double orc_norm_v(int n, const double *v)
{
    const double part = orc_dnrm2(n, v);
    const double square = part * part;
    return sqrt(square);
}

/* mycblas.c:99-133: ||[y_k; y_d]|| = sqrt(pk^2 + pd^2) */
double orc_norm_u(int m_k, const double *y_k, int m_d, const double *y_d)
{
    const double pk = orc_dnrm2(m_k, y_k);
    const double pd = orc_norm_v(m_d, y_d);
    return sqrt(pk * pk + pd * pd);
}

/* mycblas.c:286-309 */
void orc_dscal(int n, double alpha, double *v)
{
    int i;
    for (i = 0; i < n; i++) v[i] *= alpha;
}

/* cblas.h:53-63 */
double orc_d2norm(double a, double b)
{
    const double scale = fabs(a) + fabs(b);
    if (scale == 0.0) return 0.0;
    return scale * sqrt((a / scale) * (a / scale) + (b / scale) * (b / scale));
}

static void trace_row(double *trace, int row, double alpha, double beta, double rnorm,
                      double test1, double test2, double anorm, double acond, double xnorm)
{
    if (!trace) return;
    double *t = trace + (size_t)row * ORC_TRACE_WIDTH;
    t[0] = alpha; t[1] = beta; t[2] = rnorm; t[3] = test1;
    t[4] = test2; t[5] = anorm; t[6] = acond; t[7] = xnorm;
}

/* lsqr.c:195-917 restated for one rank.  u = [y_k; y_d] lives in the caller's
 * y_k (initially b) and y_d (initially 0), v in the caller's v, as in
 * PLSQR3_main (plsqr3.c:575-603).  The loop runs exactly itn iterations
 * (lsqr.c:833-834) unless tol > 0, in which case it also stops at the first
 * iteration with test2 <= tol (a measurement knob; the reference has none).
 * se may be NULL (the reference passes NULL). */
int orc_lsqr(const orc_sys *s, double *y_k, double *y_d, double *v, double *w, double *x,
             double *se, int itn_max, double damp, double tol, double *trace, orc_out *out)
{
    const double atol = 1e-10, btol = 1e-10, conlim = 100.0;   /* lsqr.c:47-49 */
    const int itnlim = 1000;                                    /* lsqr.c:50 */
    const int n = s->n;
    const int damped = damp > 0.0;
    const int wantse = se != NULL;
    int istop = 0, itn = 0, i, maxdx = 0, nstop = 0;
    double alpha, beta, anorm = 0.0, acond = 0.0, dnorm = 0.0, dxmax = 0.0, res2 = 0.0,
           psi = 0.0, xnorm = 0.0, xnorm1 = 0.0, cs2 = -1.0, sn2 = 0.0, z = 0.0,
           ctol = 0.0, rnorm, bnorm, rhobar, phibar, arnorm, arnorm0,
           test1 = 0.0, test2 = 0.0, test3 = 0.0;
    double cs1, sn1, rhbar1, rho, cs, sn, theta, phi, tau, t1, t2, t3, t, dknorm, dxk,
           delta, gambar, rhs, zbar, gamma, rtol, temp;

    if (conlim > 0.0) ctol = 1.0 / conlim;                     /* :283-284 */
    for (i = 0; i < n; i++) { v[i] = 0.0; x[i] = 0.0; }         /* :311-312 */
    if (wantse) for (i = 0; i < n; i++) se[i] = 0.0;
    alpha = 0.0;

    beta = orc_norm_u(s->m_k, y_k, s->m_d, y_d);                /* :343 */
    if (beta > 0.0) {
        const double ib = 1.0 / beta;
        orc_dscal(s->m_k, ib, y_k);                             /* :361 */
        orc_dscal(s->m_d, ib, y_d);
        if (orc_aprod(s, 2, v, y_k, y_d)) return -1;            /* :366 */
        alpha = orc_norm_v(n, v);                               /* :375 */
    }
    if (alpha > 0.0) {                                          /* :392-395 */
        orc_dscal(n, 1.0 / alpha, v);
        for (i = 0; i < n; i++) w[i] = v[i];
    }
    arnorm = arnorm0 = alpha * beta;                            /* :396 */
    (void)arnorm0;
    if (arnorm == 0.0) {                                        /* :400-401 -> goto_800 */
        trace_row(trace, 0, alpha, beta, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
        if (out) {
            memset(out, 0, sizeof(*out));
            out->istop = 0; out->itn = 0;
        }
        return 0;
    }
    rhobar = alpha;                                             /* :425-428 */
    phibar = beta;
    bnorm = beta;
    rnorm = beta;
    test1 = 1.0;                                                /* :437-438 */
    test2 = alpha / beta;
    trace_row(trace, 0, alpha, beta, rnorm, test1, test2, anorm, acond, xnorm);

    if (itn_max > 0)
        while (1) {
            itn = itn + 1;                                      /* :487 */
            {
                const double na = -alpha;                       /* :502 */
                orc_dscal(s->m_k, na, y_k);
                orc_dscal(s->m_d, na, y_d);
            }
            if (orc_aprod(s, 1, v, y_k, y_d)) return -1;        /* :507 */
            beta = orc_norm_u(s->m_k, y_k, s->m_d, y_d);        /* :516 */
            temp = orc_d2norm(alpha, beta);                     /* :526-528 */
            temp = orc_d2norm(temp, damp);
            anorm = orc_d2norm(anorm, temp);
            if (beta > 0.0) {                                   /* :530 */
                const double ib = 1.0 / beta;
                orc_dscal(s->m_k, ib, y_k);                     /* :538 */
                orc_dscal(s->m_d, ib, y_d);
                orc_dscal(n, -beta, v);                         /* :558 */
                if (orc_aprod(s, 2, v, y_k, y_d)) return -1;    /* :564 */
                alpha = orc_norm_v(n, v);                       /* :577 */
                if (alpha > 0.0) orc_dscal(n, 1.0 / alpha, v);  /* :590-597 */
            }
            rhbar1 = rhobar;                                    /* :609-616 */
            if (damped) {
                rhbar1 = orc_d2norm(rhobar, damp);
                cs1 = rhobar / rhbar1;
                sn1 = damp / rhbar1;
                psi = sn1 * phibar;
                phibar = cs1 * phibar;
            }
            rho = orc_d2norm(rhbar1, beta);                     /* :622-629 */
            cs = rhbar1 / rho;
            sn = beta / rho;
            theta = sn * alpha;
            rhobar = -cs * alpha;
            phi = cs * phibar;
            phibar = sn * phibar;
            tau = sn * phi;
            t1 = phi / rho;                                     /* :634-637 */
            t2 = -theta / rho;
            t3 = 1.0 / rho;
            dknorm = 0.0;
            if (wantse) {                                       /* :662-669 */
                for (i = 0; i < n; i++) {
                    t = w[i];
                    x[i] = t1 * t + x[i];
                    w[i] = t2 * t + v[i];
                    t = (t3 * t) * (t3 * t);
                    se[i] = t + se[i];
                    dknorm = t + dknorm;
                }
            } else {                                            /* :685-690 */
                for (i = 0; i < n; i++) {
                    t = w[i];
                    x[i] = t1 * t + x[i];
                    w[i] = t2 * t + v[i];
                    dknorm = (t3 * t) * (t3 * t) + dknorm;
                }
            }
            dknorm = sqrt(dknorm);                              /* :703-709 */
            dnorm = orc_d2norm(dnorm, dknorm);
            dxk = fabs(phi * dknorm);
            if (dxmax < dxk) { dxmax = dxk; maxdx = itn; }
            delta = sn2 * rho;                                  /* :716-725 */
            gambar = -cs2 * rho;
            rhs = phi - delta * z;
            zbar = rhs / gambar;
            xnorm = orc_d2norm(xnorm1, zbar);
            gamma = orc_d2norm(gambar, theta);
            cs2 = gambar / gamma;
            sn2 = theta / gamma;
            z = rhs / gamma;
            xnorm1 = orc_d2norm(xnorm1, z);
            acond = anorm * dnorm;                              /* :732-748 */
            res2 = orc_d2norm(res2, psi);
            rnorm = orc_d2norm(res2, phibar);
            arnorm = alpha * fabs(tau);
            test1 = rnorm / bnorm;
            test2 = 0.0;
            if (rnorm > 0.0) test2 = arnorm / (anorm * rnorm);
            test3 = 1.0 / acond;
            t1 = test1 / (1.0 + anorm * xnorm / bnorm);
            rtol = btol + atol * anorm * xnorm / bnorm;
            t3 = 1.0 + test3;                                   /* :756-775 */
            t2 = 1.0 + test2;
            t1 = 1.0 + t1;
            if (itn >= itnlim) istop = 5;
            if (t3 <= 1.0) istop = 4;
            if (t2 <= 1.0) istop = 2;
            if (t1 <= 1.0) istop = 1;
            if (test3 <= ctol) istop = 4;
            if (test2 <= atol) istop = 2;
            if (test1 <= rtol) istop = 1;
            if (istop == 0) nstop = 0;                          /* :820-827 */
            else { nstop = nstop + 1; if (nstop < 1 && itn < itnlim) istop = 0; }
            trace_row(trace, itn, alpha, beta, rnorm, test1, test2, anorm, acond, xnorm);
            if (out) out->arnorm_true = arnorm;
            if (itn == itn_max) break;                          /* :833-834 */
            if (tol > 0.0 && test2 <= tol) break;
        }
    if (wantse) {                                               /* :878-891 */
        int m = s->m_k + s->m_d;
        t = 1.0;
        if (m > n) t = m - n;
        if (damped) t = m;
        t = rnorm / sqrt(t);
        for (i = 0; i < n; i++) se[i] = t * sqrt(se[i]);
    }
    if (damped && istop == 2) istop = 3;                        /* :895-896 */
    if (out) {                                                  /* :908-914 */
        out->istop = istop;
        out->itn = itn;
        out->anorm = anorm;
        out->acond = acond;
        out->rnorm = rnorm;
        out->arnorm = test2;
        out->xnorm = xnorm;
        out->bnorm = bnorm;
        out->dxmax = dxmax;
        out->maxdx = maxdx;
    }
    return 0;
}
