/*
 * plsqr_oracle.h -- CPU restatement of PLSQR3's LSQR hot path.
 * TEST INFRASTRUCTURE ONLY (the checker, never the product); see
 * plsqr_oracle.c for the reference citations and what pins each function.
 */
#ifndef PLSQR_ORACLE_H_
#define PLSQR_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys {
    int m_k, m_d, n;
    const int *k_col_ptr, *k_row_idx;
    const double *k_val;
    const int *d_row_ptr, *d_col_idx;
    const double *d_val;
    const int *dc_col_ptr, *dc_row_idx;
    const double *dc_val;
} orc_sys;

typedef struct orc_out {
    int istop, itn;
    double anorm, acond, rnorm, arnorm, xnorm, bnorm, dxmax;
    int maxdx;
    double arnorm_true;
} orc_out;

#define ORC_TRACE_WIDTH 8 /* alpha, beta, rnorm, test1, test2, anorm, acond, xnorm */

void orc_spmv_csc(int n, const int *col_ptr, const int *row_idx, const double *val,
                  const double *x, double *Ax);
void orc_spmv_csr_v2(int m, const int *row_ptr, const int *col_idx, const double *val,
                     const double *x, int start, double *Ax);
void orc_spmvT_csc_v2(int n, const int *col_ptr, const int *row_idx, const double *val,
                      const double *y, int start, double *tAy);
int orc_aprod(const orc_sys *s, int mode, double *x, double *y_k, double *y_d);
double orc_dnrm2(int n, const double *v);
double orc_norm_v(int n, const double *v);
double orc_norm_u(int m_k, const double *y_k, int m_d, const double *y_d);
void orc_dscal(int n, double alpha, double *v);
double orc_d2norm(double a, double b);
int orc_lsqr(const orc_sys *s, double *y_k, double *y_d, double *v, double *w, double *x,
             double *se, int itn_max, double damp, double tol, double *trace, orc_out *out);

#ifdef __cplusplus
}
#endif
#endif
