/*
 * aprod_plsqr3.c -- PLSQR3's aprod callback on libplsqr_hip.so.
 *
 * A maintainer builds PLSQR3 (the C files of source/) with this file in place of
 * source/aprod.c and links libplsqr_hip.so: lsqr() (lsqr.c:195-917) keeps
 * its host loop, norms and scalings, and every operator application it
 * requests -- lsqr.c:366 (mode 2, start), :507 (mode 1) and :564 (mode 2)
 * -- runs on the GPU.  The signature is the reference's (aprod.h:10-13,
 * callback type lsqr.h:38-40).  As in source/aprod.c:24-84 the x and y
 * arguments are ignored: v = p->vec_x and u = [p->vec_y_k ; p->vec_y_d]
 * live in the PPDataV3_TRI work struct (main.h:397-432).
 *
 * The device operator is created on the first call from the matrices
 * PLSQR3 loaded (kernel CSC, damping CSR and damping CSC by column,
 * main.h:400-414) and destroyed after the last mode-2 call
 * (isFinalIteration, as aprod_impl.c:501-523 frees its scratch).
 * PLSQR_HIP_OPERATOR=exact selects PLSQR_OP_EXACT_ORDER (every row and
 * column summed in the reference's order: x bit for bit the reference's);
 * unset or "default" the fast operator.
 *
 * Requirements: the FP64 build (main.h:26 SDVERSION 2, MY_FLOAT_TYPE
 * double) and one MPI rank (the whole system on one GPU).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "main.h"
#include "aprod.h"
#include "plsqr_hip.h"

typedef char plsqr_hip_needs_the_fp64_build[sizeof(MY_FLOAT_TYPE) == sizeof(double) ? 1 : -1];

static plsqr_op *g_op;

static void plsqr_hip_fail(const char *what)
{
    fprintf(stderr, "aprod (libplsqr_hip): %s: %s\n", what,
            g_op ? plsqr_last_error(g_op) : plsqr_last_global_error());
    exit(-1);
}

static void plsqr_hip_init(int numProc, PPDataV3_TRI *p)
{
    if (numProc != 1) {
        fprintf(stderr, "aprod (libplsqr_hip): one MPI rank holds the whole system; got %d\n", numProc);
        exit(-1);
    }
    plsqr_system s;
    memset(&s, 0, sizeof(s));
    s.m_k = p->matCSC_k->numRow;
    s.m_d = p->vec_y_d_len;
    s.n = p->vec_x_len;
    s.nnz_k = p->matCSC_k->numNonzero;
    s.nnz_d = p->matCSR_d->numNonzero;
    s.k_col_ptr = p->matCSC_k->col_ptr;
    s.k_row_idx = p->matCSC_k->row_idx;
    s.k_val = p->matCSC_k->val;
    s.d_row_ptr = p->matCSR_d->row_ptr;
    s.d_col_idx = p->matCSR_d->col_idx;
    s.d_val = p->matCSR_d->val;
    s.dc_col_ptr = p->matCSC_d_bycol->col_ptr;
    s.dc_row_idx = p->matCSC_d_bycol->row_idx;
    s.dc_val = p->matCSC_d_bycol->val;
    const char *kind = getenv("PLSQR_HIP_OPERATOR");
    const int flags = (kind && strcmp(kind, "exact") == 0) ? PLSQR_OP_EXACT_ORDER : PLSQR_OP_DEFAULT;
    if (plsqr_op_create(&s, flags, &g_op) != PLSQR_OK) plsqr_hip_fail("plsqr_op_create");
}

void aprod(int version, int procRank, int numProc, int mode, int m, int n, MY_FLOAT_TYPE x[],
           MY_FLOAT_TYPE y[], void *UsrWork, int isFinalIteration, int iteration)
{
    PPDataV3_TRI *p = (PPDataV3_TRI *)UsrWork;
    (void)version; (void)procRank; (void)m; (void)n; (void)x; (void)y; (void)iteration;
    if (!g_op) plsqr_hip_init(numProc, p);
    if (plsqr_aprod_split(g_op, mode, p->vec_x, p->vec_y_k, p->vec_y_d) != PLSQR_OK)
        plsqr_hip_fail("plsqr_aprod_split");
    if (isFinalIteration && mode == 2) {
        plsqr_op_destroy(g_op);
        g_op = NULL;
    }
}
