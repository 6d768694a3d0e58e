/*
 * plsqr_hip.h -- C-ABI of the MI355X (gfx950) LSQR hot path.
 *
 * The hot path of PLSQR3 is the LSQR bidiagonalisation loop (lsqr.c:195-917)
 * and its operator callback aprod (aprod.c:24-84 -> aprod_impl.c:22-524),
 * which computes  u += A v  (mode 1) and  v += A^T u  (mode 2) for
 * A = [A_k ; D]: a column-sorted kernel matrix A_k (m_k x n) stacked on a
 * damping matrix D (m_d x n).  This header is the drop-in boundary: plain C
 * types, host or device pointers and sizes, no C++ or torch types.
 *
 * Conventions
 *   - Matrices enter exactly as the reference loaders build them: 0-based
 *     int32 pointers and indices, FP64 values (plsqr3.c:1515-1626 for the
 *     kernel CSC, plsqr3.c:1133-1189 for the damping CSR and CSC).
 *   - u = [y_k ; y_d] is one vector of m_k + m_d doubles, kernel rows first
 *     (the split of PPDataV3_TRI, main.h:421-432).
 *   - Functions returning int return PLSQR_OK (0) or a PLSQR_ERR_* code and
 *     never call exit(); the message is kept on the handle
 *     (plsqr_last_error) or, for calls without a handle, in thread-local
 *     storage (plsqr_last_global_error).
 *   - A handle is bound to one device and one HIP stream.  It is
 *     thread-compatible, not thread-safe: one host thread per handle.
 */
#ifndef PLSQR_HIP_H_
#define PLSQR_HIP_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status codes                                                            */
/* ---------------------------------------------------------------------- */
enum {
    PLSQR_OK = 0,
    PLSQR_ERR_ARG = 1,     /* invalid argument or size mismatch            */
    PLSQR_ERR_BOUNDS = 2,  /* index out of range in matrix data            */
    PLSQR_ERR_IO = 3,      /* file open/read/write failure                 */
    PLSQR_ERR_HIP = 4,     /* HIP runtime failure (no device, launch, ...) */
    PLSQR_ERR_NOMEM = 5,   /* host or device allocation failed             */
    PLSQR_ERR_STATE = 6,   /* call not valid in the handle's current state */
    PLSQR_ERR_FORMAT = 7,  /* malformed file contents                      */
    PLSQR_ERR_NONFINITE = 8/* NaN or Inf in the input vectors or matrix    */
};

/* ---------------------------------------------------------------------- */
/* Host-side system in the reference's loaded layout                      */
/* ---------------------------------------------------------------------- */

/* Kernel CSC + damping CSR + damping CSC-by-column + kernel right-hand side.
 * Replaces the three SpMatCSC/SpMatCSR objects the reference driver builds
 * before initPPDataV3_TRI (plsqr3.c:320-335, 420-529; main.h:123-165) and the
 * b vector it reads (plsqr3.c:580-589).  Arrays owned by the library when the
 * struct was filled by plsqr_gen_system / plsqr_read_reference_files (free
 * with plsqr_system_free); arrays supplied by the caller stay the caller's. */
typedef struct plsqr_system {
    int m_k;            /* kernel rows  (-row_k)                               */
    int m_d;            /* damping rows (-row_d)                               */
    int n;              /* columns      (-col)                                 */
    int owned;          /* 1 if the arrays below were allocated by the library */
    long long nnz_k;
    long long nnz_d;
    int    *k_col_ptr;  /* n+1                                                 */
    int    *k_row_idx;  /* nnz_k, 0-based kernel row, ascending per column     */
    double *k_val;      /* nnz_k                                               */
    int    *d_row_ptr;  /* m_d+1                                               */
    int    *d_col_idx;  /* nnz_d, 0-based column, ascending per row            */
    double *d_val;      /* nnz_d                                               */
    int    *dc_col_ptr; /* n+1                                                 */
    int    *dc_row_idx; /* nnz_d, 0-based damping row, ascending per column    */
    double *dc_val;     /* nnz_d                                               */
    double *b;          /* m_k: kernel right-hand side (damping rhs is zero)   */
    double *x_true;     /* n: model used to synthesise b, or NULL              */
} plsqr_system;

void plsqr_system_free(plsqr_system *sys);

/* ---------------------------------------------------------------------- */
/* Synthetic straight-ray tomography generator                             */
/* ---------------------------------------------------------------------- */
/* Replaces the offline chain ker_ascii2bin -> Ker2PLSQR3 -> damping_binary.py
 * (PLSQR3_tools/kernel_format/Ker2PLSQR3.c:358-415,595-606;
 *  PLSQR3_tools/damping_format/damping_binary.py:84-289) for seeded
 * synthetic systems.  Columns are iz*ny*nx + iy*nx + ix (damping_binary.py:54).
 * Damping rows per cell, x fastest: identity*lambda, then the 1,-2,1 second
 * difference along x (if ix<nx-2), y (if iy<ny-2), z (if iz<nz-2), each
 * scaled by its weight; no cross terms.  All damping values are exact FP64
 * in both the row and the column layout. */
enum {
    PLSQR_GEOM_OPPOSITE_FACES = 0, /* axis drawn per ray; endpoints on the two faces normal to it */
    PLSQR_GEOM_TOP_SOURCES = 1     /* source on the z=nz face; receiver on one of the 5 other faces */
};

typedef struct plsqr_gen_params {
    int nx, ny, nz;
    long long n_rays;
    int geometry;                 /* PLSQR_GEOM_*                                  */
    unsigned long long seed;
    double sigma;                 /* std of the Gaussian noise added to b          */
    double lambda;                /* identity damping weight                       */
    double w_xx, w_yy, w_zz;      /* second-difference weights                     */
    int n_blobs;                  /* Gaussian blobs in x_true (5)                  */
    double blob_width;            /* blob standard deviation in voxels (3)         */
} plsqr_gen_params;

int plsqr_gen_system(const plsqr_gen_params *p, plsqr_system *out);

/* New right-hand side for an existing matrix: b = A_k x_true + N(0, sigma^2)
 * with x_true drawn from p->seed, p->n_blobs, p->blob_width and noise from
 * p->seed, p->sigma (p->nx*p->ny*p->nz must equal sys->n).  b and x_true must
 * hold m_k and n doubles. */
int plsqr_gen_rhs(const plsqr_system *sys, const plsqr_gen_params *p,
                  double *b, double *x_true);

/* Thread-local message of the last failing call that has no handle. */
const char *plsqr_last_global_error(void);

/* ---------------------------------------------------------------------- */
/* Reference on-disk formats                                               */
/* ---------------------------------------------------------------------- */
/* Files, all written 1-based like the reference tools:
 *   kernel  .bin  16-byte records {int32 row, int32 col, float64 val},
 *                 column-major, rows ascending      (Ker2PLSQR3.c:595-606)
 *   kernel  .info ASCII "col nnz start" per column; start is the 1-based
 *                 record offset, 0 when nnz==0      (Ker2PLSQR3.c:505-511)
 *   damping row data/info, col data/info            (damping_binary.py:97-101,523-551)
 *   b       ASCII, one value per line ("%.17e")     (read at plsqr3.c:580-589)
 * Any name may be NULL to skip that file when writing. */
int plsqr_write_reference_files(const plsqr_system *sys, const char *dir,
                                const char *ker_f, const char *ker_i,
                                const char *damp_f, const char *damp_i,
                                const char *damp_f_bycol, const char *damp_i_bycol,
                                const char *b_k);

/* Loads the same seven files the reference's CLI names (plsqr3.c:123-186:
 * -dir -ker_f -ker_i -damp_f -damp_i -damp_f_bycol -damp_i_bycol -b_k
 * -row_k -row_d -col) into a library-owned plsqr_system, with the reference
 * loaders' bounds checks (plsqr3.c:1149-1153,1178-1182,1606-1615) reported as
 * PLSQR_ERR_BOUNDS instead of exit(-1). */
int plsqr_read_reference_files(const char *dir,
                               const char *ker_f, const char *ker_i,
                               const char *damp_f, const char *damp_i,
                               const char *damp_f_bycol, const char *damp_i_bycol,
                               const char *b_k, int row_k, int row_d, int col,
                               plsqr_system *out);

/* Kernel re-layout (PLSQR3_tools/kernel_format/Ker2PLSQR3.c:358-415 reads
 * the rows in order and hands every entry to its column; :595-606 writes
 * the column-major records, rows ascending within a column), done on the
 * current HIP device as a stable radix sort of the entries by column.
 *   in : m_k rows, row_ptr (m_k+1, from 0), col1 (1-based columns, any order
 *        within a row), val;
 *   out: col_ptr (n+1), row_idx (0-based, ascending within a column), val_out
 *        (nnz each): the kernel CSC of plsqr_system.
 * Columns outside 1..n -> PLSQR_ERR_BOUNDS; NaN/Inf -> PLSQR_ERR_NONFINITE. */
int plsqr_kernel_rows_to_csc(int m_k, int n, const int *row_ptr, const int *col1, const double *val,
                             int *col_ptr, int *row_idx, double *val_out);

/* The tool itself: reads the kernel rows named in dir/list_file (one "file
 * nnz" pair per line, FileIO.c:353-402; file r holds row r as nnz 16-byte
 * records {int32 aux, int32 col (1-based), float64 val}), re-lays them with
 * plsqr_kernel_rows_to_csc and writes dir/out_bin and dir/out_info in the
 * formats above.  n_col: the number of columns (Ker2PLSQR3's argv[3]; the
 * per-column counts of its argv[4] are computed instead of read). */
int plsqr_ker2plsqr3(const char *dir, const char *list_file, int n_col, const char *out_bin,
                     const char *out_info);

/* Raw float64[n] in global column order, the layout of mpiio_write_result
 * (plsqr3.c:1986-2013) for the FP64 build. */
int plsqr_write_x(const char *path, const double *x, int n);
int plsqr_read_x(const char *path, double *x, int n);

/* The reference's ASCII-to-binary kernel step (PLSQR3_tools/kernel_format/
 * ker_ascii2bin.c:18-94).  list_path names the row files: first token their
 * count, then one name per kernel row (:41-42,58).  Each in_dir/<name> holds
 * entries "col ix iy iz value" (:76-77) and becomes out_dir/<name>.bin, one
 * 16-byte record {int32 ix, int32 col, float64 value} per entry (:81-83),
 * the row-file format plsqr_ker2plsqr3 reads.  out_list (may be NULL) is
 * written into out_dir with one "<name>.bin nnz" line per row: the list
 * plsqr_ker2plsqr3 takes (the example dataset ships it by hand,
 * data/ker_bin.list).  An entry that does not parse as four integers and a
 * number -> PLSQR_ERR_FORMAT (the tool itself would loop on it). */
int plsqr_ker_ascii2bin(const char *list_path, const char *in_dir, const char *out_dir, const char *out_list);

/* Damping matrix of the reference tool PLSQR3_tools/damping_format/
 * damping_binary.py (its argv[4..14], :36-47).  n_param parameter blocks
 * follow one another (block p owns columns p*nx*ny*nz .. (p+1)*nx*ny*nz-1,
 * :93-95); in each block every cell, x fastest (:89-92), gets
 *   identity                           w[0]                       (:93-113)
 *   xx if ix < nx-2                    (1,-2,1)*w[1] at ix..ix+2  (:115-178)
 *   yy if iy < ny-2, zz if iz < nz-2   the same along y (w[2]) and z (w[3])
 *   with cross != 0 also
 *   xy if ix < nx-2 and iy < ny-2      (0.5,-0.5,-0.5,0.5)*w[4] at (ix,iy),
 *                                      (ix+2,iy), (ix,iy+2), (ix+2,iy+2)
 *   xz, yz                             the same in those planes (w[5], w[6])
 *                                                                 (:291-479)
 * The reference tool always writes the cross rows; the synthetic workloads
 * have none.  All values are FP64 in both layouts (the tool narrows its
 * column copy to FP32, damping_binary.py:70; identical whenever FP32 holds
 * the weights exactly, as for 1.0). */
typedef struct plsqr_damping_spec {
    int nx, ny, nz;   /* grid of one parameter block */
    int n_param;      /* parameter blocks (argv[7])  */
    double w[7];      /* identity, xx, yy, zz, xy, xz, yz weights (argv[8..14]) */
    int cross;        /* 1: with the xy, xz, yz rows */
} plsqr_damping_spec;

/* Fills a library-owned plsqr_system with the damping part only (m_k = 0,
 * kernel arrays and b NULL; n = nx*ny*nz*n_param). */
int plsqr_gen_damping(const plsqr_damping_spec *spec, plsqr_system *out);

/* The four damping files of the tool (row data/info, column data/info;
 * 1-based rows and columns, damping-local rows), as plsqr_write_reference_files
 * writes them.  Any name may be NULL to skip that file. */
int plsqr_write_damping_files(const plsqr_damping_spec *spec, const char *dir, const char *damp_f,
                              const char *damp_i, const char *damp_f_bycol, const char *damp_i_bycol);

/* ---------------------------------------------------------------------- */
/* Device operator                                                         */
/* ---------------------------------------------------------------------- */
typedef struct plsqr_op plsqr_op;

enum {
    PLSQR_OP_DEFAULT = 0,      /* FP64 operator, wavefront-parallel deterministic sums    */
    PLSQR_OP_EXACT_ORDER = 1,  /* FP64, every row/column summed sequentially in the
                                  reference's order (bit-identical aprod)                */
    PLSQR_OP_WITH_FP32 = 2     /* also keep FP32 copies of the values (mixed solves)     */
};

/* Uploads the system to the current HIP device, re-lays the kernel CSC as
 * CSR on the device (stable counting sort, bit-exact triplets) and allocates
 * the resident u (m_k+m_d) and v (n) vectors, both zero like
 * initPPDataV3_TRI (tridiagonal.c:33-57).  Replaces the GPU set-up the
 * reference calls InitPPDataV3_TRI_GPU (lsqr.c:451-457). */
int  plsqr_op_create(const plsqr_system *sys, int flags, plsqr_op **out);

/* Grid the columns come from (column = iz*ny*nx + iy*nx + ix).  With it the
 * default operator renumbers the columns into bricks of 4 x 4 x 8 voxels so
 * that one ray crossing a brick is gathered once for all the brick's voxels
 * in x += A^T y.  All vectors crossing this API stay in the caller's order. */
typedef struct plsqr_grid { int nx, ny, nz; } plsqr_grid;
int  plsqr_op_create_ex(const plsqr_system *sys, int flags, const plsqr_grid *grid,
                        plsqr_op **out);
void plsqr_op_destroy(plsqr_op *op);
const char *plsqr_last_error(const plsqr_op *op);
void plsqr_clear_error(plsqr_op *op);
int  plsqr_op_dims(const plsqr_op *op, int *m_k, int *m_d, int *n,
                   long long *nnz_k, long long *nnz_d);
/* Device memory held by the handle, in bytes. */
size_t plsqr_op_device_bytes(const plsqr_op *op);
/* Attach a HIP stream (hipStream_t as void*; NULL = the default stream). */
int  plsqr_op_set_stream(plsqr_op *op, void *stream);
/* The stream the handle enqueues on (hipStream_t as void*), e.g. for events. */
void *plsqr_op_get_stream(const plsqr_op *op);

/* Resident vectors used when aprod is called with NULL x/y (PLSQR3's
 * convention, lsqr.c:366,507,564).  NULL arguments are skipped. */
int plsqr_op_set_vectors(plsqr_op *op, const double *u_host, const double *v_host);
int plsqr_op_get_vectors(plsqr_op *op, double *u_host, double *v_host);

/* ---------------------------------------------------------------------- */
/* aprod: the callback boundary                                            */
/* ---------------------------------------------------------------------- */
/* Exact signature of the reference callback (lsqr.h:38-40, aprod.h:10-13),
 * with MY_FLOAT_TYPE = double.  UsrWrk is a plsqr_op*.
 *   mode 1:  y <- y + A x     mode 2:  x <- x + A^T y
 * x (n) and y (m = m_k + m_d) are host pointers; when NULL the handle's
 * resident v and u are used instead.  version, procRank, numProc,
 * isFinalIteration and iteration are accepted and ignored: scratch memory
 * belongs to the handle, so there is no final-iteration free
 * (aprod_impl.c:237-261,501-523) and a second solve is safe.  m and n are
 * checked against the handle; errors are recorded on the handle. */
void plsqr_aprod_ref(int version, int procRank, int numProc, int mode,
                     int m, int n, double x[], double y[], void *UsrWrk,
                     int isFinalIteration, int iteration);

/* Classic Paige-Saunders callback shape (clsqr2/lsqr_s.h:15-16,
 * clsqr2/aprod_s.h:12-13): host x and y, both required. */
void plsqr_aprod(int mode, int m, int n, double x[], double y[], void *UsrWrk);

/* PLSQR3's split u: x = vec_x (n), y_k = vec_y_k (m_k), y_d = vec_y_d (m_d),
 * host pointers (main.h:421-432).  This is the call a PLSQR3 aprod.c body
 * maps onto. */
int plsqr_aprod_split(plsqr_op *op, int mode, double *x, double *y_k, double *y_d);

/* Device pointers (x: n doubles, y: m_k+m_d doubles), enqueued on the
 * handle's stream; returns without synchronising. */
int plsqr_aprod_device(plsqr_op *op, int mode, double *x_dev, double *y_dev);

/* The same operator in the mixed-precision solver's arithmetic: FP32 matrix
 * values and vectors, every row / column dot product accumulated in FP64 from
 * the exact products and rounded to FP32 once (aprod.c:24-84 in FP32; no counterpart in the
 * reference, which is FP64 only).  mode 1: y += A x; mode 2: x += A^T y; x: n
 * floats, y: m_k+m_d floats, caller order.  The _device form takes device
 * pointers and returns without synchronising; the host form copies in and
 * out.  PLSQR_ERR_STATE on an exact-order handle. */
int plsqr_aprod_f32(plsqr_op *op, int mode, float *x, float *y);
int plsqr_aprod_device_f32(plsqr_op *op, int mode, float *x_dev, float *y_dev);

/* The handle's own vector order.  The default operator renumbers the columns
 * (voxel bricks) and the kernel rows (ray order), and the caller-order entry
 * points above convert x and y at every call, as the reference's aprod
 * gathers x into its compressed local numbering at every call
 * (aprod_impl.c:145-175, vector_comm.c GatherVecXCompressed).  A caller that
 * keeps its device vectors in the handle's order converts once:
 *   space 0: column space (n doubles), 1: row space (m_k+m_d doubles);
 *   the two pointers must differ (PLSQR_ERR_ARG otherwise);
 * and then calls the _internal aprod, which launches the SpMV kernels only.
 * All enqueue on the handle's stream and return without synchronising. */
int plsqr_op_to_internal(plsqr_op *op, int space, const double *ext_dev, double *int_dev);
int plsqr_op_to_external(plsqr_op *op, int space, const double *int_dev, double *ext_dev);
int plsqr_aprod_device_internal(plsqr_op *op, int mode, double *x_dev, double *y_dev);
int plsqr_aprod_device_internal_f32(plsqr_op *op, int mode, float *x_dev, float *y_dev);

/* ---------------------------------------------------------------------- */
/* Device-resident LSQR                                                    */
/* ---------------------------------------------------------------------- */
enum {
    PLSQR_PREC_FP64 = 0,   /* plain FP64, the reference's arithmetic order          */
    PLSQR_PREC_MIXED = 1   /* FP32 values and u, v with FP64-accumulated dot products,
                              FP64 scalars and an FP64 residual refinement          */
};

/* Preconditioning (no counterpart in the reference, whose lsqr() always runs
 * on A itself).  A column-scaled solve runs its inner LSQR on A D in the
 * solve's precision (FP32 operator for PLSQR_PREC_MIXED, FP64 operator for
 * PLSQR_PREC_FP64) and checks test2 on A in FP64 every refine_every inner
 * iterations (refine_adaptive as for the mixed solve); no trace, no
 * standard errors. */
enum {
    PLSQR_PRECOND_NONE = 0,      /* LSQR on A                                          */
    PLSQR_PRECOND_COLSCALE = 1   /* LSQR on A D with D = diag(||A e_j||^-p), p =
                                    precond_power (1: columns of [A_k; damping] scaled
                                    to unit 2-norm; zero columns keep 1), x = D y; the
                                    FP64 checks of test2 stay on A                     */
};

typedef struct plsqr_solve_opts {
    int    itn;            /* iteration count (-itn); the loop runs exactly itn
                              iterations when tol == 0, like lsqr.c:833-834      */
    double tol;            /* > 0: stop at the first iteration with
                              test2 <= tol (test2 as lsqr.c:742-744)            */
    int    precision;      /* PLSQR_PREC_*                                       */
    int    refine_every;   /* mixed only: inner iterations per FP64 refinement   */
    double damp;           /* LSQR damp parameter (the reference passes 0.0)     */
    double atol, btol, conlim; /* istop tests (lsqr.c:759-775); reference values
                              1e-10, 1e-10, 100 when all three are 0            */
    int    itnlim;         /* istop=5 threshold; 1000 (lsqr.c:50) when 0          */
    int    precond;        /* PLSQR_PRECOND_*                                     */
    double anorm_ref;      /* with PLSQR_PRECOND_COLSCALE: the ||A|| of the
                              FP64 test2 = ||A^T r|| / (anorm_ref ||r||); 0 = ||A||_F.
                              Passing the anorm a plain FP64 solve reported makes
                              the stop exactly as strict as that solve's         */
    int    refine_adaptive; /* mixed with tol > 0: 1 = after the first check
                              (refine_every inner iterations), check when the inner
                              estimate of test2 times the last measured ratio
                              FP64 test2 / inner estimate reaches tol, at most
                              4 * refine_every apart; 0 = every refine_every   */
    double precond_power;  /* PLSQR_PRECOND_COLSCALE: D = diag(||A e_j||^-precond_power);
                              0 means 1 (unit column norms)                     */
} plsqr_solve_opts;

typedef struct plsqr_solve_out {
    int    istop, itn;     /* istop as computed (the reference never acts on it) */
    double anorm, acond, rnorm, arnorm, xnorm, bnorm;
                           /* arnorm holds test2, as lsqr.c:913 assigns         */
    double arnorm_true;    /* alpha*|tau| of the last iteration                  */
    double elapsed_ms;     /* device time of the iteration loop                  */
    int    refinements;    /* mixed: FP64 residual refinements performed         */
    int    inner_total;    /* mixed: total inner iterations                      */
} plsqr_solve_out;

/* Per-iteration trace row: alpha, beta, rnorm, test1, test2, anorm, acond, xnorm */
#define PLSQR_TRACE_WIDTH 8

/* Solve min ||A x - [b;0]|| from x = 0 (lsqr.c:311-312).  b_host: m_k
 * doubles.  x_host (n) receives x; w_host (n) may be NULL.  trace may be
 * NULL or hold (itn+1) * PLSQR_TRACE_WIDTH doubles (row 0 = initial state,
 * lsqr.c:425-445).  The resident u and v are overwritten. */
int plsqr_lsqr_solve(plsqr_op *op, const double *b_host, const plsqr_solve_opts *opts,
                     double *x_host, double *w_host, plsqr_solve_out *out,
                     double *trace);

/* Same on device pointers (b_dev: m_k, x_dev and w_dev: n; w_dev may be
 * NULL), enqueued on the handle's stream; the call synchronises that
 * stream before it returns so that *out is complete. */
int plsqr_lsqr_solve_device(plsqr_op *op, const double *b_dev, const plsqr_solve_opts *opts,
                            double *x_dev, double *w_dev, plsqr_solve_out *out,
                            double *trace);

/* Plain FP64 solve as plsqr_lsqr_solve that also returns the standard-error
 * estimates of the reference's wantse branch (lsqr.c:644-670 accumulate
 * var += (w/rho)^2 per iteration; lsqr.c:878-891 scale se = rnorm / sqrt(t)
 * * sqrt(var) with t = m - n when m > n, m when damp > 0, else 1; m = m_k +
 * m_d).  The reference itself always passes se = NULL.  x_host and se_host:
 * n doubles, caller's column order.  PLSQR_ERR_ARG for the mixed precision. */
int plsqr_lsqr_solve_se(plsqr_op *op, const double *b_host, const plsqr_solve_opts *opts,
                        double *x_host, double *se_host, plsqr_solve_out *out);

/* Residual and normal-equation residual in FP64 on the device:
 * r = [b;0] - A x,  rnorm = ||r||,  arnorm = ||A^T r||,  anorm_f = ||A||_F.
 * Used to confirm a test2 reading with ||A^T r|| / (||A||_F ||r||). */
int plsqr_residual_norms(plsqr_op *op, const double *b_host, const double *x_host,
                         double *rnorm, double *arnorm, double *anorm_f);

/* Warm start: solve from x0 (host, n doubles, caller's column order; NULL =
 * zero start).  x0 is kept only when ||[b;0] - A x0|| <= ||[b;0]||, the
 * residual of the zero start; otherwise the solve starts from 0.  *used_x0
 * (may be NULL) reports which.  FP64: LSQR on r0 = [b;0] - A x0, x = x0 +
 * its solution (test2 is that of the full problem, since the residuals are
 * the same); MIXED: the FP64 refinement loop starts from x0.  out->bnorm is
 * then ||r0||.  No counterpart in the reference (lsqr.c:311-312 always
 * starts from zero); used with the warm-start network below. */
int plsqr_lsqr_solve_x0(plsqr_op *op, const double *b_host, const double *x0_host,
                        const plsqr_solve_opts *opts, double *x_host, plsqr_solve_out *out,
                        int *used_x0);
int plsqr_lsqr_solve_x0_device(plsqr_op *op, const double *b_dev, const double *x0_dev,
                               const plsqr_solve_opts *opts, double *x_dev, plsqr_solve_out *out,
                               int *used_x0);

/* x_true field of the generator alone (the blob sum of plsqr_gen_system for
 * p's grid, seed, n_blobs and blob_width): nx*ny*nz doubles. */
int plsqr_gen_field(const plsqr_gen_params *p, double *x_true);

/* ---------------------------------------------------------------------- */
/* Warm-start network (no counterpart in the reference)                    */
/* ---------------------------------------------------------------------- */
/* A 3-D convolutional network on the nx*ny*nz voxel grid maps the
 * back-projection A^T b to a starting point x0: 4 layers of 3x3x3
 * convolutions with zero padding, 32 channels, ReLU between layers, bf16
 * weights and activations with FP32 accumulation on the matrix cores, FP32
 * master weights trained with Adam on the GPU.  Each sample is scaled by the
 * RMS s of its A^T b (input A^T b / s, prediction s * output), so the map
 * commutes with scaling b.  Parameters (plsqr_cnn_param_count() floats):
 * for each layer l = 0..3, weights w[t][n][k] (tap t = (dz+1)*9 + (dy+1)*3
 * + (dx+1) reading voxel (x+dx, y+dy, z+dz), output channel n, input
 * channel k) then bias[n]; the input has one real channel (k = 0) and the
 * output one (n = 0), the other entries of the first and last layer are 0.
 * The last parameter is the output gain g (not trained; plsqr_cnn_set_data
 * sets it to the RMS of the scaled targets xstar / s): x0 = s * g * output. */
typedef struct plsqr_cnn plsqr_cnn;
int  plsqr_cnn_param_count(void);
int  plsqr_cnn_create(int nx, int ny, int nz, int max_batch, unsigned long long seed,
                      plsqr_cnn **out);
void plsqr_cnn_destroy(plsqr_cnn *h);
/* last error of h; with h == NULL the last failure of a call without one */
const char *plsqr_cnn_last_error(const plsqr_cnn *h);
size_t plsqr_cnn_device_bytes(const plsqr_cnn *h);
int  plsqr_cnn_get_params(plsqr_cnn *h, float *params);
int  plsqr_cnn_set_params(plsqr_cnn *h, const float *params);
/* Training pairs (host, count * nx*ny*nz doubles each): A^T b and the LSQR
 * solution x*; kept on the device, scaled per sample. */
int  plsqr_cnn_set_data(plsqr_cnn *h, int count, const double *atb, const double *xstar);
/* steps Adam steps on minibatches of `batch` pairs drawn without
 * replacement per epoch (seeded shuffle); loss_out (may be NULL, steps
 * doubles) receives each step's loss, the mean over the minibatch of
 * ||x0 - x*||^2 / ||x*||^2 in the scaled variables. */
int  plsqr_cnn_train(plsqr_cnn *h, int steps, int batch, float lr, unsigned long long seed,
                     double *loss_out);
/* x0 for count samples of A^T b (host arrays). */
int  plsqr_cnn_predict(plsqr_cnn *h, int count, const double *atb, double *x0);
/* Loss and its gradient with respect to every parameter (params layout) for
 * one minibatch (batch <= max_batch), without updating the parameters. */
int  plsqr_cnn_gradient(plsqr_cnn *h, int batch, const double *atb, const double *xstar,
                        float *grads, double *loss);
/* One 32 -> 32 layer on host arrays in[g][k], out[g][n] (g = sample*V + voxel):
 * operands are rounded to bf16, products summed in FP32.  backward: din is
 * the data gradient of dout (no activation derivative), dw and dbias the
 * parameter gradients. */
int  plsqr_cnn_conv_forward(int nx, int ny, int nz, int batch, const float *in, const float *w,
                            const float *bias, int relu, float *out);
int  plsqr_cnn_conv_backward(int nx, int ny, int nz, int batch, const float *in, const float *w,
                             const float *dout, float *din, float *dw, float *dbias);
/* plsqr_cnn_conv_forward on device pointers that allocates nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops).  in_dev and out_dev
 * hold batch*nx*ny*nz*32 floats, w_dev 27*32*32, bias_dev 32 (or NULL).
 * Enqueued on stream (hipStream_t as void*, NULL = the default stream); no
 * synchronisation.  Its scratch (the bf16 operands) is one fixed array per
 * device inside the library, so batch*nx*ny*nz may not exceed 131072
 * (PLSQR_ERR_ARG), and calls that overlap on one device race. */
int  plsqr_cnn_conv_forward_noalloc(int nx, int ny, int nz, int batch, const float *in_dev,
                                    const float *w_dev, const float *bias_dev, int relu, float *out_dev,
                                    void *stream);

/* ---------------------------------------------------------------------- */
/* Batched solves: many independent systems, up to 4 right-hand sides each */
/* ---------------------------------------------------------------------- */
/* The reference solves one system per process run (plsqr3.c main ->
 * call_lsqr, lsqr.c:34-184); a batch runs lsqr.c's iteration for every
 * (system, right-hand side) pair at once.  Each pair has its own scalar
 * state and stops on its own test (tol, itn); a stopped pair keeps its x
 * while the others go on, and a system whose pairs have all stopped costs
 * no further matrix traffic.  Every iteration of the whole batch is one
 * captured HIP graph of six launches, whatever the number of systems. */
typedef struct plsqr_batch plsqr_batch;

#define PLSQR_BATCH_MAX_RHS 4
enum {
    PLSQR_BATCH_NO_TUNE = 1,     /* keep the default launch shapes at create      */
    PLSQR_BATCH_TUNE_MIXED = 2   /* also tune the FP32 kernels at create           */
};

/* Launch shapes the batch's autotuner chooses among (mode 1: y += A x over
 * rows; mode 2: x += A^T y over 4 x 4 x 8 voxel tiles). */
typedef struct plsqr_batch_tuning {
    int    rows_per_wave;   /* mode 1: rows per 64-lane wavefront: 2, 4, 8, 16  */
    int    nnz_per_lane;    /* mode 1: nonzeros each lane keeps in flight: 2, 4 */
    int    block_rows;      /* mode 1 workgroup size: 128 or 256               */
    int    block_cols;      /* mode 2 workgroup size: 128, 256 or 512          */
    double ms_rows, ms_cols; /* measured time per launch of the chosen shapes  */
    int    candidates;      /* shapes timed (0: defaults, not measured)        */
    int    gather_kib;      /* mode 2: LDS per workgroup for the gathered ray
                               values, 32 or 64 KiB (0 in set_tuning: 32)     */
} plsqr_batch_tuning;

/* systems[p]: the p-th system in the loaded layout (b and x_true unused).
 * grids: NULL or n_problems entries (a zero-sized entry: no grid).  n_rhs:
 * right-hand sides per system (1..4).  Validates every system on the host,
 * builds the block-diagonal device layout and, unless PLSQR_BATCH_NO_TUNE,
 * times the candidate launch shapes on it and keeps the fastest. */
int  plsqr_batch_create(const plsqr_system *const *systems, const plsqr_grid *grids,
                        int n_problems, int n_rhs, int flags, plsqr_batch **out);
void plsqr_batch_destroy(plsqr_batch *h);
/* last error of h; with h == NULL the last plsqr_batch_create failure */
const char *plsqr_batch_last_error(const plsqr_batch *h);
void plsqr_batch_clear_error(plsqr_batch *h);
int  plsqr_batch_dims(const plsqr_batch *h, int *n_problems, int *n_rhs,
                      long long *rows, long long *cols, long long *nnz);
size_t plsqr_batch_device_bytes(const plsqr_batch *h);

/* Times every candidate shape (reps launches each, median) for one
 * precision and keeps the fastest; *out (may be NULL) receives it. */
int plsqr_batch_autotune(plsqr_batch *h, int precision, int reps, plsqr_batch_tuning *out);
int plsqr_batch_get_tuning(const plsqr_batch *h, int precision, plsqr_batch_tuning *out);
/* Force a shape (its ms_* and candidates fields are ignored). */
int plsqr_batch_set_tuning(plsqr_batch *h, int precision, const plsqr_batch_tuning *t);

/* Array layout of every batch vector: system after system, and inside a
 * system right-hand side after right-hand side; x parts hold n_p values in
 * the caller's column order, y parts m_k + m_d (kernel rows, then damping
 * rows), b parts m_k.
 * aprod of every system and right-hand side at once, host arrays:
 *   mode 1: y <- y + A x      mode 2: x <- x + A^T y
 * precision PLSQR_PREC_MIXED runs the FP32 kernels (operands rounded to FP32). */
int plsqr_batch_aprod(plsqr_batch *h, int mode, int precision, double *x, double *y);

/* LSQR from x = 0 for every (system, right-hand side): opts as for
 * plsqr_lsqr_solve (precision FP64 or MIXED; MIXED refines in FP64 every
 * refine_every inner iterations and stops on the FP64 test2 of the true
 * residual).  out: n_problems * n_rhs results in the array order above;
 * out[i].elapsed_ms is the device time from the start of the solve until
 * pair i stopped, and *total_ms (may be NULL) the time of the whole batch.
 * MIXED accepts precond = PLSQR_PRECOND_COLSCALE (precond_power as for
 * plsqr_lsqr_solve): each problem's inner LSQR runs on A_p D_p and pair i's
 * FP64 test2 takes the anorm set by plsqr_batch_set_anorm_refs, or
 * ||A_p||_F (anorm_ref is not used here).  refine_adaptive = 1 (MIXED,
 * tol > 0): after the first check, the next one comes as soon as any live
 * pair's inner test2 estimate times its FP64 / inner ratio at the last check
 * reaches tol (polled every 8 iterations), at most 4 * refine_every apart. */
int plsqr_batch_solve(plsqr_batch *h, const double *b_host, const plsqr_solve_opts *opts,
                      double *x_host, plsqr_solve_out *out, double *total_ms);
/* The anorm of the FP64 test2 of every (problem, right-hand side) pair of a
 * column-scaled batch solve (refs: n_problems * n_rhs values in the order of
 * out, e.g. the out[i].anorm of a plain FP64 batch solve of the same b, so
 * that each pair stops exactly as strictly; a value <= 0 means ||A_p||_F).
 * refs == NULL clears them.  No counterpart in the reference. */
int plsqr_batch_set_anorm_refs(plsqr_batch *h, const double *refs);
/* Same on device arrays (b_dev and x_dev in the array layout above). */
int plsqr_batch_solve_device(plsqr_batch *h, const double *b_dev, const plsqr_solve_opts *opts,
                             double *x_dev, plsqr_solve_out *out, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* PLSQR_HIP_H_ */
