// C-ABI of batched solves: many independent systems, several right-hand sides
// each, every LSQR iteration of the whole batch one captured graph of six
// launches (include/plsqr_hip.h, plsqr_batch_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "batch_build.h"
#include "plsqr_batch.h"
#include "plsqr_hip.h"
#include "plsqr_kernels.h"

using plsqr::Scal;

namespace {

constexpr int TW = 128;         // voxels per tile (one 4 x 4 x 8 brick)
constexpr int ROW_CHUNK = 512;  // kernel rows per mode-1 piece
constexpr int DAMP_CHUNK = 2048;
constexpr int COL_CHUNK = 1024;
constexpr int SEG_CH = 512;        // mode-1 segment entries staged per wavefront copy

struct Shape {
    int L = 16, bs1 = 256, U = 4, bs2 = 256, kib2 = 32;
    double ms1 = 0.0, ms2 = 0.0;
    int candidates = 0;
};

struct GraphCache {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

}  // namespace

struct plsqr_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    int cu = 256;
    int P = 0, K = 1, Kr = 1;
    long long rows = 0, cols = 0, nnz = 0, dnnz = 0, slots = 0, segs = 0, cuts = 0;
    int items = 0, nblk = 0, nrch = 0, ncch = 0;
    long long partials = 0;
    // caller layout offsets (host)
    std::vector<int> mk, md, np;
    std::vector<long long> row0, col0, xb, yb, bb;
    // matrices
    int *rp = nullptr, *ci = nullptr;
    double *rv = nullptr;
    float *rv32 = nullptr;
    int *dcp = nullptr, *dri = nullptr;
    double *dcv = nullptr;
    int *i_prob = nullptr, *i_c0 = nullptr, *i_c1 = nullptr, *blk_item = nullptr, *seg_ray = nullptr;
    int *blk_rch = nullptr;  // row chunk of each mode-1 per-ray-sum workgroup
    int nblk_r = 0;
    long long *i_pbase = nullptr;
    long long *slot_base = nullptr, *seg_start = nullptr, *cut_base = nullptr;
    uint16_t *vox_len = nullptr, *sseg = nullptr, *vox_cut = nullptr;
    double *sval = nullptr;
    float *sval32 = nullptr;
    // mode 1 as per-brick ray segments (RowTiles of every problem) + per-ray sums
    bool rtiled = false;
    long long groups = 0, ents = 0;
    int *rt_gbase = nullptr, *rt_rptr = nullptr, *rt_rseg = nullptr;
    long long *rt_ebase = nullptr;
    uint8_t *rt_len = nullptr, *rt_vox = nullptr;
    double *rt_val = nullptr;
    double *segv = nullptr;  // segment partials: segs*K doubles, or segs*K floats on the FP32 path
    float *rt_val32 = nullptr;
    // pieces and ranges
    int *rch_row = nullptr, *rch_prob = nullptr, *cch_col = nullptr, *cch_prob = nullptr;
    uint8_t *rch_kind = nullptr;
    int *r_rch = nullptr, *r_cch = nullptr;
    long long *d_pofs = nullptr;
    int *d_nslice = nullptr;
    double *pp = nullptr;  // slice partials of mode 2
    // caller layout maps
    int *row_prob = nullptr, *row_old = nullptr, *col_prob = nullptr, *col_old = nullptr;
    long long *d_row0 = nullptr, *d_col0 = nullptr, *d_xb = nullptr, *d_yb = nullptr, *d_bb = nullptr;
    int *d_mk = nullptr, *d_mrow = nullptr, *d_ncol = nullptr;
    // vectors (interleaved K)
    double *u64 = nullptr, *v64 = nullptr, *b0 = nullptr, *x = nullptr, *w = nullptr, *dx = nullptr, *xt = nullptr;
    float *u32 = nullptr, *v32 = nullptr;
    double *part = nullptr, *sums = nullptr, *stage = nullptr;
    long long stage_len = 0;
    int part_len = 0;
    Scal *sc = nullptr, *sc_plain = nullptr;
    int *live_d = nullptr, *live_h = nullptr;  // live states, and those predicted to have converged
    double *ratio_d = nullptr;                  // S: FP64 / inner test2 at the last check (adaptive checks)
    void *hstage = nullptr;  // pinned: the states (S Scal) or the 2 S state sums on their way to the host
    // column scaling of the mixed solve: cs (cols) holds the scales during a
    // scaled solve and 1 otherwise (the FP32 combine always reads it);
    // colss the column sums of squares, fro_p each problem's ||A_p||_F,
    // anorm_refs the caller's per-(problem, rhs) anorm for the FP64 test2
    double *cs = nullptr, *colss = nullptr;
    bool cs_scaled = false;  // cs holds scales (a scaled solve ran or stopped early)
    std::vector<double> fro_p, anorm_refs;
    Shape shape[2];
    GraphCache graphs[2];
    std::vector<hipEvent_t> evs;
    std::vector<void *> allocs;
    size_t bytes = 0;
    std::string err;
};

namespace {

using namespace plsqr::batch;

int set_err(plsqr_batch *b, int code, const std::string &msg) {
    if (b && b->err.empty()) b->err = msg;
    return code;
}

thread_local std::string g_batch_err;

#define BT_TRY(b, expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_err(b, PLSQR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BT_LAUNCH(b) BT_TRY(b, hipGetLastError())

template <class T>
int balloc(plsqr_batch *b, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(b, PLSQR_ERR_NOMEM,
                       "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed: " + hipGetErrorString(e));
    }
    b->allocs.push_back((void *)*p);
    b->bytes += sizeof(T) * count;
    return PLSQR_OK;
}

template <class T>
void bfree(plsqr_batch *b, T *p, size_t count) {
    if (!p) return;
    auto it = std::find(b->allocs.begin(), b->allocs.end(), (void *)p);
    if (it != b->allocs.end()) {
        b->allocs.erase(it);
        b->bytes -= sizeof(T) * std::max<size_t>(count, 1);
    }
    (void)hipFree(p);
}

template <class T>
int upload(plsqr_batch *b, T *dst, const T *src, size_t count) {
    if (count == 0) return PLSQR_OK;
    BT_TRY(b, hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyHostToDevice, b->stream));
    return PLSQR_OK;
}

int grid_for(long long work, int per_block, int cap) {
    long long g = (work + per_block - 1) / per_block;
    return (int)std::max(1LL, std::min<long long>(g, cap));
}

int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

ExtMap ext_map(const plsqr_batch *b) {
    ExtMap m;
    m.row_prob = b->row_prob;
    m.row_old = b->row_old;
    m.col_prob = b->col_prob;
    m.col_old = b->col_old;
    m.row0 = b->d_row0;
    m.col0 = b->d_col0;
    m.xb = b->d_xb;
    m.yb = b->d_yb;
    m.bb = b->d_bb;
    m.mk = b->d_mk;
    m.mrow = b->d_mrow;
    m.ncol = b->d_ncol;
    return m;
}

// -------------------------------------------------------------------------
// kernel dispatch over (T, K) and the launch shape
// -------------------------------------------------------------------------
template <class T, int K>
int rows_k(plsqr_batch *b, const Shape &s, const T *av, const T *xv, T *yv, const Scal *sc) {
    const dim3 g(b->nrch), blk1(s.bs1);
    if (b->rtiled) {
        const T *ev = std::is_same<T, float>::value ? (const T *)b->rt_val32 : (const T *)b->rt_val;
        hipLaunchKernelGGL((kb_seg<T, K, 256, SEG_CH>), dim3(b->nblk), dim3(256), 0, b->stream, b->blk_item,
                           b->i_prob, b->i_c0, b->i_c1, b->seg_start, b->rt_gbase, b->rt_ebase, b->rt_len, b->rt_vox,
                           ev, xv, sc, (T *)b->segv);
        BT_LAUNCH(b);
        const dim3 gr(b->nblk_r);
        if (s.bs1 == 128)
            hipLaunchKernelGGL((kb_rows_comb<T, K, 4, 128>), gr, dim3(128), 0, b->stream, b->blk_rch, b->rch_row,
                               b->rch_prob, b->rch_kind, b->rt_rptr, b->rt_rseg, (const T *)b->segv, b->rp,
                               b->ci, av, xv, yv, sc, b->part);
        else
            hipLaunchKernelGGL((kb_rows_comb<T, K, 4, 256>), gr, dim3(256), 0, b->stream, b->blk_rch, b->rch_row,
                               b->rch_prob, b->rch_kind, b->rt_rptr, b->rt_rseg, (const T *)b->segv, b->rp,
                               b->ci, av, xv, yv, sc, b->part);
        BT_LAUNCH(b);
        return PLSQR_OK;
    }
#define ROWS_CASE(L_, BS_, U_)                                                                                  \
    if (s.L == L_ && s.bs1 == BS_ && s.U == U_) {                                                              \
        hipLaunchKernelGGL((kb_rows<T, K, L_, BS_, U_>), g, blk1, 0, b->stream, b->nrch, b->rch_row, b->rch_prob, \
                           b->rch_kind, b->rp, b->ci, av, xv, yv, sc, b->part);                                  \
        BT_LAUNCH(b);                                                                                           \
        return PLSQR_OK;                                                                                        \
    }
    ROWS_CASE(4, 128, 2) ROWS_CASE(4, 128, 4) ROWS_CASE(4, 256, 2) ROWS_CASE(4, 256, 4)
    ROWS_CASE(8, 128, 2) ROWS_CASE(8, 128, 4) ROWS_CASE(8, 256, 2) ROWS_CASE(8, 256, 4)
    ROWS_CASE(16, 128, 2) ROWS_CASE(16, 128, 4) ROWS_CASE(16, 256, 2) ROWS_CASE(16, 256, 4)
    ROWS_CASE(32, 128, 2) ROWS_CASE(32, 128, 4) ROWS_CASE(32, 256, 2) ROWS_CASE(32, 256, 4)
#undef ROWS_CASE
    return set_err(b, PLSQR_ERR_ARG, "unsupported mode-1 launch shape");
}

// kib2 KiB of LDS hold CAP = kib2 * 1024 / (K * sizeof(T)) gathered
// segments; items with more are staged in rounds
template <class T, int K>
int cols_k(plsqr_batch *b, const Shape &s, const T *sv, const T *uv, T *vv, const Scal *sc) {
    const dim3 g(b->nblk);
    bool done = false;
#define COLS_CASE(BS_, KIB_)                                                                                     \
    if (s.bs2 == BS_ && s.kib2 == KIB_) {                                                                       \
        constexpr int CAP = KIB_ * 1024 / (K * (int)sizeof(T));                                                 \
        hipLaunchKernelGGL((kb_cols<T, K, TW, BS_, CAP, plsqr::TileLayout::CUT>), g, dim3(BS_), 0, b->stream,     \
                           b->blk_item, b->i_prob, b->i_c0, b->i_c1, b->i_pbase, b->slot_base, b->seg_start,       \
                           b->vox_len, b->seg_ray, sv, b->sseg, b->cut_base, b->vox_cut, uv, sc, b->pp);          \
        done = true;                                                                                           \
    }
    COLS_CASE(128, 32) COLS_CASE(256, 32) COLS_CASE(512, 32)
    COLS_CASE(128, 64) COLS_CASE(256, 64) COLS_CASE(512, 64)
#undef COLS_CASE
    if (!done) return set_err(b, PLSQR_ERR_ARG, "unsupported mode-2 launch shape");
    BT_LAUNCH(b);
    hipLaunchKernelGGL((kb_combine<T, K, 256>), dim3(b->ncch), dim3(256), 0, b->stream, b->ncch, b->cch_col,
                       b->cch_prob, b->d_pofs, b->d_nslice, b->d_col0, b->d_ncol, (const double *)b->pp, b->dcp,
                       b->dri, b->dcv, uv, vv, sc, b->part,
                       std::is_same<T, float>::value ? (const double *)b->cs : (const double *)nullptr);
    BT_LAUNCH(b);
    return PLSQR_OK;
}

template <class T>
int launch_rows(plsqr_batch *b, const Shape &s, const T *av, const T *xv, T *yv, const Scal *sc) {
    switch (b->K) {
        case 1: return rows_k<T, 1>(b, s, av, xv, yv, sc);
        case 2: return rows_k<T, 2>(b, s, av, xv, yv, sc);
        default: return rows_k<T, 4>(b, s, av, xv, yv, sc);
    }
}

template <class T>
int launch_cols(plsqr_batch *b, const Shape &s, const T *sv, const T *uv, T *vv, const Scal *sc) {
    switch (b->K) {
        case 1: return cols_k<T, 1>(b, s, sv, uv, vv, sc);
        case 2: return cols_k<T, 2>(b, s, sv, uv, vv, sc);
        default: return cols_k<T, 4>(b, s, sv, uv, vv, sc);
    }
}

template <int PH>
int launch_scalar(plsqr_batch *b, const int *rng) {
    const dim3 g(b->P * b->K), blk(256);
    switch (b->K) {
        case 1: hipLaunchKernelGGL((kb_scalar<1, PH>), g, blk, 0, b->stream, rng, b->part, b->sc); break;
        case 2: hipLaunchKernelGGL((kb_scalar<2, PH>), g, blk, 0, b->stream, rng, b->part, b->sc); break;
        default: hipLaunchKernelGGL((kb_scalar<4, PH>), g, blk, 0, b->stream, rng, b->part, b->sc); break;
    }
    BT_LAUNCH(b);
    return PLSQR_OK;
}

template <class T>
int launch_update(plsqr_batch *b, T *vv, double *xacc) {
    const dim3 g(b->ncch), blk(256);
    switch (b->K) {
        case 1: hipLaunchKernelGGL((kb_update<T, 1, 256>), g, blk, 0, b->stream, b->ncch, b->cch_col, b->cch_prob, b->sc, vv, b->w, xacc, b->part); break;
        case 2: hipLaunchKernelGGL((kb_update<T, 2, 256>), g, blk, 0, b->stream, b->ncch, b->cch_col, b->cch_prob, b->sc, vv, b->w, xacc, b->part); break;
        default: hipLaunchKernelGGL((kb_update<T, 4, 256>), g, blk, 0, b->stream, b->ncch, b->cch_col, b->cch_prob, b->sc, vv, b->w, xacc, b->part); break;
    }
    BT_LAUNCH(b);
    return PLSQR_OK;
}

template <class T>
int launch_init_vw(plsqr_batch *b, T *vv) {
    const dim3 g(b->ncch), blk(256);
    switch (b->K) {
        case 1: hipLaunchKernelGGL((kb_init_vw<T, 1>), g, blk, 0, b->stream, b->ncch, b->cch_col, b->cch_prob, b->sc, vv, b->w); break;
        case 2: hipLaunchKernelGGL((kb_init_vw<T, 2>), g, blk, 0, b->stream, b->ncch, b->cch_col, b->cch_prob, b->sc, vv, b->w); break;
        default: hipLaunchKernelGGL((kb_init_vw<T, 4>), g, blk, 0, b->stream, b->ncch, b->cch_col, b->cch_prob, b->sc, vv, b->w); break;
    }
    BT_LAUNCH(b);
    return PLSQR_OK;
}

template <class T>
int launch_ssq_rows(plsqr_batch *b, const T *y) {
    const dim3 g(b->nrch), blk(256);
    switch (b->K) {
        case 1: hipLaunchKernelGGL((kb_ssq_rows<T, 1, 256>), g, blk, 0, b->stream, b->nrch, b->rch_row, y, b->part); break;
        case 2: hipLaunchKernelGGL((kb_ssq_rows<T, 2, 256>), g, blk, 0, b->stream, b->nrch, b->rch_row, y, b->part); break;
        default: hipLaunchKernelGGL((kb_ssq_rows<T, 4, 256>), g, blk, 0, b->stream, b->nrch, b->rch_row, y, b->part); break;
    }
    BT_LAUNCH(b);
    return PLSQR_OK;
}

int launch_state_sums(plsqr_batch *b, const int *rng, double *out) {
    const dim3 g(b->P * b->K), blk(256);
    switch (b->K) {
        case 1: hipLaunchKernelGGL((kb_state_sums<1>), g, blk, 0, b->stream, rng, b->part, out); break;
        case 2: hipLaunchKernelGGL((kb_state_sums<2>), g, blk, 0, b->stream, rng, b->part, out); break;
        default: hipLaunchKernelGGL((kb_state_sums<4>), g, blk, 0, b->stream, rng, b->part, out); break;
    }
    BT_LAUNCH(b);
    return PLSQR_OK;
}

// u64 = [b;0] - A xv, per-chunk sums of squares in part
int launch_residual(plsqr_batch *b, const double *xv) {
    const dim3 g(b->nrch), blk(256);
    switch (b->K) {
        case 1: hipLaunchKernelGGL((kb_residual<1, 256>), g, blk, 0, b->stream, b->nrch, b->rch_row, b->rp, b->ci, b->rv, xv, b->b0, b->u64, b->part); break;
        case 2: hipLaunchKernelGGL((kb_residual<2, 256>), g, blk, 0, b->stream, b->nrch, b->rch_row, b->rp, b->ci, b->rv, xv, b->b0, b->u64, b->part); break;
        default: hipLaunchKernelGGL((kb_residual<4, 256>), g, blk, 0, b->stream, b->nrch, b->rch_row, b->rp, b->ci, b->rv, xv, b->b0, b->u64, b->part); break;
    }
    BT_LAUNCH(b);
    return PLSQR_OK;
}

template <class T>
int launch_load_b(plsqr_batch *b, const double *ext, T *u) {
    const int g = grid_for(b->rows, 256, b->cu * 8);
    const ExtMap m = ext_map(b);
    switch (b->K) {
        case 1: hipLaunchKernelGGL((kb_load_b<T, 1>), dim3(g), dim3(256), 0, b->stream, b->rows, b->Kr, m, ext, u); break;
        case 2: hipLaunchKernelGGL((kb_load_b<T, 2>), dim3(g), dim3(256), 0, b->stream, b->rows, b->Kr, m, ext, u); break;
        default: hipLaunchKernelGGL((kb_load_b<T, 4>), dim3(g), dim3(256), 0, b->stream, b->rows, b->Kr, m, ext, u); break;
    }
    BT_LAUNCH(b);
    return PLSQR_OK;
}

template <class T>
int launch_conv(plsqr_batch *b, bool rows, int dir, double *ext, T *vec) {
    const long long len = rows ? b->rows : b->cols;
    const int g = grid_for(len, 256, b->cu * 8);
    const ExtMap m = ext_map(b);
#define CONV(K_)                                                                                                \
    if (rows)                                                                                                   \
        hipLaunchKernelGGL((kb_conv_rows<T, K_>), dim3(g), dim3(256), 0, b->stream, len, b->Kr, m, dir, ext, vec); \
    else                                                                                                        \
        hipLaunchKernelGGL((kb_conv_cols<T, K_>), dim3(g), dim3(256), 0, b->stream, len, b->Kr, m, dir, ext, vec);
    switch (b->K) {
        case 1: CONV(1) break;
        case 2: CONV(2) break;
        default: CONV(4) break;
    }
#undef CONV
    BT_LAUNCH(b);
    return PLSQR_OK;
}

template <class T>
int fill(plsqr_batch *b, T *a, long long len, T val) {
    hipLaunchKernelGGL((kb_fill<T>), dim3(grid_for(len, 256, b->cu * 8)), dim3(256), 0, b->stream, len, val, a);
    BT_LAUNCH(b);
    return PLSQR_OK;
}

int ensure_fp32(plsqr_batch *b) {
    if (b->rv32) return PLSQR_OK;
    int rc;
    if ((rc = balloc(b, &b->cs, (size_t)b->cols)) || (rc = fill(b, b->cs, b->cols, 1.0)) ||
        (rc = balloc(b, &b->rv32, (size_t)b->nnz)) || (rc = balloc(b, &b->sval32, (size_t)b->slots)) ||
        (rc = balloc(b, &b->u32, (size_t)b->rows * b->K)) || (rc = balloc(b, &b->v32, (size_t)b->cols * b->K)))
        return rc;
    hipLaunchKernelGGL((kb_convert<float>), dim3(grid_for(b->nnz, 256, b->cu * 8)), dim3(256), 0, b->stream, b->nnz,
                       b->rv, b->rv32);
    hipLaunchKernelGGL((kb_convert<float>), dim3(grid_for(b->slots, 256, b->cu * 8)), dim3(256), 0, b->stream,
                       b->slots, b->sval, b->sval32);
    BT_LAUNCH(b);
    if (b->rtiled) {
        if ((rc = balloc(b, &b->rt_val32, (size_t)b->ents + 4))) return rc;
        hipLaunchKernelGGL((kb_convert<float>), dim3(grid_for(b->ents, 256, b->cu * 8)), dim3(256), 0, b->stream,
                           b->ents, b->rt_val, b->rt_val32);
        BT_LAUNCH(b);
    }
    BT_TRY(b, hipStreamSynchronize(b->stream));
    return PLSQR_OK;
}

void drop_graphs(plsqr_batch *b) {
    for (auto &g : b->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
        g.graph = nullptr;
    }
}

// one LSQR iteration of every live state (lsqr.c:487-775)
template <class T>
int enqueue_iteration(plsqr_batch *b, const Shape &s, const T *av, const T *sv, T *u, T *v, double *xacc) {
    using namespace plsqr::batch;
    int rc;
    if ((rc = launch_rows<T>(b, s, av, v, u, b->sc))) return rc;
    if ((rc = launch_scalar<PH_BETA>(b, b->r_rch))) return rc;
    if ((rc = launch_cols<T>(b, s, sv, u, v, b->sc))) return rc;
    if ((rc = launch_scalar<PH_ALPHA>(b, b->r_cch))) return rc;
    if ((rc = launch_update<T>(b, v, xacc))) return rc;
    if ((rc = launch_scalar<PH_EST>(b, b->r_cch))) return rc;
    return PLSQR_OK;
}

// u, v, w hold the start (u = [b;0] or r, v = w = xacc = 0); beta, alpha
// and the first v, w (lsqr.c:343-445)
template <class T>
int enqueue_start(plsqr_batch *b, const Shape &s, const T *sv, T *u, T *v) {
    int rc;
    if ((rc = launch_ssq_rows<T>(b, u))) return rc;
    if ((rc = launch_scalar<PH_INIT_BETA>(b, b->r_rch))) return rc;
    if ((rc = launch_cols<T>(b, s, sv, u, v, b->sc))) return rc;
    if ((rc = launch_scalar<PH_INIT_ALPHA>(b, b->r_cch))) return rc;
    return launch_init_vw<T>(b, v);
}

template <class T>
int get_graph(plsqr_batch *b, int prec, const T *av, const T *sv, T *u, T *v, double *xacc, hipGraphExec_t *out) {
    GraphCache &gc = b->graphs[prec];
    if (!gc.exec) {
        BT_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_iteration<T>(b, b->shape[prec], av, sv, u, v, xacc);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return set_err(b, PLSQR_ERR_HIP, std::string("capture: ") + hipGetErrorString(ce));
        gc.graph = g;
        BT_TRY(b, hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0));
    }
    *out = gc.exec;
    return PLSQR_OK;
}

int ensure_events(plsqr_batch *b, size_t count) {
    while (b->evs.size() < count) {
        hipEvent_t e;
        BT_TRY(b, hipEventCreate(&e));
        b->evs.push_back(e);
    }
    return PLSQR_OK;
}

Scal fresh_state(const plsqr_solve_opts *o, int itn_max) {
    Scal z;
    memset(&z, 0, sizeof(z));
    z.damp = o->damp;
    z.atol = o->atol;
    z.btol = o->btol;
    z.ctol = o->conlim > 0.0 ? 1.0 / o->conlim : 0.0;
    z.tol = o->tol;
    z.itnlim = o->itnlim;
    z.itn_max = itn_max;
    z.cs2 = -1.0;  // lsqr.c:293
    return z;
}

void default_opts(const plsqr_solve_opts *o, plsqr_solve_opts *d) {
    *d = *o;
    if (d->itnlim <= 0) d->itnlim = d->itn > 0 ? d->itn : 1;
    if (d->refine_every <= 0) d->refine_every = 25;
}

// runs `count` graph launches, polling the live count every 8; returns the
// number launched.  Records b->evs[first + i] after launch i.
// ratio_d (may be NULL): per state the FP64 / inner test2 ratio of the last
// check; with it the chunk also ends once a live state's predicted test2
// reaches tol, but not before min_pred launches: a check costs the whole
// batch, and the pairs of a batch reach tol at different times, so checks
// stay at most as frequent as the fixed schedule.
int run_launches(plsqr_batch *b, hipGraphExec_t exec, int count, size_t first, int *launched,
                 const double *ratio_d = nullptr, double tol = 0.0, int min_pred = 0) {
    int rc;
    if ((rc = ensure_events(b, first + (size_t)count + 1))) return rc;
    *launched = 0;
    for (int i = 0; i < count; ++i) {
        BT_TRY(b, hipGraphLaunch(exec, b->stream));
        BT_TRY(b, hipEventRecord(b->evs[first + (size_t)i], b->stream));
        ++*launched;
        if ((i + 1) % 8 == 0 && i + 1 < count) {
            if (ratio_d)
                hipLaunchKernelGGL(kb_count_pred, dim3(1), dim3(256), 0, b->stream, b->P * b->K, (const Scal *)b->sc,
                                   ratio_d, tol, b->live_d);
            else
                hipLaunchKernelGGL(kb_count_live, dim3(1), dim3(256), 0, b->stream, b->P * b->K, (const Scal *)b->sc,
                                   b->live_d);
            BT_LAUNCH(b);
            BT_TRY(b, hipMemcpyAsync(b->live_h, b->live_d, sizeof(int) * 2, hipMemcpyDeviceToHost, b->stream));
            BT_TRY(b, hipStreamSynchronize(b->stream));
            if (b->live_h[0] == 0) break;
            if (ratio_d && b->live_h[1] > 0 && i + 1 >= min_pred) break;
        }
    }
    return PLSQR_OK;
}

// every state (S Scal) to the host through the pinned staging
int read_states(plsqr_batch *b, std::vector<Scal> &h) {
    const size_t S = (size_t)b->P * b->K;
    h.resize(S);
    BT_TRY(b, hipMemcpyAsync(b->hstage, b->sc, sizeof(Scal) * S, hipMemcpyDeviceToHost, b->stream));
    BT_TRY(b, hipStreamSynchronize(b->stream));
    std::memcpy(h.data(), b->hstage, sizeof(Scal) * S);
    return PLSQR_OK;
}

// Timing events of one solve, destroyed on every exit path once the stream
// has drained.
struct EventPair {
    plsqr_batch *b;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    explicit EventPair(plsqr_batch *o) : b(o) {}
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
    int create() {
        BT_TRY(b, hipEventCreate(&e0));
        BT_TRY(b, hipEventCreate(&e1));
        return PLSQR_OK;
    }
    ~EventPair() {
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

float elapsed(hipEvent_t a, hipEvent_t c) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, c) != hipSuccess) return 0.f;
    return ms;
}

int solve_fp64(plsqr_batch *b, const double *b_dev, const plsqr_solve_opts *o, plsqr_solve_out *out, double *total_ms) {
    int rc;
    const int itn = std::max(0, o->itn);
    const int S = b->P * b->K;
    std::vector<Scal> h((size_t)S, fresh_state(o, itn));
    if ((rc = upload(b, b->sc, h.data(), h.size()))) return rc;
    EventPair ev(b);
    if ((rc = ensure_events(b, 2)) || (rc = ev.create())) return rc;
    hipEvent_t e0 = ev.e0, e_start = ev.e1;
    BT_TRY(b, hipEventRecord(e0, b->stream));
    if ((rc = launch_load_b<double>(b, b_dev, b->u64)) || (rc = fill(b, b->v64, b->cols * b->K, 0.0)) ||
        (rc = fill(b, b->x, b->cols * b->K, 0.0)) || (rc = fill(b, b->w, b->cols * b->K, 0.0)) ||
        (rc = enqueue_start<double>(b, b->shape[0], b->sval, b->u64, b->v64)))
        return rc;
    BT_TRY(b, hipEventRecord(e_start, b->stream));
    int launched = 0;
    if (itn > 0) {
        hipGraphExec_t exec;
        if ((rc = get_graph<double>(b, 0, b->rv, b->sval, b->u64, b->v64, b->x, &exec))) return rc;
        if ((rc = run_launches(b, exec, itn, 0, &launched))) return rc;
    }
    if ((rc = read_states(b, h))) return rc;
    const float t_start = elapsed(e0, e_start);
    float t_all = t_start;
    for (int p = 0; p < b->P; ++p)
        for (int k = 0; k < b->Kr; ++k) {
            const Scal &s = h[(size_t)p * b->K + k];
            plsqr_solve_out &r = out[(size_t)p * b->Kr + k];
            memset(&r, 0, sizeof(r));
            int istop = s.istop;
            if (s.damp > 0.0 && istop == 2) istop = 3;  // lsqr.c:895-896
            r.istop = s.init_zero ? 0 : istop;
            r.itn = s.itn;
            r.anorm = s.anorm;
            r.acond = s.acond;
            r.rnorm = s.rnorm;
            r.arnorm = s.test2;  // lsqr.c:913
            r.xnorm = s.xnorm;
            r.bnorm = s.bnorm;
            r.arnorm_true = s.arnorm;
            r.inner_total = s.itn;
            const int last = std::min(s.itn, launched);
            r.elapsed_ms = last > 0 ? elapsed(e0, b->evs[(size_t)last - 1]) : t_start;
            t_all = std::max(t_all, (float)r.elapsed_ms);
        }
    if (launched > 0) t_all = std::max(t_all, elapsed(e0, b->evs[(size_t)launched - 1]));
    if (total_ms) *total_ms = t_all;
    return PLSQR_OK;
}

// per-state ||r|| and ||A^T r|| in FP64 for r = [b;0] - A xt (r into u64,
// A^T r into v64); host arrays rn2[S], an2[S] receive the squares
int true_norms(plsqr_batch *b, const double *xv, std::vector<double> &sums) {
    int rc;
    const int S = b->P * b->K;
    if ((rc = launch_residual(b, xv)) || (rc = launch_state_sums(b, b->r_rch, b->sums)) || (rc = fill(b, b->v64, b->cols * b->K, 0.0)) ||
        (rc = launch_cols<double>(b, b->shape[0], b->sval, b->u64, b->v64, b->sc_plain)) ||
        (rc = launch_state_sums(b, b->r_cch, b->sums + S)))
        return rc;
    sums.resize((size_t)S * 2);
    BT_TRY(b, hipMemcpyAsync(b->hstage, b->sums, sizeof(double) * 2 * S, hipMemcpyDeviceToHost, b->stream));
    BT_TRY(b, hipStreamSynchronize(b->stream));
    std::memcpy(sums.data(), b->hstage, sizeof(double) * 2 * S);
    return PLSQR_OK;
}

// Column sums of squares of every problem's [A_k; D] (once per handle: the
// FP64 mode 2 on the squared values with u = 1), ||A_p||_F, and the scales
// cs = colss^(-power/2) for a scaled mixed solve.
int set_colscale(plsqr_batch *b, double power) {
    int rc;
    if (!(power > 0.0) || !std::isfinite(power)) return set_err(b, PLSQR_ERR_ARG, "precond_power must be > 0");
    if (b->fro_p.size() != (size_t)b->P) {  // not formed yet (or an earlier attempt failed)
        double *sq_s = nullptr, *sq_d = nullptr;
        if ((!b->colss && (rc = balloc(b, &b->colss, (size_t)b->cols))) || (rc = balloc(b, &sq_s, (size_t)b->slots)) ||
            (rc = balloc(b, &sq_d, (size_t)std::max(b->dnnz, 1LL))))
            return rc;
        hipLaunchKernelGGL(kb_square, dim3(grid_for(b->slots, 256, b->cu * 8)), dim3(256), 0, b->stream, b->slots,
                           (const double *)b->sval, sq_s);
        hipLaunchKernelGGL(kb_square, dim3(grid_for(b->dnnz, 256, b->cu * 8)), dim3(256), 0, b->stream, b->dnnz,
                           (const double *)b->dcv, sq_d);
        BT_LAUNCH(b);
        if ((rc = fill(b, b->u64, b->rows * b->K, 1.0)) || (rc = fill(b, b->v64, b->cols * b->K, 0.0))) return rc;
        std::swap(b->dcv, sq_d);
        rc = launch_cols<double>(b, b->shape[0], sq_s, b->u64, b->v64, b->sc_plain);
        std::swap(b->dcv, sq_d);
        BT_TRY(b, hipStreamSynchronize(b->stream));
        bfree(b, sq_s, (size_t)b->slots);
        bfree(b, sq_d, (size_t)std::max(b->dnnz, 1LL));
        if (rc) return rc;
        hipLaunchKernelGGL(kb_take_first, dim3(grid_for(b->cols, 256, b->cu * 8)), dim3(256), 0, b->stream, b->cols,
                           b->K, (const double *)b->v64, b->colss);
        BT_LAUNCH(b);
        std::vector<double> h((size_t)b->cols);
        BT_TRY(b, hipMemcpyAsync(h.data(), b->colss, sizeof(double) * b->cols, hipMemcpyDeviceToHost, b->stream));
        BT_TRY(b, hipStreamSynchronize(b->stream));
        b->fro_p.assign((size_t)b->P, 0.0);
        for (int p = 0; p < b->P; ++p) {
            double f = 0.0;
            for (long long c = b->col0[(size_t)p]; c < b->col0[(size_t)p] + b->np[(size_t)p]; ++c) f += h[(size_t)c];
            b->fro_p[(size_t)p] = std::sqrt(f);
        }
    }
    hipLaunchKernelGGL(kb_colscale, dim3(grid_for(b->cols, 256, b->cu * 8)), dim3(256), 0, b->stream, b->cols,
                       (const double *)b->colss, power, b->cs);
    BT_LAUNCH(b);
    b->cs_scaled = true;
    return PLSQR_OK;
}

// cs back to 1 (unscaled FP32 operator)
int clear_colscale(plsqr_batch *b) {
    if (!b->cs_scaled) return PLSQR_OK;
    int rc;
    if ((rc = fill(b, b->cs, b->cols, 1.0))) return rc;
    b->cs_scaled = false;
    return PLSQR_OK;
}

// Mixed precision per (system, right-hand side), as the single-system solve:
// a continuous inner FP32 LSQR on r = [b;0] - A x for every live pair, an
// FP64 check of xt = x + dx every refine_every launches (a pair stops there
// once its FP64 test2 = ||A^T r|| / (anorm ||r||) <= tol, anorm the largest
// inner estimate), and a restart on the refined residuals (x = xt) when
// every pair still live has stopped its inner run without meeting tol.
// With PLSQR_PRECOND_COLSCALE every problem's inner LSQR runs on A_p D_p
// and pair i's test2 takes anorm_refs[i] (or ||A_p||_F) as its anorm.
int solve_mixed(plsqr_batch *b, const double *b_dev, const plsqr_solve_opts *o, plsqr_solve_out *out, double *total_ms) {
    int rc;
    if ((rc = ensure_fp32(b))) return rc;
    if (!b->xt && (rc = balloc(b, &b->xt, (size_t)b->cols * b->K))) return rc;
    const bool colscale = o->precond == PLSQR_PRECOND_COLSCALE;
    if (o->precond != PLSQR_PRECOND_NONE && !colscale) return set_err(b, PLSQR_ERR_ARG, "unknown preconditioner");
    if ((rc = colscale ? set_colscale(b, o->precond_power == 0.0 ? 1.0 : o->precond_power) : clear_colscale(b)))
        return rc;
    // the ||A|| of the FP64 test2 of state (p, k) on a scaled solve
    auto anorm_fixed = [&](int p, int k) {
        const size_t i = (size_t)p * b->Kr + k;
        return i < b->anorm_refs.size() && b->anorm_refs[i] > 0.0 ? b->anorm_refs[i] : b->fro_p[(size_t)p];
    };
    const int itn = std::max(0, o->itn);
    const int period = o->refine_every > 0 ? o->refine_every : std::max(itn, 1);
    const int S = b->P * b->K;
    // adaptive checks: after the first, a chunk runs from 1 to 4 periods and
    // ends (after the first period) once a live pair's inner test2 estimate
    // times its last FP64 / inner ratio reaches tol
    const bool adaptive = o->refine_adaptive != 0 && o->tol > 0.0;
    bool have_ratio = false;
    std::vector<double> ratio((size_t)S, -1.0);
    std::vector<Scal> h((size_t)S);
    std::vector<double> anorm_max((size_t)S, 0.0), rn((size_t)S, 0.0), an((size_t)S, 0.0), t2((size_t)S, INFINITY);
    std::vector<double> rn0((size_t)S, 0.0), sums;
    std::vector<int> odone((size_t)S, 0), inner((size_t)S, 0), prev((size_t)S, 0), checks((size_t)S, 0),
        fin_ev((size_t)S, -1);
    for (int p = 0; p < b->P; ++p)
        for (int k = b->Kr; k < b->K; ++k) odone[(size_t)p * b->K + k] = 1;
    EventPair ev(b);
    if ((rc = ev.create())) return rc;
    const hipEvent_t e0 = ev.e0;
    BT_TRY(b, hipEventRecord(e0, b->stream));
    if ((rc = launch_load_b<double>(b, b_dev, b->b0)) || (rc = fill(b, b->x, b->cols * b->K, 0.0)) ||
        (rc = true_norms(b, b->x, sums)))
        return rc;
    std::vector<hipEvent_t> check_ev;
    size_t ev_next = 0;
    auto record = [&](int *idx) -> int {
        int r2;
        if ((r2 = ensure_events(b, ev_next + 1))) return r2;
        BT_TRY(b, hipEventRecord(b->evs[ev_next], b->stream));
        check_ev.push_back(b->evs[ev_next]);
        ++ev_next;
        *idx = (int)check_ev.size() - 1;
        return PLSQR_OK;
    };
    int ev_idx;
    if ((rc = record(&ev_idx))) return rc;
    for (int st = 0; st < S; ++st) {
        rn0[(size_t)st] = rn[(size_t)st] = std::sqrt(sums[(size_t)st]);
        if (odone[(size_t)st]) continue;
        if (rn[(size_t)st] == 0.0 || itn == 0) {
            odone[(size_t)st] = 1;
            t2[(size_t)st] = rn[(size_t)st] == 0.0 ? 0.0 : INFINITY;
            fin_ev[(size_t)st] = ev_idx;
        }
    }
    int total = 0;
    hipGraphExec_t exec = nullptr;
    const int one = 1;
    for (;;) {
        int live = 0;
        for (int st = 0; st < S; ++st) live += odone[(size_t)st] ? 0 : 1;
        if (live == 0 || total >= itn) break;
        // (re)start the inner runs of the live pairs on r (in u64)
        for (int st = 0; st < S; ++st) {
            h[(size_t)st] = fresh_state(o, itn - total);
            if (odone[(size_t)st]) h[(size_t)st].done = 1;
            prev[(size_t)st] = 0;
        }
        if ((rc = upload(b, b->sc, h.data(), h.size()))) return rc;
        hipLaunchKernelGGL((kb_convert<float>), dim3(grid_for(b->rows * b->K, 256, b->cu * 8)), dim3(256), 0, b->stream,
                           b->rows * b->K, (const double *)b->u64, b->u32);
        BT_LAUNCH(b);
        if ((rc = fill(b, b->v32, b->cols * b->K, 0.f)) || (rc = fill(b, b->dx, b->cols * b->K, 0.0)) ||
            (rc = fill(b, b->w, b->cols * b->K, 0.0)) ||
            (rc = enqueue_start<float>(b, b->shape[1], b->sval32, b->u32, b->v32)))
            return rc;
        if (!exec && (rc = get_graph<float>(b, 1, b->rv32, b->sval32, b->u32, b->v32, b->dx, &exec))) return rc;
        bool restart = false;
        while (!restart) {
            const bool pred = adaptive && have_ratio;
            const int k = std::min(pred ? 4 * period : period, itn - total);
            int launched = 0;
            if ((rc = run_launches(b, exec, k, ev_next, &launched, pred ? b->ratio_d : nullptr, o->tol, period)))
                return rc;
            ev_next += (size_t)launched;
            total += launched;
            if ((rc = read_states(b, h))) return rc;
            hipLaunchKernelGGL(kb_add3, dim3(grid_for(b->cols * b->K, 256, b->cu * 8)), dim3(256), 0, b->stream,
                               b->cols * b->K, (const double *)b->x, (const double *)b->dx, b->xt);
            BT_LAUNCH(b);
            if ((rc = true_norms(b, b->xt, sums)) || (rc = record(&ev_idx))) return rc;
            bool all_inner_done = true;
            live = 0;
            for (int st = 0; st < S; ++st) {
                if (odone[(size_t)st]) continue;
                const Scal &q = h[(size_t)st];
                inner[(size_t)st] += std::max(q.itn, 0) - prev[(size_t)st];
                prev[(size_t)st] = std::max(q.itn, 0);
                anorm_max[(size_t)st] = std::max(anorm_max[(size_t)st], q.anorm);
                rn[(size_t)st] = std::sqrt(sums[(size_t)st]);
                an[(size_t)st] = std::sqrt(sums[(size_t)S + st]);
                checks[(size_t)st] += 1;
                const double anorm_t = colscale ? anorm_fixed(st / b->K, st % b->K) : anorm_max[(size_t)st];
                t2[(size_t)st] = (rn[(size_t)st] > 0.0 && anorm_t > 0.0)
                                     ? an[(size_t)st] / (anorm_t * rn[(size_t)st])
                                     : 0.0;
                if (adaptive && q.test2 > 0.0 && std::isfinite(q.test2)) ratio[(size_t)st] = t2[(size_t)st] / q.test2;
                const bool stuck = q.done && q.itn == 0;  // the inner run could not start
                if (t2[(size_t)st] <= o->tol || total >= itn || stuck) {
                    odone[(size_t)st] = 1;
                    fin_ev[(size_t)st] = ev_idx;
                    // stop its inner run where it is: x + dx is its answer
                    BT_TRY(b, hipMemcpyAsync(&b->sc[st].done, &one, sizeof(int), hipMemcpyHostToDevice, b->stream));
                } else {
                    ++live;
                    if (!q.done) all_inner_done = false;
                }
            }
            if (adaptive) {
                if ((rc = upload(b, b->ratio_d, ratio.data(), ratio.size()))) return rc;
                have_ratio = true;
            }
            if (live == 0 || total >= itn) {
                restart = true;
            } else if (all_inner_done || launched == 0) {
                restart = true;  // refine: x = x + dx, r (in u64) is its residual
            }
        }
        // x = x + dx for every pair (finished pairs keep their final x + dx)
        BT_TRY(b, hipMemcpyAsync(b->x, b->xt, sizeof(double) * b->cols * b->K, hipMemcpyDeviceToDevice, b->stream));
    }
    BT_TRY(b, hipStreamSynchronize(b->stream));
    float t_all = 0.f;
    for (size_t i = 0; i < check_ev.size(); ++i) t_all = std::max(t_all, elapsed(e0, check_ev[i]));
    for (int p = 0; p < b->P; ++p)
        for (int k = 0; k < b->Kr; ++k) {
            const size_t st = (size_t)p * b->K + k;
            plsqr_solve_out &r = out[(size_t)p * b->Kr + k];
            memset(&r, 0, sizeof(r));
            r.itn = inner[st];
            r.inner_total = inner[st];
            r.refinements = checks[st];
            r.anorm = colscale ? anorm_fixed(p, k) : anorm_max[st];
            r.rnorm = rn[st];
            r.bnorm = rn0[st];
            r.arnorm = t2[st];
            r.arnorm_true = an[st];
            r.istop = t2[st] <= o->tol ? 2 : 5;
            const int fe = fin_ev[st];
            r.elapsed_ms = fe >= 0 && (size_t)fe < check_ev.size() ? elapsed(e0, check_ev[(size_t)fe]) : t_all;
        }
    if (total_ms) *total_ms = t_all;
    return clear_colscale(b);
}

int check_handle(plsqr_batch *b) {
    if (!b) return PLSQR_ERR_ARG;
    if (b->allocs.empty()) return set_err(b, PLSQR_ERR_STATE, "batch failed to initialise");
    return PLSQR_OK;
}

int shape_from_tuning(const plsqr_batch_tuning *t, Shape *s) {
    const int rpw = t->rows_per_wave;
    if (rpw != 2 && rpw != 4 && rpw != 8 && rpw != 16) return PLSQR_ERR_ARG;
    if (t->nnz_per_lane != 2 && t->nnz_per_lane != 4) return PLSQR_ERR_ARG;
    if (t->block_rows != 128 && t->block_rows != 256) return PLSQR_ERR_ARG;
    if (t->block_cols != 128 && t->block_cols != 256 && t->block_cols != 512) return PLSQR_ERR_ARG;
    if (t->gather_kib != 0 && t->gather_kib != 32 && t->gather_kib != 64) return PLSQR_ERR_ARG;
    s->L = 64 / rpw;
    s->U = t->nnz_per_lane;
    s->bs1 = t->block_rows;
    s->bs2 = t->block_cols;
    s->kib2 = t->gather_kib ? t->gather_kib : 32;
    return PLSQR_OK;
}

void tuning_from_shape(const Shape &s, plsqr_batch_tuning *t) {
    t->rows_per_wave = 64 / s.L;
    t->nnz_per_lane = s.U;
    t->block_rows = s.bs1;
    t->block_cols = s.bs2;
    t->ms_rows = s.ms1;
    t->ms_cols = s.ms2;
    t->candidates = s.candidates;
    t->gather_kib = s.kib2;
}

// median time per launch of fn over reps launches (after one warm-up)
template <class F>
int time_launches(plsqr_batch *b, int reps, F &&fn, double *ms) {
    int rc;
    if ((rc = fn())) return rc;
    if ((rc = ensure_events(b, (size_t)reps + 1))) return rc;
    BT_TRY(b, hipEventRecord(b->evs[0], b->stream));
    for (int i = 0; i < reps; ++i) {
        if ((rc = fn())) return rc;
        BT_TRY(b, hipEventRecord(b->evs[(size_t)i + 1], b->stream));
    }
    BT_TRY(b, hipStreamSynchronize(b->stream));
    std::vector<double> t((size_t)reps);
    for (int i = 0; i < reps; ++i) t[(size_t)i] = elapsed(b->evs[(size_t)i], b->evs[(size_t)i + 1]);
    std::sort(t.begin(), t.end());
    *ms = t[(size_t)reps / 2];
    return PLSQR_OK;
}

template <class T>
int autotune(plsqr_batch *b, int prec, int reps, const T *av, const T *sv, T *u, T *v) {
    int rc;
    if ((rc = fill<T>(b, u, b->rows * b->K, T(1))) || (rc = fill<T>(b, v, b->cols * b->K, T(1)))) return rc;
    Shape best = b->shape[prec];
    double best1 = INFINITY, best2 = INFINITY;
    int n = 0;
    for (int L : {4, 8, 16, 32})
        for (int bs : {128, 256})
            for (int U : {2, 4}) {
                Shape s = best;
                s.L = L;
                s.bs1 = bs;
                s.U = U;
                double ms;
                // restart from y = 1 so every candidate sees the same values
                if ((rc = fill<T>(b, u, b->rows * b->K, T(1)))) return rc;
                if ((rc = time_launches(b, reps, [&] { return launch_rows<T>(b, s, av, v, u, b->sc_plain); }, &ms)))
                    return rc;
                ++n;
                if (ms < best1) {
                    best1 = ms;
                    best.L = L;
                    best.bs1 = bs;
                    best.U = U;
                }
            }
    for (int bs : {128, 256, 512})
        for (int kib : {32, 64}) {
            Shape s = best;
            s.bs2 = bs;
            s.kib2 = kib;
            double ms;
            if ((rc = fill<T>(b, v, b->cols * b->K, T(1)))) return rc;
            if ((rc = time_launches(b, reps, [&] { return launch_cols<T>(b, s, sv, u, v, b->sc_plain); }, &ms)))
                return rc;
            ++n;
            if (ms < best2) {
                best2 = ms;
                best.bs2 = bs;
                best.kib2 = kib;
            }
        }
    best.ms1 = best1;
    best.ms2 = best2;
    best.candidates = n;
    b->shape[prec] = best;
    GraphCache &g = b->graphs[prec];
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
    return PLSQR_OK;
}

int build_batch(plsqr_batch *b, plsqr::BatchPlan &plan) {
    int rc;
    const int P = plan.P;
    b->P = P;
    b->rows = plan.rows;
    b->cols = plan.cols;
    b->nnz = plan.nnz;
    b->dnnz = plan.dnnz;
    b->slots = plan.slots;
    b->segs = plan.segs;
    b->cuts = plan.cuts;
    b->items = plan.items;
    b->nblk = (int)plan.blk_item.size();
    b->nblk_r = (int)plan.blk_rch.size();
    b->partials = plan.partials;
    b->nrch = (int)plan.rch_prob.size();
    b->ncch = (int)plan.cch_prob.size();
    b->mk.resize((size_t)P);
    b->md.resize((size_t)P);
    b->np.resize((size_t)P);
    b->row0 = plan.row0;
    b->col0 = plan.col0;
    b->xb.assign((size_t)P + 1, 0);
    b->yb.assign((size_t)P + 1, 0);
    b->bb.assign((size_t)P + 1, 0);
    for (int p = 0; p < P; ++p) {
        const plsqr::ProblemLayout &L = plan.prob[(size_t)p];
        b->mk[(size_t)p] = L.m_k;
        b->md[(size_t)p] = L.m_d;
        b->np[(size_t)p] = L.n;
        b->xb[(size_t)p + 1] = b->xb[(size_t)p] + (long long)b->Kr * L.n;
        b->yb[(size_t)p + 1] = b->yb[(size_t)p] + (long long)b->Kr * (L.m_k + L.m_d);
        b->bb[(size_t)p + 1] = b->bb[(size_t)p] + (long long)b->Kr * L.m_k;
    }
    const int K = b->K;
    if ((rc = balloc(b, &b->rp, (size_t)b->rows + 1)) || (rc = balloc(b, &b->ci, (size_t)b->nnz)) ||
        (rc = balloc(b, &b->rv, (size_t)b->nnz)) || (rc = balloc(b, &b->dcp, (size_t)b->cols + 1)) ||
        (rc = balloc(b, &b->dri, (size_t)b->dnnz)) || (rc = balloc(b, &b->dcv, (size_t)b->dnnz)) ||
        (rc = balloc(b, &b->i_prob, (size_t)b->items)) || (rc = balloc(b, &b->i_c0, (size_t)b->items)) ||
        (rc = balloc(b, &b->i_c1, (size_t)b->items)) || (rc = balloc(b, &b->i_pbase, (size_t)b->items)) ||
        (rc = balloc(b, &b->blk_item, (size_t)b->nblk)) || (rc = balloc(b, &b->blk_rch, (size_t)b->nblk_r)) || (rc = balloc(b, &b->pp, (size_t)b->partials * K)) ||
        (rc = balloc(b, &b->d_pofs, (size_t)P + 1)) || (rc = balloc(b, &b->d_nslice, (size_t)P)) ||
        (rc = balloc(b, &b->slot_base, (size_t)b->items + 1)) || (rc = balloc(b, &b->seg_start, (size_t)b->items + 1)) ||
        (rc = balloc(b, &b->vox_len, (size_t)b->items * TW)) || (rc = balloc(b, &b->seg_ray, (size_t)b->segs)) ||
        (rc = balloc(b, &b->sval, (size_t)b->slots)) || (rc = balloc(b, &b->sseg, (size_t)b->slots)) ||
        (rc = balloc(b, &b->cut_base, (size_t)b->items + 1)) || (rc = balloc(b, &b->vox_cut, (size_t)b->cuts * TW)) ||
        (rc = balloc(b, &b->rch_row, (size_t)b->nrch + 1)) || (rc = balloc(b, &b->rch_prob, (size_t)b->nrch)) ||
        (rc = balloc(b, &b->rch_kind, (size_t)b->nrch)) || (rc = balloc(b, &b->cch_col, (size_t)b->ncch + 1)) ||
        (rc = balloc(b, &b->cch_prob, (size_t)b->ncch)) || (rc = balloc(b, &b->r_rch, (size_t)P + 1)) ||
        (rc = balloc(b, &b->r_cch, (size_t)P + 1)) ||
        (rc = balloc(b, &b->row_prob, (size_t)b->rows)) || (rc = balloc(b, &b->row_old, (size_t)b->rows)) ||
        (rc = balloc(b, &b->col_prob, (size_t)b->cols)) ||
        (rc = balloc(b, &b->col_old, (size_t)b->cols)) || (rc = balloc(b, &b->d_row0, (size_t)P + 1)) ||
        (rc = balloc(b, &b->d_col0, (size_t)P + 1)) || (rc = balloc(b, &b->d_xb, (size_t)P + 1)) ||
        (rc = balloc(b, &b->d_yb, (size_t)P + 1)) || (rc = balloc(b, &b->d_bb, (size_t)P + 1)) ||
        (rc = balloc(b, &b->d_mk, (size_t)P)) || (rc = balloc(b, &b->d_mrow, (size_t)P)) ||
        (rc = balloc(b, &b->d_ncol, (size_t)P)))
        return rc;
    b->rtiled = env_int("PLSQR_BATCH_ROWTILES", 1) != 0;
    b->groups = plan.groups;
    b->ents = plan.ents;
    if (b->rtiled &&
        ((rc = balloc(b, &b->rt_gbase, (size_t)b->items + 1)) || (rc = balloc(b, &b->rt_ebase, (size_t)b->groups + 1)) ||
         (rc = balloc(b, &b->rt_len, (size_t)b->groups * plsqr::RowTiles::G)) ||
         // entries carry 4 of slack: the 16-byte staging loads may read past the last group
         (rc = balloc(b, &b->rt_vox, (size_t)b->ents + 4)) || (rc = balloc(b, &b->rt_val, (size_t)b->ents + 4)) ||
         (rc = balloc(b, &b->rt_rptr, (size_t)b->rows + 1)) || (rc = balloc(b, &b->rt_rseg, (size_t)b->segs)) ||
         (rc = balloc(b, &b->segv, (size_t)b->segs * K))))
        return rc;
    // per-problem pieces, shifted to global numbering on the host and
    // uploaded one problem at a time (the host copy is released after)
    for (int p = 0; p < P; ++p) {
        plsqr::ProblemLayout &L = plan.prob[(size_t)p];
        const long long r0 = plan.row0[(size_t)p], c0 = plan.col0[(size_t)p];
        const int m = L.m_k + L.m_d;
        std::vector<int> ti((size_t)m);
        for (int i = 0; i < m; ++i) ti[(size_t)i] = (int)(plan.nnz0[(size_t)p] + L.rp[(size_t)i]);
        if ((rc = upload(b, b->rp + r0, ti.data(), ti.size()))) return rc;
        for (auto &c : L.ci) c += (int)c0;
        if ((rc = upload(b, b->ci + plan.nnz0[(size_t)p], L.ci.data(), L.ci.size())) ||
            (rc = upload(b, b->rv + plan.nnz0[(size_t)p], L.rv.data(), L.rv.size())))
            return rc;
        ti.resize((size_t)L.n);
        for (int c = 0; c < L.n; ++c) ti[(size_t)c] = (int)(plan.dnnz0[(size_t)p] + L.dcp[(size_t)c]);
        if ((rc = upload(b, b->dcp + c0, ti.data(), ti.size()))) return rc;
        for (auto &r : L.dri) r += (int)(r0 + L.m_k);
        if ((rc = upload(b, b->dri + plan.dnnz0[(size_t)p], L.dri.data(), L.dri.size())) ||
            (rc = upload(b, b->dcv + plan.dnnz0[(size_t)p], L.dcv.data(), L.dcv.size())))
            return rc;
        const plsqr::TileLayout &T = L.tiles;
        const int t0 = plan.item0[(size_t)p];  // first item of the problem
        const int nit = T.T * T.S;
        std::vector<int> tp((size_t)nit, p), ic0((size_t)nit), ic1((size_t)nit);
        std::vector<long long> sb((size_t)nit), sg((size_t)nit), cb((size_t)nit), pbase((size_t)nit);
        for (int t = 0; t < T.T; ++t)
            for (int q = 0; q < T.S; ++q) {
                const size_t it = (size_t)t * T.S + q;
                ic0[it] = (int)(c0 + T.tile_start[(size_t)t]);
                ic1[it] = (int)(c0 + T.tile_start[(size_t)t + 1]);
                pbase[it] = plan.pofs[(size_t)p] + (long long)q * L.n - c0;
                sb[it] = plan.slot0[(size_t)p] + T.slot_base[it];
                sg[it] = plan.seg0[(size_t)p] + T.seg_start[it];
                cb[it] = plan.cut0[(size_t)p] + T.cut_base[it];
            }
        std::vector<int> ray(T.seg_ray.size());
        for (size_t i = 0; i < ray.size(); ++i) ray[i] = (int)(r0 + T.seg_ray[i]);
        if ((rc = upload(b, b->i_prob + t0, tp.data(), tp.size())) || (rc = upload(b, b->i_c0 + t0, ic0.data(), ic0.size())) ||
            (rc = upload(b, b->i_c1 + t0, ic1.data(), ic1.size())) ||
            (rc = upload(b, b->i_pbase + t0, pbase.data(), pbase.size())) ||
            (rc = upload(b, b->slot_base + t0, sb.data(), sb.size())) ||
            (rc = upload(b, b->seg_start + t0, sg.data(), sg.size())) ||
            (rc = upload(b, b->cut_base + t0, cb.data(), cb.size())) ||
            (rc = upload(b, b->vox_cut + (size_t)plan.cut0[(size_t)p] * TW, T.vox_cut.data(), T.vox_cut.size())) ||
            (rc = upload(b, b->vox_len + (size_t)t0 * TW, T.vox_len.data(), T.vox_len.size())) ||
            (rc = upload(b, b->seg_ray + plan.seg0[(size_t)p], ray.data(), ray.size())) ||
            (rc = upload(b, b->sval + plan.slot0[(size_t)p], T.sval.data(), T.sval.size())) ||
            (rc = upload(b, b->sseg + plan.slot0[(size_t)p], T.sseg.data(), T.sseg.size())))
            return rc;
        if (b->rtiled) {
            // mode-1 segment layout of this problem's tiles, shifted to global numbering
            plsqr::RowTiles R;
            std::string why;
            if (!plsqr::build_row_tiles(T, L.m_k, &R, &why)) return set_err(b, PLSQR_ERR_FORMAT, why);
            if ((long long)R.n_grp != plan.grp0[(size_t)p + 1] - plan.grp0[(size_t)p] ||
                R.n_ent != plan.ent0[(size_t)p + 1] - plan.ent0[(size_t)p])
                return set_err(b, PLSQR_ERR_STATE, "mode-1 segment layout disagrees with the plan");
            std::vector<int> gb((size_t)nit), rpt((size_t)m), rsg(R.ray_seg.size());
            std::vector<long long> eb((size_t)R.n_grp);
            for (int i = 0; i < nit; ++i) gb[(size_t)i] = (int)(plan.grp0[(size_t)p] + R.grp_base[(size_t)i]);
            for (int g = 0; g < R.n_grp; ++g) eb[(size_t)g] = plan.ent0[(size_t)p] + R.ent_base[(size_t)g];
            for (int i = 0; i < m; ++i)
                rpt[(size_t)i] = (int)(plan.seg0[(size_t)p] + R.ray_ptr[(size_t)std::min(i, L.m_k)]);
            for (size_t j = 0; j < rsg.size(); ++j) rsg[j] = (int)(plan.seg0[(size_t)p] + R.ray_seg[j]);
            if ((rc = upload(b, b->rt_gbase + t0, gb.data(), gb.size())) ||
                (rc = upload(b, b->rt_ebase + plan.grp0[(size_t)p], eb.data(), eb.size())) ||
                (rc = upload(b, b->rt_len + (size_t)plan.grp0[(size_t)p] * plsqr::RowTiles::G, R.seg_len.data(),
                             R.seg_len.size())) ||
                (rc = upload(b, b->rt_vox + plan.ent0[(size_t)p], R.ent_vox.data(), R.ent_vox.size())) ||
                (rc = upload(b, b->rt_val + plan.ent0[(size_t)p], R.ent_val.data(), R.ent_val.size())) ||
                (rc = upload(b, b->rt_rptr + r0, rpt.data(), rpt.size())) ||
                (rc = upload(b, b->rt_rseg + plan.seg0[(size_t)p], rsg.data(), rsg.size())))
                return rc;
            BT_TRY(b, hipStreamSynchronize(b->stream));
        }
        std::vector<int> rpb((size_t)m, p), cpb((size_t)L.n, p), rold((size_t)m);
        for (int i = 0; i < m; ++i) rold[(size_t)i] = i < L.m_k ? L.rperm[(size_t)i] : i;
        if ((rc = upload(b, b->row_prob + r0, rpb.data(), rpb.size())) ||
            (rc = upload(b, b->row_old + r0, rold.data(), rold.size())) ||
            (rc = upload(b, b->col_prob + c0, cpb.data(), cpb.size())) ||
            (rc = upload(b, b->col_old + c0, L.inv.data(), L.inv.size())))
            return rc;
        // the copies above are asynchronous from pageable memory: wait before
        // the host vectors are released
        BT_TRY(b, hipStreamSynchronize(b->stream));
        L = plsqr::ProblemLayout();
    }
    const int rows_i = (int)b->rows, cols_i = (int)b->cols;
    const int nnz_i = (int)b->nnz, dnnz_i = (int)b->dnnz;
    std::vector<long long> tail_sb{b->slots}, tail_sg{b->segs}, tail_cb{b->cuts};
    std::vector<int> mrow((size_t)P), nsl((size_t)P);
    for (int p = 0; p < P; ++p) {
        mrow[(size_t)p] = b->mk[(size_t)p] + b->md[(size_t)p];
        nsl[(size_t)p] = plan.nslice[(size_t)p];
    }
    if ((rc = upload(b, b->rp + rows_i, &nnz_i, 1)) || (rc = upload(b, b->dcp + cols_i, &dnnz_i, 1)) ||
        (rc = upload(b, b->slot_base + b->items, tail_sb.data(), 1)) ||
        (rc = upload(b, b->seg_start + b->items, tail_sg.data(), 1)) ||
        (rc = upload(b, b->cut_base + b->items, tail_cb.data(), 1)) ||
        (rc = upload(b, b->blk_item, plan.blk_item.data(), plan.blk_item.size())) ||
        (rc = upload(b, b->blk_rch, plan.blk_rch.data(), plan.blk_rch.size())) ||
        (rc = upload(b, b->d_pofs, plan.pofs.data(), plan.pofs.size())) ||
        (rc = upload(b, b->rch_row, plan.rch_row.data(), plan.rch_row.size())) ||
        (rc = upload(b, b->rch_prob, plan.rch_prob.data(), plan.rch_prob.size())) ||
        (rc = upload(b, b->rch_kind, plan.rch_kind.data(), plan.rch_kind.size())) ||
        (rc = upload(b, b->cch_col, plan.cch_col.data(), plan.cch_col.size())) ||
        (rc = upload(b, b->cch_prob, plan.cch_prob.data(), plan.cch_prob.size())) ||
        (rc = upload(b, b->r_rch, plan.rch0.data(), plan.rch0.size())) ||
        (rc = upload(b, b->r_cch, plan.cch0.data(), plan.cch0.size())) ||
        (rc = upload(b, b->d_row0, b->row0.data(), b->row0.size())) ||
        (rc = upload(b, b->d_col0, b->col0.data(), b->col0.size())) ||
        (rc = upload(b, b->d_xb, b->xb.data(), b->xb.size())) || (rc = upload(b, b->d_yb, b->yb.data(), b->yb.size())) ||
        (rc = upload(b, b->d_bb, b->bb.data(), b->bb.size())) || (rc = upload(b, b->d_mk, b->mk.data(), b->mk.size())) ||
        (rc = upload(b, b->d_mrow, mrow.data(), mrow.size())) || (rc = upload(b, b->d_ncol, b->np.data(), b->np.size())) ||
        (rc = upload(b, b->d_nslice, nsl.data(), nsl.size())))
        return rc;
    if (b->rtiled) {
        const int groups_i = (int)b->groups, segs_i = (int)b->segs;
        const long long ents_l = b->ents;
        if ((rc = upload(b, b->rt_gbase + b->items, &groups_i, 1)) || (rc = upload(b, b->rt_ebase + b->groups, &ents_l, 1)) ||
            (rc = upload(b, b->rt_rptr + rows_i, &segs_i, 1)))
            return rc;
    }
    BT_TRY(b, hipStreamSynchronize(b->stream));
    // vectors and scratch
    b->part_len = std::max(b->nrch, b->ncch) * K;
    b->stage_len = std::max(b->yb[(size_t)P], b->xb[(size_t)P]);
    const int S = P * K;
    if ((rc = balloc(b, &b->u64, (size_t)b->rows * K)) || (rc = balloc(b, &b->v64, (size_t)b->cols * K)) ||
        (rc = balloc(b, &b->b0, (size_t)b->rows * K)) || (rc = balloc(b, &b->x, (size_t)b->cols * K)) ||
        (rc = balloc(b, &b->w, (size_t)b->cols * K)) || (rc = balloc(b, &b->dx, (size_t)b->cols * K)) ||
        (rc = balloc(b, &b->part, (size_t)b->part_len)) || (rc = balloc(b, &b->sums, (size_t)S * 2)) ||
        (rc = balloc(b, &b->stage, (size_t)b->stage_len)) || (rc = balloc(b, &b->sc, (size_t)S)) ||
        (rc = balloc(b, &b->sc_plain, (size_t)S)) || (rc = balloc(b, &b->live_d, 2)) ||
        (rc = balloc(b, &b->ratio_d, (size_t)S)))
        return rc;
    BT_TRY(b, hipHostMalloc((void **)&b->live_h, sizeof(int) * 2));
    b->live_h[0] = b->live_h[1] = 0;
    BT_TRY(b, hipHostMalloc(&b->hstage, std::max(sizeof(Scal), 2 * sizeof(double)) * (size_t)S));
    // states of a plain operator application: y*1*1 + A x, v*1 + A^T u
    Scal z;
    memset(&z, 0, sizeof(z));
    z.ib = 1.0;
    z.alpha = -1.0;
    z.vscale = 1.0;
    z.beta_pos = 1;
    std::vector<Scal> hz((size_t)S, z);
    if ((rc = upload(b, b->sc_plain, hz.data(), hz.size()))) return rc;
    BT_TRY(b, hipStreamSynchronize(b->stream));
    return PLSQR_OK;
}

}  // namespace

extern "C" {

const char *plsqr_batch_last_error(const plsqr_batch *b) {
    if (!b) return g_batch_err.c_str();
    return b->err.c_str();
}

void plsqr_batch_clear_error(plsqr_batch *b) {
    if (b) b->err.clear();
}

int plsqr_batch_create(const plsqr_system *const *systems, const plsqr_grid *grids, int n_problems, int n_rhs,
                       int flags, plsqr_batch **out) {
    g_batch_err.clear();
    if (!out) return PLSQR_ERR_ARG;
    *out = nullptr;
    if (n_problems <= 0 || !systems) {
        g_batch_err = "no systems";
        return PLSQR_ERR_ARG;
    }
    if (n_rhs < 1 || n_rhs > PLSQR_BATCH_MAX_RHS) {
        g_batch_err = "n_rhs must be 1.." + std::to_string(PLSQR_BATCH_MAX_RHS);
        return PLSQR_ERR_ARG;
    }
    plsqr::BatchPlan plan;
    std::string why;
    const int Kp = n_rhs == 1 ? 1 : (n_rhs == 2 ? 2 : 4);
    // rows per slice: the slice's u (Kp FP64 values per row) in 2 MB, half an XCD's L2
    const int slice_rows = std::max(4096, env_int("PLSQR_BATCH_SLICE_BYTES", 2 << 20) / (Kp * 8));
    int rc = plsqr::build_batch_plan(systems, grids, n_problems, TW, ROW_CHUNK, DAMP_CHUNK, COL_CHUNK, slice_rows,
                                     8, &plan, &why);
    if (rc) {
        g_batch_err = why;
        return rc;
    }
    plsqr_batch *b = new plsqr_batch();
    b->Kr = n_rhs;
    b->K = Kp;
    auto fail = [&](int code) {
        g_batch_err = b->err;
        plsqr_batch_destroy(b);
        return code;
    };
    hipError_t e = hipGetDevice(&b->dev);
    if (e != hipSuccess) return fail(set_err(b, PLSQR_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e)));
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, b->dev);
    if (e != hipSuccess)
        return fail(set_err(b, PLSQR_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e)));
    b->cu = std::max(1, prop.multiProcessorCount);
    e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(set_err(b, PLSQR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    if ((rc = build_batch(b, plan))) return fail(rc);
    if (!(flags & PLSQR_BATCH_NO_TUNE) && env_int("PLSQR_BATCH_TUNE", 1) != 0) {
        plsqr_batch_tuning t;
        if ((rc = plsqr_batch_autotune(b, PLSQR_PREC_FP64, 5, &t))) return fail(rc);
        if ((flags & PLSQR_BATCH_TUNE_MIXED) && (rc = plsqr_batch_autotune(b, PLSQR_PREC_MIXED, 5, &t)))
            return fail(rc);
    }
    *out = b;
    return PLSQR_OK;
}

void plsqr_batch_destroy(plsqr_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (hipEvent_t e : b->evs) (void)hipEventDestroy(e);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->live_h) (void)hipHostFree(b->live_h);
    if (b->hstage) (void)hipHostFree(b->hstage);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int plsqr_batch_dims(const plsqr_batch *b, int *n_problems, int *n_rhs, long long *rows, long long *cols,
                     long long *nnz) {
    if (!b) return PLSQR_ERR_ARG;
    if (n_problems) *n_problems = b->P;
    if (n_rhs) *n_rhs = b->Kr;
    if (rows) *rows = b->rows;
    if (cols) *cols = b->cols;
    if (nnz) *nnz = b->nnz;
    return PLSQR_OK;
}

size_t plsqr_batch_device_bytes(const plsqr_batch *b) { return b ? b->bytes : 0; }

int plsqr_batch_autotune(plsqr_batch *b, int precision, int reps, plsqr_batch_tuning *out) {
    int rc;
    if ((rc = check_handle(b))) return rc;
    reps = std::max(1, std::min(reps, 1000));
    if (precision == PLSQR_PREC_FP64) {
        rc = autotune<double>(b, 0, reps, b->rv, b->sval, b->u64, b->v64);
    } else if (precision == PLSQR_PREC_MIXED) {
        if ((rc = ensure_fp32(b))) return rc;
        rc = autotune<float>(b, 1, reps, b->rv32, b->sval32, b->u32, b->v32);
    } else {
        return set_err(b, PLSQR_ERR_ARG, "unknown precision");
    }
    if (rc) return rc;
    BT_TRY(b, hipStreamSynchronize(b->stream));
    if (out) tuning_from_shape(b->shape[precision == PLSQR_PREC_MIXED ? 1 : 0], out);
    return PLSQR_OK;
}

int plsqr_batch_get_tuning(const plsqr_batch *b, int precision, plsqr_batch_tuning *out) {
    if (!b || !out || (precision != PLSQR_PREC_FP64 && precision != PLSQR_PREC_MIXED)) return PLSQR_ERR_ARG;
    tuning_from_shape(b->shape[precision == PLSQR_PREC_MIXED ? 1 : 0], out);
    return PLSQR_OK;
}

int plsqr_batch_set_tuning(plsqr_batch *b, int precision, const plsqr_batch_tuning *t) {
    int rc;
    if ((rc = check_handle(b))) return rc;
    if (!t || (precision != PLSQR_PREC_FP64 && precision != PLSQR_PREC_MIXED))
        return set_err(b, PLSQR_ERR_ARG, "bad tuning arguments");
    Shape s;
    if (shape_from_tuning(t, &s)) return set_err(b, PLSQR_ERR_ARG, "unsupported launch shape");
    const int i = precision == PLSQR_PREC_MIXED ? 1 : 0;
    b->shape[i] = s;
    GraphCache &g = b->graphs[i];
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
    return PLSQR_OK;
}

int plsqr_batch_aprod(plsqr_batch *b, int mode, int precision, double *x, double *y) {
    int rc;
    if ((rc = check_handle(b))) return rc;
    if (mode != 1 && mode != 2) return set_err(b, PLSQR_ERR_ARG, "mode must be 1 or 2");
    if (!x || !y) return set_err(b, PLSQR_ERR_ARG, "x and y are required");
    const long long ny = b->yb[(size_t)b->P], nx = b->xb[(size_t)b->P];
    const bool mixed = precision == PLSQR_PREC_MIXED;
    if (!mixed && precision != PLSQR_PREC_FP64) return set_err(b, PLSQR_ERR_ARG, "unknown precision");
    if (mixed && ((rc = ensure_fp32(b)) || (rc = clear_colscale(b)))) return rc;
    const int pi = mixed ? 1 : 0;
    // both operands to the device
    if ((rc = upload(b, b->stage, x, (size_t)nx))) return rc;
    rc = mixed ? launch_conv<float>(b, false, 0, b->stage, b->v32) : launch_conv<double>(b, false, 0, b->stage, b->v64);
    if (rc) return rc;
    if ((rc = upload(b, b->stage, y, (size_t)ny))) return rc;
    rc = mixed ? launch_conv<float>(b, true, 0, b->stage, b->u32) : launch_conv<double>(b, true, 0, b->stage, b->u64);
    if (rc) return rc;
    if (mode == 1) {
        rc = mixed ? launch_rows<float>(b, b->shape[pi], b->rv32, b->v32, b->u32, b->sc_plain)
                   : launch_rows<double>(b, b->shape[pi], b->rv, b->v64, b->u64, b->sc_plain);
        if (rc) return rc;
        rc = mixed ? launch_conv<float>(b, true, 1, b->stage, b->u32) : launch_conv<double>(b, true, 1, b->stage, b->u64);
        if (rc) return rc;
        BT_TRY(b, hipMemcpyAsync(y, b->stage, sizeof(double) * ny, hipMemcpyDeviceToHost, b->stream));
    } else {
        rc = mixed ? launch_cols<float>(b, b->shape[pi], b->sval32, b->u32, b->v32, b->sc_plain)
                   : launch_cols<double>(b, b->shape[pi], b->sval, b->u64, b->v64, b->sc_plain);
        if (rc) return rc;
        rc = mixed ? launch_conv<float>(b, false, 1, b->stage, b->v32) : launch_conv<double>(b, false, 1, b->stage, b->v64);
        if (rc) return rc;
        BT_TRY(b, hipMemcpyAsync(x, b->stage, sizeof(double) * nx, hipMemcpyDeviceToHost, b->stream));
    }
    BT_TRY(b, hipStreamSynchronize(b->stream));
    return PLSQR_OK;
}

int plsqr_batch_set_anorm_refs(plsqr_batch *b, const double *refs) {
    int rc;
    if ((rc = check_handle(b))) return rc;
    if (!refs) {
        b->anorm_refs.clear();
        return PLSQR_OK;
    }
    b->anorm_refs.assign(refs, refs + (size_t)b->P * b->Kr);
    return PLSQR_OK;
}

int plsqr_batch_solve_device(plsqr_batch *b, const double *b_dev, const plsqr_solve_opts *opts, double *x_dev,
                             plsqr_solve_out *out, double *total_ms) {
    int rc;
    if ((rc = check_handle(b))) return rc;
    if (!b_dev || !opts || !x_dev || !out) return set_err(b, PLSQR_ERR_ARG, "b, opts, x and out are required");
    plsqr_solve_opts o;
    default_opts(opts, &o);
    if (o.precond != PLSQR_PRECOND_NONE && o.precision != PLSQR_PREC_MIXED)
        return set_err(b, PLSQR_ERR_ARG, "column scaling is offered on the mixed-precision solve");
    if (o.precision == PLSQR_PREC_FP64) {
        rc = solve_fp64(b, b_dev, &o, out, total_ms);
    } else if (o.precision == PLSQR_PREC_MIXED) {
        rc = solve_mixed(b, b_dev, &o, out, total_ms);
    } else {
        return set_err(b, PLSQR_ERR_ARG, "unknown precision");
    }
    if (rc) return rc;
    if ((rc = launch_conv<double>(b, false, 1, x_dev, b->x))) return rc;
    BT_TRY(b, hipStreamSynchronize(b->stream));
    return PLSQR_OK;
}

int plsqr_batch_solve(plsqr_batch *b, const double *b_host, const plsqr_solve_opts *opts, double *x_host,
                      plsqr_solve_out *out, double *total_ms) {
    int rc;
    if ((rc = check_handle(b))) return rc;
    if (!b_host || !opts || !x_host || !out) return set_err(b, PLSQR_ERR_ARG, "b, opts, x and out are required");
    const long long nb = b->bb[(size_t)b->P], nx = b->xb[(size_t)b->P];
    for (long long i = 0; i < nb; ++i)
        if (!std::isfinite(b_host[i]))
            return set_err(b, PLSQR_ERR_NONFINITE, "b holds a NaN or Inf at " + std::to_string(i));
    // b is staged in the v64/u64-independent staging buffer (nb <= stage_len)
    if ((rc = upload(b, b->stage, b_host, (size_t)nb))) return rc;
    BT_TRY(b, hipStreamSynchronize(b->stream));
    // the solve loads b into u (or b0) before it uses the staging buffer again
    if ((rc = plsqr_batch_solve_device(b, b->stage, opts, b->stage, out, total_ms))) return rc;
    BT_TRY(b, hipMemcpyAsync(x_host, b->stage, sizeof(double) * nx, hipMemcpyDeviceToHost, b->stream));
    BT_TRY(b, hipStreamSynchronize(b->stream));
    return PLSQR_OK;
}

}  // extern "C"
