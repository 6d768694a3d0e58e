// C-ABI of the LSQR hot path: operator handle, aprod entry points and the
// device-resident LSQR loop (include/plsqr_hip.h).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "plsqr_hip.h"
#include "plsqr_kernels.h"
#include "plsqr_spmv.h"
#include "plsqr_mixed.h"
#include "tile_build.h"

using plsqr::Scal;

struct plsqr_op {
    int dev = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int m_k = 0, m_d = 0, n = 0, m = 0;
    long long nnz_k = 0, nnz_d = 0;
    int flags = 0;
    bool exact = false;
    // stacked CSR of [A_k ; D]: kernel rows re-laid on the device, then damping rows
    int *rp = nullptr, *ci = nullptr;
    double *rv = nullptr;
    // kernel CSC as loaded (kept for the exact-order kernels)
    int *kcp = nullptr, *kri = nullptr;
    double *kv = nullptr;
    // kernel CSC split into row slices (one per XCD)
    int nsl = 1;
    int *scp = nullptr, *sri = nullptr;
    double *sv = nullptr;
    int *bounds = nullptr;
    // damping CSC by column
    int *dcp = nullptr, *dri = nullptr;
    double *dcv = nullptr;
    // the same columns interleaved by groups of 64 (tiled operator): column c's
    // k-th entry at dpos[c] + 64 k, dlen[c] entries; NULL when not built
    int *dpos = nullptr, *dri_il = nullptr;
    double *dcv_il = nullptr;
    uint8_t *dlen = nullptr;
    long long d_il = 0;  // slots of the interleaved arrays
    // vectors
    double *u = nullptr, *v = nullptr, *x = nullptr, *w = nullptr, *se = nullptr;
    double *bst = nullptr;  // b staging (m_k)
    double *P = nullptr;    // nsl * n slice partials
    double *part = nullptr;
    int part_len = 0;
    Scal *sc = nullptr;
    double *trace_d = nullptr;
    int trace_rows = 0;
    double *hx = nullptr, *hy = nullptr;  // staging for host-pointer aprod (n, m)
    // column renumbering (bricks of voxels); every column-space vector on the
    // device is in the new order, the API boundary converts
    bool permuted = false;
    int *perm = nullptr, *inv = nullptr;  // old -> new, new -> old
    double *hx2 = nullptr;                // n, new-order staging
    // tiled mode-2 layout (tile_build.h)
    bool tiled = false;
    int tS = 1, tT = 0, tmax_seg = 0;
    int *t_start = nullptr, *t_ray = nullptr;
    long long *t_slot = nullptr, *t_seg = nullptr, *t_cut = nullptr;
    uint16_t *t_vlen = nullptr, *t_sseg = nullptr, *t_vcut = nullptr;
    double *t_sval = nullptr;
    long long t_nslot = 0, t_nseg = 0;
    // slot form of the tiled mode-2 layout (SlotRows): 0 by voxel, 1 by
    // position in packed rows, 2 by position in TW-wide rows
    int t_form = 0;
    uint8_t *t_pvox = nullptr;
    long long *t_rbase = nullptr;
    int *t_roff = nullptr;
    bool nt_loads = false;
    int tw = 128;  // voxels per tile (brick) of the tiled layouts: 128 or 256
    bool rows_nt = false;  // streaming loads for the mode-1 matrix data
    // mode 1 on the bricked layout (segment partials + per-ray combine)
    bool rtiled = false;
    long long rt_nent = 0;
    int rt_lc = 4;
    int *rt_gbase = nullptr, *rt_rptr = nullptr, *rt_rseg = nullptr;
    // the mode-1 layout's own tiles (stacks of two bricks, or the mode-2 tiles)
    int rt_tw = 256, rtT = 0;
    int *rt_start = nullptr;
    long long *rt_seg = nullptr;
    uint8_t *rt_lo = nullptr;  // per-segment counts of entries below voxel 256 (512, 768)
    uint16_t *rt_dst = nullptr;  // per lane: the partial's ray-order index (length-sorted windows)
    // pass 1 over chunks of at most rt_chunk groups of one item (0: one
    // workgroup per item); chunk c = 8k + s is the k-th of row slice s
    int rt_chunk = 0, rt_nch = 0;
    int *rt_chitem = nullptr, *rt_chg0 = nullptr;
    long long *rt_ebase = nullptr;
    uint8_t *rt_len = nullptr, *rt_vox = nullptr;
    double *rt_val = nullptr;
    double *rt_segv = nullptr;  // segment partials: n_seg doubles, or n_seg floats on the FP32 path
    // FP32 copies for mixed-precision solves (allocated on first use)
    float *rv32 = nullptr, *t_sval32 = nullptr, *u32 = nullptr, *v32 = nullptr, *rt_val32 = nullptr;
    float *fx = nullptr, *fy = nullptr;  // caller-order FP32 staging of the host FP32 aprod
    plsqr::Scal *sc_plain = nullptr;     // ib = 1, alpha = -1, vscale = 1: plain y += A x, x += A^T y
    double *dx = nullptr;
    // column scaling of the mixed solve: d = ||A e_j||^-power (device column
    // order), ||A||_F, and the scales the FP32 combine applies during a solve
    double *cscale = nullptr, *colss = nullptr;  // scales, column sums of squares
    double cs_power = 0.0;                        // exponent the scales were made for
    double fro = 0.0;
    const double *cs_active = nullptr;
    // warm starts: x0 in the device column order and r0 = [b;0] - A x0
    double *x0i = nullptr, *r0 = nullptr;
    // ray order: kernel rows renumbered so that rays crossing the same tiles
    // are adjacent; every row-space vector on the device is in the new order
    bool rowperm = false;
    int *rperm = nullptr, *rinv = nullptr;  // new -> old, old -> new (m rows: kernel, then damping)
    double *hy2 = nullptr;                  // m, new-order staging
    double *bint = nullptr;                 // m_k, b in the new order
    double *xt = nullptr;                   // n: the mixed solve's checked iterate x + dx
    // pinned host staging owned by the handle (hipHostMalloc at create, freed
    // at destroy): every read of device state or partial sums goes through it
    plsqr::Scal *hscal = nullptr;           // copy of the device Scal
    double *hpart = nullptr;                // part_len doubles
    std::vector<void *> allocs;
    size_t bytes = 0;
    // launch shapes
    int cu = 256;
    int L1 = 16, L2 = 16, G1 = 0, G2 = 0, G4 = 0, G6 = 0, Gs = 0;
    std::string err;
};

namespace {

int set_err(plsqr_op *op, int code, const std::string &msg) {
    if (op && op->err.empty()) op->err = msg;
    return code;
}

#define HIP_TRY(op, expr)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_err(op, PLSQR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK(op) HIP_TRY(op, hipGetLastError())

template <class T>
int dalloc(plsqr_op *op, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(op, PLSQR_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
                                                " bytes failed: " + hipGetErrorString(e));
    }
    op->allocs.push_back((void *)*p);
    op->bytes += sizeof(T) * count;
    return PLSQR_OK;
}

template <class T>
void dfree(plsqr_op *op, T *&p, size_t count) {
    if (!p) return;
    auto it = std::find(op->allocs.begin(), op->allocs.end(), (void *)p);
    if (it != op->allocs.end()) {
        op->allocs.erase(it);
        op->bytes -= sizeof(T) * std::max<size_t>(count, 1);
        (void)hipFree(p);
    }
    p = nullptr;
}

int grid_for(long long work, int per_block, int cap) {
    long long g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// damping columns as the combine kernels read them: interleaved when built,
// else the CSC (dlen NULL)
inline const int *dcol_start(const plsqr_op *op) { return op->dlen ? op->dpos : op->dcp; }
inline const int *dcol_row(const plsqr_op *op) { return op->dlen ? op->dri_il : op->dri; }
inline const double *dcol_val(const plsqr_op *op) { return op->dlen ? op->dcv_il : op->dcv; }

// Pinned host staging of the handle, allocated once at create.
int alloc_host_staging(plsqr_op *op) {
    hipError_t e = hipHostMalloc((void **)&op->hscal, sizeof(plsqr::Scal));
    if (e == hipSuccess) e = hipHostMalloc((void **)&op->hpart, sizeof(double) * std::max(op->part_len, 1));
    if (e != hipSuccess)
        return set_err(op, PLSQR_ERR_NOMEM, std::string("hipHostMalloc of the staging buffers: ") + hipGetErrorString(e));
    return PLSQR_OK;
}

void free_host_staging(plsqr_op *op) {
    if (op->hscal) (void)hipHostFree(op->hscal);
    if (op->hpart) (void)hipHostFree(op->hpart);
    op->hscal = nullptr;
    op->hpart = nullptr;
}

// The device Scal copied to the host through the pinned staging (the stream
// is synchronised before returning).
int read_scal(plsqr_op *op, plsqr::Scal *h) {
    HIP_TRY(op, hipMemcpyAsync(op->hscal, op->sc, sizeof(plsqr::Scal), hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    std::memcpy(h, op->hscal, sizeof(plsqr::Scal));
    return PLSQR_OK;
}

// Sum, in a fixed order on the host, of count <= part_len device partials.
int host_sum(plsqr_op *op, const double *part, int count, double *out) {
    if (count < 0 || count > op->part_len) return set_err(op, PLSQR_ERR_STATE, "partial count beyond the staging buffer");
    HIP_TRY(op, hipMemcpyAsync(op->hpart, part, sizeof(double) * count, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    double s = 0.0;
    for (int i = 0; i < count; ++i) s += op->hpart[i];
    *out = s;
    return PLSQR_OK;
}

// Owns what one solve creates (timing events, the captured iteration graph
// and its executable) and releases it on every exit path, after the work
// already on the stream has drained; also clears the column scales the FP32
// combine reads during a scaled solve.
struct SolveScope {
    plsqr_op *op;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    explicit SolveScope(plsqr_op *o) : op(o) {}
    SolveScope(const SolveScope &) = delete;
    SolveScope &operator=(const SolveScope &) = delete;
    int create_events() {
        HIP_TRY(op, hipEventCreate(&e0));
        HIP_TRY(op, hipEventCreate(&e1));
        return PLSQR_OK;
    }
    // capture one iteration enqueued by f() and instantiate it
    template <class F>
    int capture(F &&f) {
        HIP_TRY(op, hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
        const int rc = f();
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(op->stream, &g);
        graph = g;
        if (rc) return rc;
        if (ce != hipSuccess) return set_err(op, PLSQR_ERR_HIP, std::string("capture: ") + hipGetErrorString(ce));
        HIP_TRY(op, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return PLSQR_OK;
    }
    ~SolveScope() {
        if (op->stream) (void)hipStreamSynchronize(op->stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        op->cs_active = nullptr;
    }
};

// -------------------------------------------------------------------------
// kernel dispatch over the lanes-per-row template parameter
// -------------------------------------------------------------------------
constexpr int BS = 256;

// Mode-1 pass 1 on the bricked layout (FP64 or the FP32 copy).  A wavefront
// stages its group's entries in LDS rounds of CH; a round costs one memory
// wait, so CH is as large as the loads per lane allow without losing
// occupancy: 512-voxel tiles (64 segments hold 474 entries on average at
// 67^3 / 2M rays) stage 510 FP64 entries (4 loads per lane, 22.6 KB of LDS,
// 7 workgroups per CU) or 764 FP32 entries (3 loads, 19.5 KB, 8 per CU);
// 1024-voxel tiles 382 FP64 / 764 FP32; 256- and 128-voxel tiles 448 / 512.
template <class T>
void launch_pass1(plsqr_op *op, const T *val, const T *x, const Scal *sc, T *segv) {
    using namespace plsqr;
    const dim3 g(op->rt_chunk ? op->rt_nch : op->rtT * op->tS), b(BS);
#define PASS1(CH64, CH32, TW)                                                                                       \
    if constexpr (std::is_same<T, double>::value)                                                                     \
        hipLaunchKernelGGL((k_rows_tiled_c<BS, CH64, TW>), g, b, 0, op->stream, op->tS, op->rt_start, op->rt_seg,     \
                           op->rt_gbase, op->rt_chitem, op->rt_chg0, op->rt_chunk, op->rt_ebase, op->rt_len,         \
                           op->rt_lo, op->rt_dst, op->rt_vox, val, x, sc, segv);                                      \
    else                                                                                                              \
        hipLaunchKernelGGL((k_rows_tiled_c32<BS, CH32, TW>), g, b, 0, op->stream, op->tS, op->rt_start, op->rt_seg,   \
                           op->rt_gbase, op->rt_chitem, op->rt_chg0, op->rt_chunk, op->rt_ebase, op->rt_len,         \
                           op->rt_lo, op->rt_dst, op->rt_vox, val, x, sc, segv);
    if (op->rt_tw == 1024) {
        PASS1(382, 764, 1024)
    } else if (op->rt_tw == 512) {
        PASS1(510, 764, 512)
    } else if (op->rt_tw == 256) {
        PASS1(448, 448, 256)
    } else {
        PASS1(512, 512, 128)
    }
#undef PASS1
}

int launch_rows(plsqr_op *op, const double *x, double *y, const Scal *sc, double *part) {
    using namespace plsqr;
    dim3 g(op->G1), b(BS);
    if (op->rtiled) {
        launch_pass1<double>(op, op->rt_val, x, sc, op->rt_segv);
        LAUNCH_CHECK(op);
        switch (op->rt_lc) {
            case 1: hipLaunchKernelGGL((k_rows_tiled_combine<1, BS>), g, b, 0, op->stream, op->m_k, op->m, op->tS, op->bounds, op->rt_rptr, op->rt_rseg, op->rt_segv, op->rp, op->ci, op->rv, x, y, sc, part); break;
            case 2: hipLaunchKernelGGL((k_rows_tiled_combine<2, BS>), g, b, 0, op->stream, op->m_k, op->m, op->tS, op->bounds, op->rt_rptr, op->rt_rseg, op->rt_segv, op->rp, op->ci, op->rv, x, y, sc, part); break;
            case 8: hipLaunchKernelGGL((k_rows_tiled_combine<8, BS>), g, b, 0, op->stream, op->m_k, op->m, op->tS, op->bounds, op->rt_rptr, op->rt_rseg, op->rt_segv, op->rp, op->ci, op->rv, x, y, sc, part); break;
            default: hipLaunchKernelGGL((k_rows_tiled_combine<4, BS>), g, b, 0, op->stream, op->m_k, op->m, op->tS, op->bounds, op->rt_rptr, op->rt_rseg, op->rt_segv, op->rp, op->ci, op->rv, x, y, sc, part); break;
        }
        LAUNCH_CHECK(op);
        return PLSQR_OK;
    }
    if (op->rows_nt && (op->L1 == 8 || op->L1 == 16 || op->L1 == 32)) {
        if (op->L1 == 8)
            hipLaunchKernelGGL((k_rows<8, BS, true>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part);
        else if (op->L1 == 16)
            hipLaunchKernelGGL((k_rows<16, BS, true>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part);
        else
            hipLaunchKernelGGL((k_rows<32, BS, true>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part);
        LAUNCH_CHECK(op);
        return PLSQR_OK;
    }
    switch (op->L1) {
        case 2: hipLaunchKernelGGL((k_rows<2, BS>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part); break;
        case 4: hipLaunchKernelGGL((k_rows<4, BS>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part); break;
        case 8: hipLaunchKernelGGL((k_rows<8, BS>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part); break;
        case 16: hipLaunchKernelGGL((k_rows<16, BS>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part); break;
        case 32: hipLaunchKernelGGL((k_rows<32, BS>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part); break;
        case 64: hipLaunchKernelGGL((k_rows<64, BS>), g, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv, x, y, sc, part); break;
        default: return set_err(op, PLSQR_ERR_ARG, "unsupported lanes per row");
    }
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

constexpr int TILE_W = 128;

// r = [b;0] - A x in the device order (x: n, b: m_k kernel rows, r: m)
int launch_residual(plsqr_op *op, const double *x, const double *b, double *r) {
    using namespace plsqr;
    if (op->rtiled) {
        launch_pass1<double>(op, op->rt_val, x, (const Scal *)nullptr, op->rt_segv);
        hipLaunchKernelGGL((k_rows_tiled_residual<4, BS>), dim3(op->G1), dim3(BS), 0, op->stream, op->m_k, op->m,
                           op->tS, op->bounds, op->rt_rptr, op->rt_rseg, op->rt_segv, op->rp, op->ci, op->rv, x, b, r);
    } else {
        hipLaunchKernelGGL(k_residual, dim3(grid_for(op->m, 256, op->cu * 4)), dim3(256), 0, op->stream, op->m_k,
                           op->m, op->rp, op->ci, op->rv, x, b, r);
    }
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}



// cs: column scales of an LSQR on A D (the FP64 inner iteration of a scaled
// solve), NULL for A itself
int launch_cols(plsqr_op *op, const double *u, double *v, const Scal *sc, double *part,
                const double *cs = nullptr) {
    using namespace plsqr;
    if (op->tiled) {
        const SlotRows sr{op->t_pvox, op->t_rbase, op->t_roff};
#define COLS(TW, B, CAP, NT, FORM)                                                                                    \
    hipLaunchKernelGGL((k_cols_tiled<TW, B, CAP, NT, FORM>), dim3(op->tT * op->tS), dim3(B), 0, op->stream, op->n,     \
                       op->tS, op->t_start, op->t_slot, op->t_seg, op->t_vlen, op->t_ray, op->t_sval, op->t_sseg,     \
                       op->t_cut, op->t_vcut, sr, u, sc, op->P);
        // 4 threads per voxel; LDS for 8192 (4096) gathered segments, so few
        // items need a second round
        if (op->tw == 256) {
            if (op->t_form == 1) {
                COLS(256, 1024, 8192, false, 1)
            } else if (op->t_form == 2) {
                COLS(256, 1024, 8192, false, 2)
            } else {
                COLS(256, 1024, 8192, false, 0)
            }
        } else if (op->t_form == 1) {
            COLS(TILE_W, 512, 4096, false, 1)
        } else if (op->t_form == 2) {
            COLS(TILE_W, 512, 4096, false, 2)
        } else if (op->nt_loads) {
            COLS(TILE_W, 512, 4096, true, 0)
        } else {
            COLS(TILE_W, 512, 4096, false, 0)
        }
#undef COLS
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL((k_cols_combine<BS>), dim3(op->G4), dim3(BS), 0, op->stream, op->n, op->m_k, op->tS,
                           op->P, dcol_start(op), dcol_row(op), dcol_val(op), u, sc, v, part, cs,
                           (const uint8_t *)op->dlen);
        LAUNCH_CHECK(op);
        return PLSQR_OK;
    }
    dim3 g(op->G2), b(BS);
    switch (op->L2) {
        case 2: hipLaunchKernelGGL((k_cols_sliced<2, BS>), g, b, 0, op->stream, op->n, op->nsl, op->scp, op->sri, op->sv, u, sc, op->P); break;
        case 4: hipLaunchKernelGGL((k_cols_sliced<4, BS>), g, b, 0, op->stream, op->n, op->nsl, op->scp, op->sri, op->sv, u, sc, op->P); break;
        case 8: hipLaunchKernelGGL((k_cols_sliced<8, BS>), g, b, 0, op->stream, op->n, op->nsl, op->scp, op->sri, op->sv, u, sc, op->P); break;
        case 16: hipLaunchKernelGGL((k_cols_sliced<16, BS>), g, b, 0, op->stream, op->n, op->nsl, op->scp, op->sri, op->sv, u, sc, op->P); break;
        case 32: hipLaunchKernelGGL((k_cols_sliced<32, BS>), g, b, 0, op->stream, op->n, op->nsl, op->scp, op->sri, op->sv, u, sc, op->P); break;
        case 64: hipLaunchKernelGGL((k_cols_sliced<64, BS>), g, b, 0, op->stream, op->n, op->nsl, op->scp, op->sri, op->sv, u, sc, op->P); break;
        default: return set_err(op, PLSQR_ERR_ARG, "unsupported lanes per column");
    }
    LAUNCH_CHECK(op);
    hipLaunchKernelGGL((k_cols_combine<BS>), dim3(op->G4), dim3(BS), 0, op->stream, op->n, op->m_k, op->nsl,
                       op->P, op->dcp, op->dri, op->dcv, u, sc, v, part, cs);
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

// y += A x  (mode 1) on device pointers
int aprod1_dev(plsqr_op *op, const double *x, double *y) {
    if (op->exact) {
        hipLaunchKernelGGL(plsqr::k_rows_exact, dim3(grid_for(op->m, BS, op->cu * 8)), dim3(BS), 0, op->stream,
                           op->m, op->rp, op->ci, op->rv, x, y, (const Scal *)nullptr);
        LAUNCH_CHECK(op);
        return PLSQR_OK;
    }
    return launch_rows(op, x, y, nullptr, nullptr);
}

// x += A^T y  (mode 2) on device pointers
int aprod2_dev(plsqr_op *op, double *x, const double *y) {
    if (op->exact) {
        hipLaunchKernelGGL(plsqr::k_cols_exact, dim3(grid_for(op->n, BS, op->cu * 8)), dim3(BS), 0, op->stream,
                           op->n, op->m_k, op->kcp, op->kri, op->kv, op->dcp, op->dri, op->dcv, y,
                           (const Scal *)nullptr, x);
        LAUNCH_CHECK(op);
        return PLSQR_OK;
    }
    return launch_cols(op, y, x, nullptr, nullptr);
}

// Column renumbering at the API boundary: callers see the original column
// order, the device works in the renumbered one.
int to_internal(plsqr_op *op, const double *ext, double *in) {
    hipLaunchKernelGGL(plsqr::k_gather, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, op->n,
                       op->inv, ext, in);
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

int to_external(plsqr_op *op, const double *in, double *ext) {
    hipLaunchKernelGGL(plsqr::k_gather, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, op->n,
                       op->perm, in, ext);
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

// row-space vectors (m = kernel rows then damping rows) between the caller's
// order and the device order (kernel rows among themselves, damping rows
// among themselves)
int rows_to_internal(plsqr_op *op, const double *ext, double *in) {
    if (!op->rowperm) {
        if (in != ext) HIP_TRY(op, hipMemcpyAsync(in, ext, sizeof(double) * op->m, hipMemcpyDeviceToDevice, op->stream));
        return PLSQR_OK;
    }
    hipLaunchKernelGGL(plsqr::k_gather, dim3(grid_for(op->m, 256, op->cu * 4)), dim3(256), 0, op->stream, op->m,
                       op->rperm, ext, in);
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

int rows_to_external(plsqr_op *op, const double *in, double *ext) {
    if (!op->rowperm) {
        if (in != ext) HIP_TRY(op, hipMemcpyAsync(ext, in, sizeof(double) * op->m, hipMemcpyDeviceToDevice, op->stream));
        return PLSQR_OK;
    }
    hipLaunchKernelGGL(plsqr::k_gather, dim3(grid_for(op->m, 256, op->cu * 4)), dim3(256), 0, op->stream, op->m,
                       op->rinv, in, ext);
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

// b (m_k, caller order, device) in the device row order: op->bint or b itself
int b_internal(plsqr_op *op, const double *b, const double **out) {
    if (!op->rowperm) {
        *out = b;
        return PLSQR_OK;
    }
    hipLaunchKernelGGL(plsqr::k_gather, dim3(grid_for(op->m_k, 256, op->cu * 4)), dim3(256), 0, op->stream, op->m_k,
                       op->rperm, b, op->bint);
    LAUNCH_CHECK(op);
    *out = op->bint;
    return PLSQR_OK;
}

// aprod on operands each either in the caller's order (renumbered in and,
// for the output, back) or already in the device order
int aprod_any(plsqr_op *op, int mode, double *x, bool x_ext, double *y, bool y_ext) {
    int rc;
    double *xi = x, *yi = y;
    if (x_ext && op->permuted) {
        if ((rc = to_internal(op, x, op->hx2))) return rc;
        xi = op->hx2;
    }
    if (y_ext && op->rowperm) {
        if ((rc = rows_to_internal(op, y, op->hy2))) return rc;
        yi = op->hy2;
    }
    if ((rc = mode == 1 ? aprod1_dev(op, xi, yi) : aprod2_dev(op, xi, yi))) return rc;
    if (mode == 2 && xi != x) return to_external(op, xi, x);
    if (mode == 1 && yi != y) return rows_to_external(op, yi, y);
    return PLSQR_OK;
}

// aprod on caller-order x and y (device)
int aprod_ext_dev(plsqr_op *op, int mode, double *x, double *y) { return aprod_any(op, mode, x, true, y, true); }

// -------------------------------------------------------------------------
// host-side validation of the loaded arrays (the reference loaders' checks,
// plsqr3.c:1149-1153, 1178-1182, 1606-1615, as error codes)
// -------------------------------------------------------------------------
int check_ptr(plsqr_op *op, const char *what, const int *p, int len, long long nnz) {
    if (!p) return set_err(op, PLSQR_ERR_ARG, std::string(what) + " is NULL");
    if (p[0] != 0) return set_err(op, PLSQR_ERR_FORMAT, std::string(what) + "[0] != 0");
    for (int i = 0; i < len; ++i)
        if (p[i + 1] < p[i]) return set_err(op, PLSQR_ERR_FORMAT, std::string(what) + " decreases at " + std::to_string(i));
    if (p[len] != nnz) return set_err(op, PLSQR_ERR_FORMAT, std::string(what) + " does not end at nnz");
    return PLSQR_OK;
}

int check_idx(plsqr_op *op, const char *what, const int *idx, const double *val, long long nnz, int bound) {
    if (nnz > 0 && (!idx || !val)) return set_err(op, PLSQR_ERR_ARG, std::string(what) + " arrays are NULL");
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nnz < (1 << 20)) nt = 1;
    std::vector<long long> bad_idx(nt, -1), bad_val(nt, -1);
    std::vector<std::thread> th;
    const long long chunk = (nnz + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            const long long b = t * chunk, e = std::min(nnz, b + chunk);
            for (long long i = b; i < e; ++i) {
                if (idx[i] < 0 || idx[i] >= bound) {
                    bad_idx[t] = i;
                    return;
                }
                if (!std::isfinite(val[i])) {
                    bad_val[t] = i;
                    return;
                }
            }
        });
    }
    for (auto &x : th) x.join();
    for (unsigned t = 0; t < nt; ++t) {
        if (bad_idx[t] >= 0)
            return set_err(op, PLSQR_ERR_BOUNDS, std::string(what) + " index out of range at entry " + std::to_string(bad_idx[t]));
        if (bad_val[t] >= 0)
            return set_err(op, PLSQR_ERR_NONFINITE, std::string(what) + " value not finite at entry " + std::to_string(bad_val[t]));
    }
    return PLSQR_OK;
}

int bits_for(long long v) {
    int b = 1;
    while ((1LL << b) <= v) ++b;
    return b;
}

// stable sort of entry indices by key (LSD radix sort: a sequence of stable
// counting sorts), returns the permutation in *perm_out (device, len ints)
int stable_order(plsqr_op *op, const int *keys, long long len, int bits, int **perm_out) {
    int *kin = nullptr, *kout = nullptr, *vin = nullptr, *vout = nullptr;
    int rc;
    if ((rc = dalloc(op, &kin, len)) || (rc = dalloc(op, &kout, len)) || (rc = dalloc(op, &vin, len)) ||
        (rc = dalloc(op, &vout, len)))
        return rc;
    HIP_TRY(op, hipMemcpyAsync(kin, keys, sizeof(int) * len, hipMemcpyDeviceToDevice, op->stream));
    hipLaunchKernelGGL(plsqr::k_iota, dim3(grid_for(len, 256, 4096)), dim3(256), 0, op->stream, len, vin);
    LAUNCH_CHECK(op);
    size_t tmp_bytes = 0;
    HIP_TRY(op, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, vin, vout, (int)len, 0, bits,
                                                   op->stream));
    void *tmp = nullptr;
    if ((rc = dalloc(op, (char **)&tmp, tmp_bytes))) return rc;
    HIP_TRY(op, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int)len, 0, bits,
                                                   op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    char *tmpc = (char *)tmp;
    dfree(op, tmpc, tmp_bytes);
    dfree(op, kin, len);
    dfree(op, kout, len);
    dfree(op, vin, len);
    *perm_out = vout;
    return PLSQR_OK;
}

// exclusive prefix sum of counts (len+1 entries, last count 0) into out
int exclusive_scan(plsqr_op *op, const int *cnt, int *out, long long len) {
    size_t tmp_bytes = 0;
    HIP_TRY(op, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, out, (int)len, op->stream));
    char *tmp = nullptr;
    int rc;
    if ((rc = dalloc(op, &tmp, tmp_bytes))) return rc;
    HIP_TRY(op, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, out, (int)len, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    dfree(op, tmp, tmp_bytes);
    return PLSQR_OK;
}

// L: the mode-2 layout; R: the mode-1 layout, built on the tiles of L1
// (L itself, or stacks of two of its bricks).
int build(plsqr_op *op, const plsqr_system *s, const std::vector<int> &hb, const plsqr::TileLayout *L,
          const plsqr::TileLayout *L1, const plsqr::RowTiles *R) {
    using namespace plsqr;
    int rc;
    const long long nk = s->nnz_k, nd = s->nnz_d;
    const int gcap = op->cu * 16;
    // ---- kernel CSC upload
    if ((rc = dalloc(op, &op->kcp, (size_t)op->n + 1)) || (rc = dalloc(op, &op->kri, nk)) ||
        (rc = dalloc(op, &op->kv, nk)))
        return rc;
    HIP_TRY(op, hipMemcpyAsync(op->kcp, s->k_col_ptr, sizeof(int) * (op->n + 1), hipMemcpyHostToDevice, op->stream));
    HIP_TRY(op, hipMemcpyAsync(op->kri, s->k_row_idx, sizeof(int) * nk, hipMemcpyHostToDevice, op->stream));
    HIP_TRY(op, hipMemcpyAsync(op->kv, s->k_val, sizeof(double) * nk, hipMemcpyHostToDevice, op->stream));
    // ---- column index of every CSC entry
    int *col = nullptr;
    if ((rc = dalloc(op, &col, nk))) return rc;
    hipLaunchKernelGGL(k_expand_ptr, dim3(grid_for(op->n, 256, gcap)), dim3(256), 0, op->stream, op->n, op->kcp, col);
    LAUNCH_CHECK(op);
    // ---- CSR of A_k: stable sort of the column-major entries by row
    int *perm = nullptr;
    if ((rc = stable_order(op, op->kri, nk, bits_for(op->m_k), &perm))) return rc;
    const long long nr = nk + nd;
    if ((rc = dalloc(op, &op->rp, (size_t)op->m + 1)) || (rc = dalloc(op, &op->ci, nr)) || (rc = dalloc(op, &op->rv, nr)))
        return rc;
    hipLaunchKernelGGL(k_permute, dim3(grid_for(nk, 256, gcap)), dim3(256), 0, op->stream, nk, perm, col, op->kv,
                       op->ci, op->rv);
    LAUNCH_CHECK(op);
    dfree(op, perm, nk);
    {
        int *cnt = nullptr;
        if ((rc = dalloc(op, &cnt, (size_t)op->m_k + 1))) return rc;
        HIP_TRY(op, hipMemsetAsync(cnt, 0, sizeof(int) * (op->m_k + 1), op->stream));
        hipLaunchKernelGGL(k_count, dim3(grid_for(nk, 256, gcap)), dim3(256), 0, op->stream, nk, op->kri, cnt);
        LAUNCH_CHECK(op);
        if ((rc = exclusive_scan(op, cnt, op->rp, (long long)op->m_k + 1))) return rc;
        dfree(op, cnt, (size_t)op->m_k + 1);
    }
    // ---- damping rows appended
    {
        int *drp = nullptr, *dci = nullptr;
        double *dav = nullptr;
        if ((rc = dalloc(op, &drp, (size_t)op->m_d + 1)) || (rc = dalloc(op, &dci, nd)) || (rc = dalloc(op, &dav, nd)))
            return rc;
        HIP_TRY(op, hipMemcpyAsync(drp, s->d_row_ptr, sizeof(int) * (op->m_d + 1), hipMemcpyHostToDevice, op->stream));
        if (nd > 0) {
            HIP_TRY(op, hipMemcpyAsync(dci, s->d_col_idx, sizeof(int) * nd, hipMemcpyHostToDevice, op->stream));
            HIP_TRY(op, hipMemcpyAsync(dav, s->d_val, sizeof(double) * nd, hipMemcpyHostToDevice, op->stream));
        }
        hipLaunchKernelGGL(k_append_rows, dim3(grid_for(std::max<long long>(nd, op->m_d + 1), 256, gcap)), dim3(256),
                           0, op->stream, op->m_k, op->m_d, nk, nd, drp, dci, dav, op->rp, op->ci, op->rv);
        LAUNCH_CHECK(op);
        HIP_TRY(op, hipStreamSynchronize(op->stream));
        dfree(op, drp, (size_t)op->m_d + 1);
        dfree(op, dci, nd);
        dfree(op, dav, nd);
    }
    // ---- column renumbering: CSR column indices to the new order
    if (L) {
        op->permuted = true;
        if ((rc = dalloc(op, &op->perm, op->n)) || (rc = dalloc(op, &op->inv, op->n)) || (rc = dalloc(op, &op->hx2, op->n)))
            return rc;
        HIP_TRY(op, hipMemcpy(op->perm, L->perm.data(), sizeof(int) * op->n, hipMemcpyHostToDevice));
        HIP_TRY(op, hipMemcpy(op->inv, L->inv.data(), sizeof(int) * op->n, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_remap, dim3(grid_for(nr, 256, gcap)), dim3(256), 0, op->stream, nr, op->perm, op->ci);
        LAUNCH_CHECK(op);
    }
    // ---- damping CSC
    if ((rc = dalloc(op, &op->dcp, (size_t)op->n + 1)) || (rc = dalloc(op, &op->dri, nd)) || (rc = dalloc(op, &op->dcv, nd)))
        return rc;
    if (L) {
        std::vector<int> cp((size_t)op->n + 1, 0), ri((size_t)std::max<long long>(nd, 1));
        std::vector<double> cv((size_t)std::max<long long>(nd, 1));
        for (int j = 0; j < op->n; ++j) {
            const int o = L->inv[(size_t)j];
            const int b = s->dc_col_ptr[o], e = s->dc_col_ptr[o + 1];
            cp[(size_t)j + 1] = cp[(size_t)j] + (e - b);
            std::copy(s->dc_row_idx + b, s->dc_row_idx + e, ri.begin() + cp[(size_t)j]);
            std::copy(s->dc_val + b, s->dc_val + e, cv.begin() + cp[(size_t)j]);
        }
        HIP_TRY(op, hipMemcpy(op->dcp, cp.data(), sizeof(int) * (op->n + 1), hipMemcpyHostToDevice));
        if (nd > 0) {
            HIP_TRY(op, hipMemcpy(op->dri, ri.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
            HIP_TRY(op, hipMemcpy(op->dcv, cv.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
        }
        // interleaved copy for the combine kernels: the columns in groups of 64,
        // each group as rows of 64 slots (slot k of column c at dpos[c] + 64 k)
        // as deep as its longest column; PLSQR_DAMP_IL=0 keeps the CSC
        int maxlen = 0;
        for (int j = 0; j < op->n; ++j) maxlen = std::max(maxlen, cp[(size_t)j + 1] - cp[(size_t)j]);
        if (nd > 0 && maxlen <= 255 && env_int("PLSQR_DAMP_IL", 1) != 0) {
            const int ng = (op->n + 63) / 64;
            std::vector<int> pos((size_t)op->n);
            std::vector<uint8_t> len((size_t)op->n);
            long long slots = 0;
            for (int g = 0; g < ng; ++g) {
                int w = 0;
                for (int j = g * 64; j < std::min(op->n, g * 64 + 64); ++j) {
                    len[(size_t)j] = (uint8_t)(cp[(size_t)j + 1] - cp[(size_t)j]);
                    w = std::max(w, (int)len[(size_t)j]);
                    pos[(size_t)j] = (int)(slots + (j - g * 64));
                }
                slots += (long long)w * 64;
            }
            if (slots < INT_MAX) {
                std::vector<int> ril((size_t)slots, 0);
                std::vector<double> vil((size_t)slots, 0.0);
                for (int j = 0; j < op->n; ++j)
                    for (int k = 0; k < len[(size_t)j]; ++k) {
                        ril[(size_t)pos[(size_t)j] + (size_t)k * 64] = ri[(size_t)cp[(size_t)j] + k];
                        vil[(size_t)pos[(size_t)j] + (size_t)k * 64] = cv[(size_t)cp[(size_t)j] + k];
                    }
                if ((rc = dalloc(op, &op->dpos, op->n)) || (rc = dalloc(op, &op->dlen, op->n)) ||
                    (rc = dalloc(op, &op->dri_il, slots)) || (rc = dalloc(op, &op->dcv_il, slots)))
                    return rc;
                op->d_il = slots;
                HIP_TRY(op, hipMemcpy(op->dpos, pos.data(), sizeof(int) * op->n, hipMemcpyHostToDevice));
                HIP_TRY(op, hipMemcpy(op->dlen, len.data(), op->n, hipMemcpyHostToDevice));
                HIP_TRY(op, hipMemcpy(op->dri_il, ril.data(), sizeof(int) * slots, hipMemcpyHostToDevice));
                HIP_TRY(op, hipMemcpy(op->dcv_il, vil.data(), sizeof(double) * slots, hipMemcpyHostToDevice));
            }
        }
    } else {
        HIP_TRY(op, hipMemcpyAsync(op->dcp, s->dc_col_ptr, sizeof(int) * (op->n + 1), hipMemcpyHostToDevice, op->stream));
        if (nd > 0) {
            HIP_TRY(op, hipMemcpyAsync(op->dri, s->dc_row_idx, sizeof(int) * nd, hipMemcpyHostToDevice, op->stream));
            HIP_TRY(op, hipMemcpyAsync(op->dcv, s->dc_val, sizeof(double) * nd, hipMemcpyHostToDevice, op->stream));
        }
    }
    // ---- tiled mode-2 layout
    if (L) {
        if (L->TW != 128 && L->TW != 256) return set_err(op, PLSQR_ERR_STATE, "tile width must be 128 or 256");
        op->tiled = true;
        op->tw = L->TW;
        op->tS = L->S;
        op->tT = L->T;
        op->nsl = L->S;
        op->tmax_seg = L->max_seg_per_item;
        // the slot buffers end in TW zero slots: the packed-row loop of mode 2
        // lets a lane past its length load (and drop) a slot up to TW - 1
        // past its row's start, and the last row may end the array
        op->t_nslot = L->n_slot + L->TW;
        op->t_nseg = L->n_seg;
        const long long items = (long long)L->T * L->S;
        if ((rc = dalloc(op, &op->t_start, (size_t)L->T + 1)) || (rc = dalloc(op, &op->t_slot, (size_t)items + 1)) ||
            (rc = dalloc(op, &op->t_seg, (size_t)items + 1)) || (rc = dalloc(op, &op->t_vlen, (size_t)items * L->TW)) ||
            (rc = dalloc(op, &op->t_ray, (size_t)L->n_seg)) || (rc = dalloc(op, &op->t_sval, (size_t)op->t_nslot)) ||
            (rc = dalloc(op, &op->t_sseg, (size_t)op->t_nslot)) || (rc = dalloc(op, &op->t_cut, (size_t)items + 1)) ||
            (rc = dalloc(op, &op->t_vcut, L->vox_cut.size())))
            return rc;
        HIP_TRY(op, hipMemset(op->t_sval + L->n_slot, 0, sizeof(double) * L->TW));
        HIP_TRY(op, hipMemset(op->t_sseg + L->n_slot, 0, sizeof(uint16_t) * L->TW));
        HIP_TRY(op, hipMemcpy(op->t_cut, L->cut_base.data(), sizeof(long long) * (items + 1), hipMemcpyHostToDevice));
        if (!L->vox_cut.empty())
            HIP_TRY(op, hipMemcpy(op->t_vcut, L->vox_cut.data(), sizeof(uint16_t) * L->vox_cut.size(),
                                  hipMemcpyHostToDevice));
        HIP_TRY(op, hipMemcpy(op->t_start, L->tile_start.data(), sizeof(int) * (L->T + 1), hipMemcpyHostToDevice));
        HIP_TRY(op, hipMemcpy(op->t_slot, L->slot_base.data(), sizeof(long long) * (items + 1), hipMemcpyHostToDevice));
        HIP_TRY(op, hipMemcpy(op->t_seg, L->seg_start.data(), sizeof(long long) * (items + 1), hipMemcpyHostToDevice));
        HIP_TRY(op, hipMemcpy(op->t_vlen, L->vox_len.data(), sizeof(uint16_t) * items * L->TW, hipMemcpyHostToDevice));
        if (L->n_seg > 0)
            HIP_TRY(op, hipMemcpy(op->t_ray, L->seg_ray.data(), sizeof(int) * L->n_seg, hipMemcpyHostToDevice));
        if (L->n_slot > 0) {
            HIP_TRY(op, hipMemcpy(op->t_sval, L->sval.data(), sizeof(double) * L->n_slot, hipMemcpyHostToDevice));
            HIP_TRY(op, hipMemcpy(op->t_sseg, L->sseg.data(), sizeof(uint16_t) * L->n_slot, hipMemcpyHostToDevice));
        }
        if (L->compact) {
            op->t_form = L->packed ? 1 : 2;
            if ((rc = dalloc(op, &op->t_pvox, (size_t)items * L->TW))) return rc;
            HIP_TRY(op, hipMemcpy(op->t_pvox, L->pos_vox.data(), (size_t)items * L->TW, hipMemcpyHostToDevice));
        }
        if (L->packed) {
            const size_t nrow = L->row_off.size();
            if ((rc = dalloc(op, &op->t_rbase, (size_t)items + 1)) || (rc = dalloc(op, &op->t_roff, std::max<size_t>(nrow, 1))))
                return rc;
            HIP_TRY(op, hipMemcpy(op->t_rbase, L->row_base.data(), sizeof(long long) * (items + 1), hipMemcpyHostToDevice));
            if (nrow > 0) HIP_TRY(op, hipMemcpy(op->t_roff, L->row_off.data(), sizeof(int) * nrow, hipMemcpyHostToDevice));
        }
        if (R) {
            if (!L1 || L1->S != L->S) return set_err(op, PLSQR_ERR_STATE, "mode-1 layout without its tiles");
            op->rtiled = true;
            op->rt_nent = R->n_ent;
            op->rt_tw = L1->TW;
            op->rtT = L1->T;
            if ((rc = dalloc(op, &op->bounds, (size_t)L->S + 1))) return rc;
            HIP_TRY(op, hipMemcpy(op->bounds, L->slice_bounds.data(), sizeof(int) * (L->S + 1), hipMemcpyHostToDevice));
            const long long items2 = (long long)L1->T * L1->S;
            if ((rc = dalloc(op, &op->rt_start, (size_t)L1->T + 1)) || (rc = dalloc(op, &op->rt_seg, (size_t)items2 + 1)) ||
                (rc = dalloc(op, &op->rt_gbase, (size_t)items2 + 1)) ||
                (rc = dalloc(op, &op->rt_ebase, (size_t)R->n_grp + 1)) ||
                (rc = dalloc(op, &op->rt_len, (size_t)R->n_grp * plsqr::RowTiles::G)) ||
                // entries carry 4 of slack: the 16-byte staging loads may read past the last group
                (rc = dalloc(op, &op->rt_vox, (size_t)R->n_ent + 4)) || (rc = dalloc(op, &op->rt_val, (size_t)R->n_ent + 4)) ||
                (rc = dalloc(op, &op->rt_rptr, (size_t)op->m_k + 1)) || (rc = dalloc(op, &op->rt_rseg, (size_t)L1->n_seg)) ||
                (rc = dalloc(op, &op->rt_segv, (size_t)L1->n_seg)))
                return rc;
            HIP_TRY(op, hipMemcpy(op->rt_start, L1->tile_start.data(), sizeof(int) * (L1->T + 1), hipMemcpyHostToDevice));
            HIP_TRY(op, hipMemcpy(op->rt_seg, L1->seg_start.data(), sizeof(long long) * (items2 + 1), hipMemcpyHostToDevice));
            if (!R->seg_lo.empty()) {
                if ((rc = dalloc(op, &op->rt_lo, R->seg_lo.size()))) return rc;
                HIP_TRY(op, hipMemcpy(op->rt_lo, R->seg_lo.data(), R->seg_lo.size(), hipMemcpyHostToDevice));
            }
            if (!R->seg_dst.empty()) {
                if ((rc = dalloc(op, &op->rt_dst, R->seg_dst.size()))) return rc;
                HIP_TRY(op, hipMemcpy(op->rt_dst, R->seg_dst.data(), sizeof(uint16_t) * R->seg_dst.size(),
                                      hipMemcpyHostToDevice));
            }
            // chunks of at most rt_chunk groups (PLSQR_M1_CHUNK, 0: whole
            // items), dealt so that chunk 8k + s belongs to row slice s and
            // runs on the XCD of the slice's other work
            op->rt_chunk = std::max(0, env_int("PLSQR_M1_CHUNK", 16));
            if (op->rt_chunk > 0) {
                const int S = L1->S;
                std::vector<std::vector<int>> ci((size_t)S), cg((size_t)S);
                for (long long it = 0; it < items2; ++it)
                    for (int g0 = R->grp_base[(size_t)it]; g0 < R->grp_base[(size_t)it + 1]; g0 += op->rt_chunk) {
                        ci[(size_t)(it % S)].push_back((int)it);
                        cg[(size_t)(it % S)].push_back(g0);
                    }
                size_t kmax = 0;
                for (int q = 0; q < S; ++q) kmax = std::max(kmax, ci[(size_t)q].size());
                const size_t nch = std::max<size_t>(kmax * (size_t)S, 1);
                std::vector<int> hitem(nch, -1), hg0(nch, 0);
                for (int q = 0; q < S; ++q)
                    for (size_t k = 0; k < ci[(size_t)q].size(); ++k) {
                        hitem[k * S + q] = ci[(size_t)q][k];
                        hg0[k * S + q] = cg[(size_t)q][k];
                    }
                if (nch > (size_t)INT_MAX) return set_err(op, PLSQR_ERR_ARG, "too many mode-1 chunks");
                op->rt_nch = (int)nch;
                if ((rc = dalloc(op, &op->rt_chitem, nch)) || (rc = dalloc(op, &op->rt_chg0, nch))) return rc;
                HIP_TRY(op, hipMemcpy(op->rt_chitem, hitem.data(), sizeof(int) * nch, hipMemcpyHostToDevice));
                HIP_TRY(op, hipMemcpy(op->rt_chg0, hg0.data(), sizeof(int) * nch, hipMemcpyHostToDevice));
            }
            HIP_TRY(op, hipMemcpy(op->rt_gbase, R->grp_base.data(), sizeof(int) * (items2 + 1), hipMemcpyHostToDevice));
            HIP_TRY(op, hipMemcpy(op->rt_ebase, R->ent_base.data(), sizeof(long long) * (R->n_grp + 1), hipMemcpyHostToDevice));
            HIP_TRY(op, hipMemcpy(op->rt_rptr, R->ray_ptr.data(), sizeof(int) * (op->m_k + 1), hipMemcpyHostToDevice));
            if (R->n_grp > 0)
                HIP_TRY(op, hipMemcpy(op->rt_len, R->seg_len.data(), (size_t)R->n_grp * plsqr::RowTiles::G, hipMemcpyHostToDevice));
            if (R->n_ent > 0) {
                HIP_TRY(op, hipMemcpy(op->rt_vox, R->ent_vox.data(), (size_t)R->n_ent, hipMemcpyHostToDevice));
                HIP_TRY(op, hipMemcpy(op->rt_val, R->ent_val.data(), sizeof(double) * R->n_ent, hipMemcpyHostToDevice));
            }
            if (L1->n_seg > 0)
                HIP_TRY(op, hipMemcpy(op->rt_rseg, R->ray_seg.data(), sizeof(int) * L1->n_seg, hipMemcpyHostToDevice));
        }
        dfree(op, col, nk);
        dfree(op, op->kcp, (size_t)op->n + 1);
        dfree(op, op->kri, nk);
        dfree(op, op->kv, nk);
        HIP_TRY(op, hipStreamSynchronize(op->stream));
        return PLSQR_OK;
    }
    // ---- row slices of the kernel CSC, balanced by nonzeros
    const int nsl = (int)hb.size() - 1;
    op->nsl = nsl;
    if ((rc = dalloc(op, &op->bounds, (size_t)nsl + 1))) return rc;
    HIP_TRY(op, hipMemcpy(op->bounds, hb.data(), sizeof(int) * (nsl + 1), hipMemcpyHostToDevice));
    if (nsl == 1) {
        op->scp = op->kcp;
        op->sri = op->kri;
        op->sv = op->kv;
    } else {
        const long long nkey = (long long)nsl * op->n;
        int *key = nullptr, *cnt = nullptr, *off = nullptr;
        if ((rc = dalloc(op, &key, nk)) || (rc = dalloc(op, &cnt, nkey + 1)) || (rc = dalloc(op, &off, nkey + 1)))
            return rc;
        hipLaunchKernelGGL(k_slice_key, dim3(grid_for(nk, 256, gcap)), dim3(256), 0, op->stream, nk, op->n, nsl,
                           op->bounds, op->kri, col, key);
        LAUNCH_CHECK(op);
        HIP_TRY(op, hipMemsetAsync(cnt, 0, sizeof(int) * (nkey + 1), op->stream));
        hipLaunchKernelGGL(k_count, dim3(grid_for(nk, 256, gcap)), dim3(256), 0, op->stream, nk, key, cnt);
        LAUNCH_CHECK(op);
        if ((rc = exclusive_scan(op, cnt, off, nkey + 1))) return rc;
        if ((rc = dalloc(op, &op->scp, (size_t)nsl * (op->n + 1)))) return rc;
        hipLaunchKernelGGL(k_slice_ptr, dim3(grid_for((long long)nsl * (op->n + 1), 256, gcap)), dim3(256), 0,
                           op->stream, op->n, nsl, off, nk, op->scp);
        LAUNCH_CHECK(op);
        int *perm2 = nullptr;
        if ((rc = stable_order(op, key, nk, bits_for(nkey), &perm2))) return rc;
        if ((rc = dalloc(op, &op->sri, nk)) || (rc = dalloc(op, &op->sv, nk))) return rc;
        hipLaunchKernelGGL(k_permute, dim3(grid_for(nk, 256, gcap)), dim3(256), 0, op->stream, nk, perm2, op->kri,
                           op->kv, op->sri, op->sv);
        LAUNCH_CHECK(op);
        HIP_TRY(op, hipStreamSynchronize(op->stream));
        dfree(op, perm2, nk);
        dfree(op, key, nk);
        dfree(op, cnt, nkey + 1);
        dfree(op, off, nkey + 1);
        if (!op->exact) {  // the sliced copy serves mode 2; the loaded CSC is no longer needed
            dfree(op, op->kcp, (size_t)op->n + 1);
            dfree(op, op->kri, nk);
            dfree(op, op->kv, nk);
        }
    }
    dfree(op, col, nk);
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

// Row slice boundaries balanced by kernel nonzeros (per-row counts from the CSC).
std::vector<int> slice_bounds_host(const plsqr_system *s, int nsl) {
    std::vector<long long> cnt((size_t)s->m_k + 1, 0);
    for (long long e = 0; e < s->nnz_k; ++e) cnt[(size_t)s->k_row_idx[e] + 1]++;
    for (int r = 0; r < s->m_k; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
    std::vector<int> hb((size_t)nsl + 1, 0);
    for (int sl = 1; sl < nsl; ++sl) {
        const long long target = s->nnz_k * sl / nsl;
        hb[(size_t)sl] = (int)(std::lower_bound(cnt.begin(), cnt.end(), target) - cnt.begin());
        hb[(size_t)sl] = std::min(std::max(hb[(size_t)sl], hb[(size_t)sl - 1]), s->m_k);
    }
    hb[(size_t)nsl] = s->m_k;
    return hb;
}

int ensure_trace(plsqr_op *op, int rows) {
    if (rows <= op->trace_rows) return PLSQR_OK;
    dfree(op, op->trace_d, (size_t)op->trace_rows * 8);
    op->trace_rows = 0;
    int rc = dalloc(op, &op->trace_d, (size_t)rows * 8);
    if (rc) return rc;
    op->trace_rows = rows;
    return PLSQR_OK;
}

void default_opts(const plsqr_solve_opts *o, plsqr_solve_opts *d) {
    *d = *o;
    if (d->atol == 0.0 && d->btol == 0.0 && d->conlim == 0.0) {
        d->atol = 1e-10;  // lsqr.c:47-49
        d->btol = 1e-10;
        d->conlim = 100.0;
    }
    if (d->itnlim <= 0) d->itnlim = 1000;  // lsqr.c:50
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char *plsqr_last_error(const plsqr_op *op) { return op ? op->err.c_str() : "null handle"; }

void plsqr_clear_error(plsqr_op *op) {
    if (op) op->err.clear();
}

int plsqr_op_create(const plsqr_system *s, int flags, plsqr_op **out) {
    return plsqr_op_create_ex(s, flags, nullptr, out);
}

int plsqr_op_create_ex(const plsqr_system *s, int flags, const plsqr_grid *grid, plsqr_op **out) {
    if (!out) return PLSQR_ERR_ARG;
    *out = nullptr;
    if (!s) return PLSQR_ERR_ARG;
    if (s->m_k <= 0 || s->n <= 0 || s->m_d < 0) return PLSQR_ERR_ARG;
    if (s->nnz_k < 0 || s->nnz_d < 0 || s->nnz_k + s->nnz_d >= (1LL << 31) - 1) return PLSQR_ERR_ARG;
    plsqr_op *op = new plsqr_op();
    op->flags = flags;
    op->exact = (flags & PLSQR_OP_EXACT_ORDER) != 0;
    op->m_k = s->m_k;
    op->m_d = s->m_d;
    op->n = s->n;
    op->m = s->m_k + s->m_d;
    op->nnz_k = s->nnz_k;
    op->nnz_d = s->nnz_d;
    int rc = PLSQR_OK;
    auto fail = [&](int code) {
        // keep the handle so the caller can read the message; it holds no device memory
        for (void *p : op->allocs) (void)hipFree(p);
        op->allocs.clear();
        op->bytes = 0;
        free_host_staging(op);
        if (op->own_stream && op->stream) (void)hipStreamDestroy(op->stream);
        op->stream = nullptr;
        op->own_stream = false;
        *out = op;
        return code;
    };
    if ((rc = check_ptr(op, "kernel col_ptr", s->k_col_ptr, s->n, s->nnz_k)) ||
        (rc = check_idx(op, "kernel row_idx", s->k_row_idx, s->k_val, s->nnz_k, s->m_k)))
        return fail(rc);
    if (s->m_d > 0) {
        if ((rc = check_ptr(op, "damping row_ptr", s->d_row_ptr, s->m_d, s->nnz_d)) ||
            (rc = check_idx(op, "damping col_idx", s->d_col_idx, s->d_val, s->nnz_d, s->n)) ||
            (rc = check_ptr(op, "damping col_ptr", s->dc_col_ptr, s->n, s->nnz_d)) ||
            (rc = check_idx(op, "damping-by-column row_idx", s->dc_row_idx, s->dc_val, s->nnz_d, s->m_d)))
            return fail(rc);
    } else if (s->nnz_d != 0 || !s->d_row_ptr || !s->dc_col_ptr) {
        return fail(set_err(op, PLSQR_ERR_ARG, "empty damping needs nnz_d == 0 and zero pointer arrays"));
    }
    if (grid && (long long)grid->nx * grid->ny * grid->nz != s->n)
        return fail(set_err(op, PLSQR_ERR_ARG, "grid dimensions do not multiply to the number of columns"));
    // row slices of A_k, balanced by nonzeros (one per XCD)
    int nsl = env_int("PLSQR_SLICES", 8);
    nsl = std::max(1, std::min(nsl, 64));
    if (s->m_k < nsl * 64) nsl = 1;
    std::vector<int> hb = slice_bounds_host(s, nsl);
    // tiled mode-2 layout on bricks of voxels (default operator only)
    std::unique_ptr<plsqr::TileLayout> L, L1;  // mode 2; mode 1 on stacked bricks (or none: L)
    std::unique_ptr<plsqr::RowTiles> R;
    std::unique_ptr<plsqr::RaySorted> rays;
    const plsqr_system *sys = s;  // s, or its ray-sorted renumbering
    int tw = TILE_W;
    if (!op->exact && env_int("PLSQR_TILES", 1) != 0) {
        std::vector<int> perm, tiles, stacks;
        if (grid) {
            // bricks of 4 x 8 x 8 voxels (PLSQR_BRICK=128: 4 x 4 x 8)
            const bool big = env_int("PLSQR_BRICK", 256) != 128;
            tw = big ? 256 : TILE_W;
            const int order = env_int("PLSQR_BRICKS", 1);
            // mode 1 on super-bricks of the bricks (PLSQR_M1_STACK: 0 none, 1
            // two stacked bricks, 4 x 8 x 16; 2 four, 8 x 8 x 16; 3, the
            // default, two side by side, 8 x 8 x 8): every super-brick is a
            // run of consecutive bricks, so its tile is a union of brick tiles
            // and one column order serves both modes.  At 67^3 / 2M rays a
            // ray crosses 13.5, 10.3 and 12.6 of them (15.9 bricks); 8 x 8 x 8
            // gives the fastest mode 1 (the 1024-voxel tiles shorten the
            // per-ray combine most but need two staging rounds per group)
            const int stack = big && order == 1 ? env_int("PLSQR_M1_STACK", 3) : 0;
            if (stack >= 1 && stack <= 3) {
                const int qx = stack == 1 ? 1 : 2, qz = stack == 3 ? 1 : 2;
                plsqr::brick_order_grouped(grid->nx, grid->ny, grid->nz, 4, 8, 8, qx, 1, qz, perm, tiles, stacks);
                bool sub = true;
                for (int b : stacks) sub = sub && std::binary_search(tiles.begin(), tiles.end(), b);
                if (!sub) return fail(set_err(op, PLSQR_ERR_STATE, "super-bricks that are not unions of bricks"));
            } else {
                plsqr::brick_order(grid->nx, grid->ny, grid->nz, 4, big ? 8 : 4, 8, order, perm, tiles);
            }
        } else {
            perm.resize((size_t)s->n);
            for (int i = 0; i < s->n; ++i) perm[(size_t)i] = i;
            for (int c = 0; c < s->n; c += TILE_W) tiles.push_back(c);
            tiles.push_back(s->n);
        }
        if (env_int("PLSQR_VOXSORT", 1) != 0) plsqr::order_tile_columns(s, tiles, perm);
        if (env_int("PLSQR_RAYSORT", 1) != 0) {
            // within the row slices, so hb stays valid; PLSQR_RAYKEY=0 sorts by
            // first and last tile even when the grid is known
            rays.reset(new plsqr::RaySorted());
            plsqr::sort_rays(s, perm, tiles, hb, env_int("PLSQR_RAYKEY", 1) != 0 ? grid : nullptr, rays.get());
            sys = &rays->sys;
        }
        L.reset(new plsqr::TileLayout());
        std::string why;
        if (!plsqr::build_tile_layout(sys, perm, tiles, hb, tw, L.get(), &why)) {
            L.reset();  // falls back to the sliced CSC
        } else if (env_int("PLSQR_TILE_CHECK", 0) != 0 && !plsqr::tile_layout_self_check(sys, *L, &why)) {
            return fail(set_err(op, PLSQR_ERR_STATE, why));
        }
        if (L && !stacks.empty()) {
            L1.reset(new plsqr::TileLayout());
            int tw1 = 0;  // the widest super-brick, rounded up to a kernel tile width
            for (size_t k = 0; k + 1 < stacks.size(); ++k) tw1 = std::max(tw1, stacks[k + 1] - stacks[k]);
            tw1 = tw1 <= 512 ? 512 : 1024;
            if (!plsqr::build_tile_layout(sys, perm, stacks, hb, tw1, L1.get(), &why)) L1.reset();
        }
        if (L && env_int("PLSQR_ROWTILES", 1) != 0) {
            R.reset(new plsqr::RowTiles());
            // segments in ray order inside each item; PLSQR_SEGWIN=w deals them
            // to the lanes sorted by length inside windows of w (the partials
            // keep ray order): the lanes of a group finish together, but pass 1
            // is slower at 67^3/2M rays (FP64 0.408 -> 0.472 ms, FP32 0.246 ->
            // 0.268 ms at w = 1024): neighbouring lanes no longer read the same
            // x entries, and the groups of long segments take two LDS rounds
            if (L1 && !plsqr::build_row_tiles(*L1, sys->m_k, R.get(), &why, env_int("PLSQR_SEGWIN", 1)))
                L1.reset();  // a segment too long for the stacks: mode 1 on the bricks
            const plsqr::TileLayout &LR = L1 ? *L1 : *L;
            if (!L1 && !plsqr::build_row_tiles(*L, sys->m_k, R.get(), &why, env_int("PLSQR_SEGWIN", 1))) {
                R.reset();  // mode 1 stays on the CSR kernel
            } else if (env_int("PLSQR_TILE_CHECK", 0) != 0 && !plsqr::row_tiles_self_check(sys, LR, *R, &why)) {
                return fail(set_err(op, PLSQR_ERR_STATE, why));
            }
        }
        // mode 2 reads the slots by position: PLSQR_COMPACT=1 in packed rows,
        // 2 in TW-wide rows, 0 by voxel
        const int form = env_int("PLSQR_COMPACT", 1);
        if (L && (form == 1 || form == 2)) {
            plsqr::compact_tile_layout(L.get(), form == 1);
            if (env_int("PLSQR_TILE_CHECK", 0) != 0 && !plsqr::tile_layout_self_check(sys, *L, &why))
                return fail(set_err(op, PLSQR_ERR_STATE, why));
        }
    }
    hipError_t e = hipGetDevice(&op->dev);
    if (e != hipSuccess) return fail(set_err(op, PLSQR_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e)));
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, op->dev);
    if (e != hipSuccess)
        return fail(set_err(op, PLSQR_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e)));
    op->cu = std::max(1, prop.multiProcessorCount);
    e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(set_err(op, PLSQR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    op->own_stream = true;
    if ((rc = build(op, sys, hb, L.get(), L1 ? L1.get() : L.get(), R.get()))) return fail(rc);
    R.reset();
    L1.reset();
    L.reset();
    if (rays) {
        op->rowperm = true;
        if ((rc = dalloc(op, &op->rperm, op->m)) || (rc = dalloc(op, &op->rinv, op->m)) ||
            (rc = dalloc(op, &op->hy2, op->m)) || (rc = dalloc(op, &op->bint, op->m_k)))
            return fail(rc);
        e = hipMemcpy(op->rperm, rays->rperm.data(), sizeof(int) * op->m, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(op->rinv, rays->rinv.data(), sizeof(int) * op->m, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(set_err(op, PLSQR_ERR_HIP, std::string("upload: ") + hipGetErrorString(e)));
        rays.reset();
    }
    // vectors and scratch
    op->nt_loads = env_int("PLSQR_NT", 0) != 0;
    op->rows_nt = env_int("PLSQR_ROWS_NT", 0) != 0;
    op->rt_lc = env_int("PLSQR_LC", 4);
    op->L1 = env_int("PLSQR_L1", 16);
    op->L2 = env_int("PLSQR_L2", 16);
    op->G1 = grid_for((long long)op->m_k * op->L1, BS, env_int("PLSQR_G1", op->cu * 8));
    if (op->rtiled) op->G1 = std::max(op->tS, op->G1 / op->tS * op->tS);  // the combine deals slices by b % S
    {
        int q = grid_for((long long)op->n * op->L2, BS, std::max(1, env_int("PLSQR_G2", op->cu * 8) / op->nsl));
        op->G2 = q * op->nsl;
    }
    op->G4 = grid_for(op->n, BS, op->cu * 4);
    op->G6 = grid_for(op->n, BS, op->cu * 4);
    op->Gs = grid_for(op->m, BS, op->cu * 4);
    op->part_len = 2 * std::max({op->G1, op->G4, op->G6, op->Gs, 1});
    if ((rc = dalloc(op, &op->u, op->m)) || (rc = dalloc(op, &op->v, op->n)) || (rc = dalloc(op, &op->x, op->n)) ||
        (rc = dalloc(op, &op->w, op->n)) || (rc = dalloc(op, &op->bst, op->m_k)) ||
        (rc = dalloc(op, &op->P, (size_t)op->nsl * op->n)) || (rc = dalloc(op, &op->part, op->part_len)) ||
        (rc = dalloc(op, &op->sc, 1)) || (rc = dalloc(op, &op->hx, op->n)) || (rc = dalloc(op, &op->hy, op->m)) ||
        (rc = alloc_host_staging(op)))
        return fail(rc);
    e = hipMemsetAsync(op->u, 0, sizeof(double) * op->m, op->stream);
    if (e == hipSuccess) e = hipMemsetAsync(op->v, 0, sizeof(double) * op->n, op->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(op->stream);
    if (e != hipSuccess) return fail(set_err(op, PLSQR_ERR_HIP, std::string("init: ") + hipGetErrorString(e)));
    *out = op;
    return PLSQR_OK;
}

void plsqr_op_destroy(plsqr_op *op) {
    if (!op) return;
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    for (void *p : op->allocs) (void)hipFree(p);
    free_host_staging(op);
    if (op->own_stream && op->stream) (void)hipStreamDestroy(op->stream);
    delete op;
}

int plsqr_op_dims(const plsqr_op *op, int *m_k, int *m_d, int *n, long long *nnz_k, long long *nnz_d) {
    if (!op) return PLSQR_ERR_ARG;
    if (m_k) *m_k = op->m_k;
    if (m_d) *m_d = op->m_d;
    if (n) *n = op->n;
    if (nnz_k) *nnz_k = op->nnz_k;
    if (nnz_d) *nnz_d = op->nnz_d;
    return PLSQR_OK;
}

size_t plsqr_op_device_bytes(const plsqr_op *op) { return op ? op->bytes : 0; }

int plsqr_op_set_stream(plsqr_op *op, void *stream) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    if (op->own_stream) (void)hipStreamDestroy(op->stream);
    op->own_stream = false;
    op->stream = (hipStream_t)stream;
    return PLSQR_OK;
}

void *plsqr_op_get_stream(const plsqr_op *op) { return op ? (void *)op->stream : nullptr; }

int plsqr_op_set_vectors(plsqr_op *op, const double *u_host, const double *v_host) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (u_host) {
        HIP_TRY(op, hipMemcpyAsync(op->rowperm ? op->hy : op->u, u_host, sizeof(double) * op->m, hipMemcpyHostToDevice,
                                   op->stream));
        if (op->rowperm) {
            int rc = rows_to_internal(op, op->hy, op->u);
            if (rc) return rc;
        }
    }
    if (v_host) {
        HIP_TRY(op, hipMemcpyAsync(op->permuted ? op->hx : op->v, v_host, sizeof(double) * op->n,
                                   hipMemcpyHostToDevice, op->stream));
        if (op->permuted) {
            int rc = to_internal(op, op->hx, op->v);
            if (rc) return rc;
        }
    }
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

int plsqr_op_get_vectors(plsqr_op *op, double *u_host, double *v_host) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (u_host) {
        if (op->rowperm) {
            int rc = rows_to_external(op, op->u, op->hy);
            if (rc) return rc;
        }
        HIP_TRY(op, hipMemcpyAsync(u_host, op->rowperm ? op->hy : op->u, sizeof(double) * op->m, hipMemcpyDeviceToHost,
                                   op->stream));
    }
    if (v_host) {
        if (op->permuted) {
            int rc = to_external(op, op->v, op->hx);
            if (rc) return rc;
        }
        HIP_TRY(op, hipMemcpyAsync(v_host, op->permuted ? op->hx : op->v, sizeof(double) * op->n,
                                   hipMemcpyDeviceToHost, op->stream));
    }
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

int plsqr_aprod_device(plsqr_op *op, int mode, double *x_dev, double *y_dev) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!x_dev || !y_dev) return set_err(op, PLSQR_ERR_ARG, "device pointers required");
    if (mode != 1 && mode != 2) return set_err(op, PLSQR_ERR_ARG, "mode must be 1 or 2");
    return aprod_ext_dev(op, mode, x_dev, y_dev);
}

static int ensure_fp32(plsqr_op *op);

// FP32 mode 2 on the tiled layout (partials into op->P)
static void launch_cols_tiled32(plsqr_op *op, const float *u, const Scal *sc) {
    using namespace plsqr;
    const SlotRows sr{op->t_pvox, op->t_rbase, op->t_roff};
#define COLS32(TW, B, CAP, FORM)                                                                                      \
    hipLaunchKernelGGL((k_cols_tiled32<TW, B, CAP, FORM>), dim3(op->tT * op->tS), dim3(B), 0, op->stream, op->n,      \
                       op->tS, op->t_start, op->t_slot, op->t_seg, op->t_vlen, op->t_ray, op->t_sval32, op->t_sseg,   \
                       op->t_cut, op->t_vcut, sr, u, sc, op->P);
    // 512 threads (two per voxel) on 256-voxel tiles: 36 KB of LDS, so four
    // items share a CU and one item's staging overlaps the others' slot loops
    if (op->tw == 256) {
        if (op->t_form == 1) {
            COLS32(256, 512, 8192, 1)
        } else if (op->t_form == 2) {
            COLS32(256, 512, 8192, 2)
        } else {
            COLS32(256, 512, 8192, 0)
        }
    } else if (op->t_form == 1) {
        COLS32(TILE_W, 512, 4096, 1)
    } else if (op->t_form == 2) {
        COLS32(TILE_W, 512, 4096, 2)
    } else {
        COLS32(TILE_W, 512, 4096, 0)
    }
#undef COLS32
}

// FP32 aprod on internal-order vectors (x: n, y: m floats) with the plain Scal
// FP32 per-ray combine with rt_lc lanes per ray
static void launch_combine32(plsqr_op *op, const float *x, float *y, const plsqr::Scal *sc) {
    using namespace plsqr;
#define COMB32(LC)                                                                                                 \
    hipLaunchKernelGGL((k_rows_tiled_combine32<LC, BS>), dim3(op->G1), dim3(BS), 0, op->stream, op->m_k, op->m,  \
                       op->tS, op->bounds, op->rt_rptr, op->rt_rseg, (const float *)op->rt_segv, op->rp, op->ci,  \
                       op->rv32, x, y, sc, op->part)
    switch (op->rt_lc) {
        case 2: COMB32(2); break;
        case 8: COMB32(8); break;
        default: COMB32(4); break;
    }
#undef COMB32
}

static int aprod32_internal(plsqr_op *op, int mode, float *x, float *y) {
    using namespace plsqr;
    const Scal *sp = op->sc_plain;
    if (mode == 1) {
        if (op->rtiled) {
            launch_pass1<float>(op, op->rt_val32, x, sp, (float *)op->rt_segv);
            launch_combine32(op, x, y, sp);
        } else {
            hipLaunchKernelGGL((k_rows32<16, BS>), dim3(op->G1), dim3(BS), 0, op->stream, op->m_k, op->m, op->rp,
                               op->ci, op->rv32, x, y, sp, op->part);
        }
    } else {
        launch_cols_tiled32(op, y, sp);
        hipLaunchKernelGGL((k_cols_combine32<BS>), dim3(op->G4), dim3(BS), 0, op->stream, op->n, op->m_k, op->tS,
                           op->P, dcol_start(op), dcol_row(op), dcol_val(op), y, sp, x, op->part,
                           (const double *)nullptr, (const uint8_t *)op->dlen);
    }
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

static int ensure_aprod32(plsqr_op *op) {
    int rc;
    if ((rc = ensure_fp32(op))) return rc;
    if (op->sc_plain) return PLSQR_OK;
    if ((rc = dalloc(op, &op->fx, op->n)) || (rc = dalloc(op, &op->fy, op->m)) || (rc = dalloc(op, &op->sc_plain, 1)))
        return rc;
    plsqr::Scal z;
    memset(&z, 0, sizeof(z));
    z.ib = 1.0;
    z.alpha = -1.0;
    z.vscale = 1.0;
    z.beta_pos = 1;
    HIP_TRY(op, hipMemcpy(op->sc_plain, &z, sizeof(z), hipMemcpyHostToDevice));
    return PLSQR_OK;
}

int plsqr_aprod_device_f32(plsqr_op *op, int mode, float *x_dev, float *y_dev) {
    using namespace plsqr;
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!x_dev || !y_dev) return set_err(op, PLSQR_ERR_ARG, "device pointers required");
    if (mode != 1 && mode != 2) return set_err(op, PLSQR_ERR_ARG, "mode must be 1 or 2");
    int rc;
    if ((rc = ensure_aprod32(op))) return rc;
    const int gn = grid_for(op->n, 256, op->cu * 4), gm = grid_for(op->m, 256, op->cu * 4);
    // caller order -> device order
    if (op->permuted)
        hipLaunchKernelGGL(k_gather32, dim3(gn), dim3(256), 0, op->stream, op->n, op->inv, x_dev, op->v32);
    else
        HIP_TRY(op, hipMemcpyAsync(op->v32, x_dev, sizeof(float) * op->n, hipMemcpyDeviceToDevice, op->stream));
    if (op->rowperm) {
        hipLaunchKernelGGL(k_gather32, dim3(gm), dim3(256), 0, op->stream, op->m, op->rperm, y_dev, op->u32);
    } else {
        HIP_TRY(op, hipMemcpyAsync(op->u32, y_dev, sizeof(float) * op->m, hipMemcpyDeviceToDevice, op->stream));
    }
    LAUNCH_CHECK(op);
    if ((rc = aprod32_internal(op, mode, op->v32, op->u32))) return rc;
    // back to the caller's order: only the output moves
    if (mode == 2) {
        if (op->permuted)
            hipLaunchKernelGGL(k_gather32, dim3(gn), dim3(256), 0, op->stream, op->n, op->perm, op->v32, x_dev);
        else
            HIP_TRY(op, hipMemcpyAsync(x_dev, op->v32, sizeof(float) * op->n, hipMemcpyDeviceToDevice, op->stream));
    } else if (op->rowperm) {
        hipLaunchKernelGGL(k_gather32, dim3(gm), dim3(256), 0, op->stream, op->m, op->rinv, op->u32, y_dev);
    } else {
        HIP_TRY(op, hipMemcpyAsync(y_dev, op->u32, sizeof(float) * op->m, hipMemcpyDeviceToDevice, op->stream));
    }
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

static int check_internal_args(plsqr_op *op, const void *a, const void *b) {
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!a || !b || a == b) return set_err(op, PLSQR_ERR_ARG, "two distinct device vectors required");
    return PLSQR_OK;
}

int plsqr_op_to_internal(plsqr_op *op, int space, const double *ext_dev, double *int_dev) {
    if (!op) return PLSQR_ERR_ARG;
    int rc;
    if ((rc = check_internal_args(op, ext_dev, int_dev))) return rc;
    if (space == 1) return rows_to_internal(op, ext_dev, int_dev);
    if (space != 0) return set_err(op, PLSQR_ERR_ARG, "space must be 0 (columns) or 1 (rows)");
    if (op->permuted) return to_internal(op, ext_dev, int_dev);
    HIP_TRY(op, hipMemcpyAsync(int_dev, ext_dev, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
    return PLSQR_OK;
}

int plsqr_op_to_external(plsqr_op *op, int space, const double *int_dev, double *ext_dev) {
    if (!op) return PLSQR_ERR_ARG;
    int rc;
    if ((rc = check_internal_args(op, int_dev, ext_dev))) return rc;
    if (space == 1) return rows_to_external(op, int_dev, ext_dev);
    if (space != 0) return set_err(op, PLSQR_ERR_ARG, "space must be 0 (columns) or 1 (rows)");
    if (op->permuted) return to_external(op, int_dev, ext_dev);
    HIP_TRY(op, hipMemcpyAsync(ext_dev, int_dev, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
    return PLSQR_OK;
}

int plsqr_aprod_device_internal(plsqr_op *op, int mode, double *x_dev, double *y_dev) {
    if (!op) return PLSQR_ERR_ARG;
    int rc;
    if ((rc = check_internal_args(op, x_dev, y_dev))) return rc;
    if (mode != 1 && mode != 2) return set_err(op, PLSQR_ERR_ARG, "mode must be 1 or 2");
    return mode == 1 ? aprod1_dev(op, x_dev, y_dev) : aprod2_dev(op, x_dev, y_dev);
}

int plsqr_aprod_device_internal_f32(plsqr_op *op, int mode, float *x_dev, float *y_dev) {
    if (!op) return PLSQR_ERR_ARG;
    int rc;
    if ((rc = check_internal_args(op, x_dev, y_dev))) return rc;
    if (mode != 1 && mode != 2) return set_err(op, PLSQR_ERR_ARG, "mode must be 1 or 2");
    if ((rc = ensure_aprod32(op))) return rc;
    return aprod32_internal(op, mode, x_dev, y_dev);
}

int plsqr_aprod_f32(plsqr_op *op, int mode, float *x, float *y) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!x || !y) return set_err(op, PLSQR_ERR_ARG, "host vectors required");
    if (mode != 1 && mode != 2) return set_err(op, PLSQR_ERR_ARG, "mode must be 1 or 2");
    int rc;
    if ((rc = ensure_aprod32(op))) return rc;
    HIP_TRY(op, hipMemcpyAsync(op->fx, x, sizeof(float) * op->n, hipMemcpyHostToDevice, op->stream));
    HIP_TRY(op, hipMemcpyAsync(op->fy, y, sizeof(float) * op->m, hipMemcpyHostToDevice, op->stream));
    if ((rc = plsqr_aprod_device_f32(op, mode, op->fx, op->fy))) return rc;
    if (mode == 1)
        HIP_TRY(op, hipMemcpyAsync(y, op->fy, sizeof(float) * op->m, hipMemcpyDeviceToHost, op->stream));
    else
        HIP_TRY(op, hipMemcpyAsync(x, op->fx, sizeof(float) * op->n, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

int plsqr_aprod_split(plsqr_op *op, int mode, double *x, double *y_k, double *y_d) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (mode != 1 && mode != 2) return set_err(op, PLSQR_ERR_ARG, "mode must be 1 or 2");
    if (!x || !y_k || (op->m_d > 0 && !y_d)) return set_err(op, PLSQR_ERR_ARG, "host vectors required");
    HIP_TRY(op, hipMemcpyAsync(op->hx, x, sizeof(double) * op->n, hipMemcpyHostToDevice, op->stream));
    HIP_TRY(op, hipMemcpyAsync(op->hy, y_k, sizeof(double) * op->m_k, hipMemcpyHostToDevice, op->stream));
    if (op->m_d > 0)
        HIP_TRY(op, hipMemcpyAsync(op->hy + op->m_k, y_d, sizeof(double) * op->m_d, hipMemcpyHostToDevice, op->stream));
    int rc = aprod_ext_dev(op, mode, op->hx, op->hy);
    if (rc) return rc;
    if (mode == 1) {
        HIP_TRY(op, hipMemcpyAsync(y_k, op->hy, sizeof(double) * op->m_k, hipMemcpyDeviceToHost, op->stream));
        if (op->m_d > 0)
            HIP_TRY(op, hipMemcpyAsync(y_d, op->hy + op->m_k, sizeof(double) * op->m_d, hipMemcpyDeviceToHost, op->stream));
    } else {
        HIP_TRY(op, hipMemcpyAsync(x, op->hx, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
    }
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

static int aprod_host(plsqr_op *op, int mode, int m, int n, double *x, double *y) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (mode != 1 && mode != 2) return set_err(op, PLSQR_ERR_ARG, "mode must be 1 or 2");
    if (m != op->m || n != op->n)
        return set_err(op, PLSQR_ERR_ARG, "aprod sizes m=" + std::to_string(m) + " n=" + std::to_string(n) +
                                              " do not match the handle (m=" + std::to_string(op->m) +
                                              " n=" + std::to_string(op->n) + ")");
    double *xd = op->v, *yd = op->u;
    if (y) {
        HIP_TRY(op, hipMemcpyAsync(op->hy, y, sizeof(double) * op->m, hipMemcpyHostToDevice, op->stream));
        yd = op->hy;
    }
    if (x) {
        HIP_TRY(op, hipMemcpyAsync(op->hx, x, sizeof(double) * op->n, hipMemcpyHostToDevice, op->stream));
        xd = op->hx;
    }
    // caller arrays are in the caller's order, the resident u and v in the device order
    const int rc = aprod_any(op, mode, xd, x != nullptr, yd, y != nullptr);
    if (rc) return rc;
    if (mode == 1 && y) HIP_TRY(op, hipMemcpyAsync(y, yd, sizeof(double) * op->m, hipMemcpyDeviceToHost, op->stream));
    if (mode == 2 && x) HIP_TRY(op, hipMemcpyAsync(x, xd, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

void plsqr_aprod_ref(int version, int procRank, int numProc, int mode, int m, int n, double x[], double y[],
                     void *UsrWrk, int isFinalIteration, int iteration) {
    (void)version;
    (void)procRank;
    (void)numProc;
    (void)isFinalIteration;
    (void)iteration;
    (void)aprod_host((plsqr_op *)UsrWrk, mode, m, n, x, y);
}

void plsqr_aprod(int mode, int m, int n, double x[], double y[], void *UsrWrk) {
    plsqr_op *op = (plsqr_op *)UsrWrk;
    if (!op) return;
    if (!x || !y) {
        set_err(op, PLSQR_ERR_ARG, "plsqr_aprod needs host x and y");
        return;
    }
    (void)aprod_host(op, mode, m, n, x, y);
}

// ---------------------------------------------------------------------------
// device-resident LSQR
// ---------------------------------------------------------------------------
// cs: column scales (an FP64 LSQR on A D), or NULL
static int enqueue_iteration(plsqr_op *op, double *x, double *w, double *se, double *trace_d,
                             const double *cs = nullptr) {
    using namespace plsqr;
    int rc;
    if (op->exact) {
        hipLaunchKernelGGL(k_rows_exact, dim3(grid_for(op->m, BS, op->cu * 8)), dim3(BS), 0, op->stream, op->m, op->rp,
                           op->ci, op->rv, op->v, op->u, (const Scal *)op->sc);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_dnrm2_seq, dim3(1), dim3(64), 0, op->stream, op->m_k, op->u, &op->sc->pk,
                           (const Scal *)op->sc, 0);
        hipLaunchKernelGGL(k_dnrm2_seq, dim3(1), dim3(64), 0, op->stream, op->m_d, op->u + op->m_k, &op->sc->pd,
                           (const Scal *)op->sc, 0);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_beta, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, 0);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_cols_exact, dim3(grid_for(op->n, BS, op->cu * 8)), dim3(BS), 0, op->stream, op->n,
                           op->m_k, op->kcp, op->kri, op->kv, op->dcp, op->dri, op->dcv, op->u, (const Scal *)op->sc,
                           op->v);
        hipLaunchKernelGGL(k_dnrm2_seq, dim3(1), dim3(64), 0, op->stream, op->n, op->v, &op->sc->pv,
                           (const Scal *)op->sc, 1);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_alpha, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, 0);
        hipLaunchKernelGGL(k_dknorm_seq, dim3(1), dim3(64), 0, op->stream, op->n, w, op->sc);
        hipLaunchKernelGGL((k_update<BS>), dim3(op->G6), dim3(BS), 0, op->stream, op->n, (const Scal *)op->sc, op->v,
                           w, x, se, (double *)nullptr);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_est, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, 0, trace_d);
        LAUNCH_CHECK(op);
        return PLSQR_OK;
    }
    if ((rc = launch_rows(op, op->v, op->u, op->sc, op->part))) return rc;
    hipLaunchKernelGGL(k_s_beta, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G1);
    LAUNCH_CHECK(op);
    if ((rc = launch_cols(op, op->u, op->v, op->sc, op->part, cs))) return rc;
    hipLaunchKernelGGL(k_s_alpha, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G4);
    LAUNCH_CHECK(op);
    hipLaunchKernelGGL((k_update<BS>), dim3(op->G6), dim3(BS), 0, op->stream, op->n, (const Scal *)op->sc, op->v, w,
                       x, se, op->part);
    LAUNCH_CHECK(op);
    hipLaunchKernelGGL(k_s_est, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G6, trace_d);
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

// u_full: NULL for u = [b;0], else the full m-vector to start from (a warm
// start's residual r0; the solve then returns the correction).
// se: NULL, or n doubles (device order) that receive the standard-error
// estimates of lsqr.c:644-670, 878-891.
static int solve_fp64(plsqr_op *op, const double *b_dev, const plsqr_solve_opts *o, double *x, double *w,
                      plsqr_solve_out *out, double *trace_host, const double *u_full = nullptr,
                      double *se = nullptr) {
    using namespace plsqr;
    int rc;
    if (o->precond != PLSQR_PRECOND_NONE)
        return set_err(op, PLSQR_ERR_ARG, "the column-scaled solve has no trace or standard errors");
    const int itn = std::max(0, o->itn);
    const bool want_trace = trace_host != nullptr;
    if (want_trace && (rc = ensure_trace(op, itn + 1))) return rc;
    double *trace_d = want_trace ? op->trace_d : nullptr;
    Scal h;
    memset(&h, 0, sizeof(h));
    h.damp = o->damp;
    h.atol = o->atol;
    h.btol = o->btol;
    h.ctol = o->conlim > 0.0 ? 1.0 / o->conlim : 0.0;
    h.tol = o->tol;
    h.itnlim = o->itnlim;
    h.itn_max = itn;
    h.exact = op->exact ? 1 : 0;
    h.cs2 = -1.0;  // lsqr.c:293
    HIP_TRY(op, hipMemcpyAsync(op->sc, &h, sizeof(h), hipMemcpyHostToDevice, op->stream));
    // u = [b ; 0], v = x = w = 0 (tridiagonal.c:33-57, lsqr.c:311-312)
    const int gz = grid_for(std::max(op->m, op->n), 256, op->cu * 4);
    if (u_full) {
        HIP_TRY(op, hipMemcpyAsync(op->u, u_full, sizeof(double) * op->m, hipMemcpyDeviceToDevice, op->stream));
    } else {
        HIP_TRY(op, hipMemcpyAsync(op->u, b_dev, sizeof(double) * op->m_k, hipMemcpyDeviceToDevice, op->stream));
        hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->m_d, 0.0, op->u + op->m_k);
    }
    hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, op->v);
    hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, x);
    hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, w);
    if (se) hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, se);
    LAUNCH_CHECK(op);
    // beta = ||u||
    if (op->exact) {
        hipLaunchKernelGGL(k_dnrm2_seq, dim3(1), dim3(64), 0, op->stream, op->m_k, op->u, &op->sc->pk,
                           (const Scal *)nullptr, 0);
        hipLaunchKernelGGL(k_dnrm2_seq, dim3(1), dim3(64), 0, op->stream, op->m_d, op->u + op->m_k, &op->sc->pd,
                           (const Scal *)nullptr, 0);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_init_beta, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, 0);
    } else {
        // partials laid out as [kernel, damping] pairs, like k_rows
        const int g = std::min(op->part_len / 2, grid_for(op->m, BS, op->cu * 4));
        hipLaunchKernelGGL((k_ssq2<BS>), dim3(g), dim3(BS), 0, op->stream, op->m_k, op->m, op->u, op->part);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_init_beta, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, g);
    }
    LAUNCH_CHECK(op);
    // v = A^T u (scaled by 1/beta on the fly), alpha
    if (op->exact) {
        hipLaunchKernelGGL(k_cols_exact, dim3(grid_for(op->n, BS, op->cu * 8)), dim3(BS), 0, op->stream, op->n,
                           op->m_k, op->kcp, op->kri, op->kv, op->dcp, op->dri, op->dcv, op->u, (const Scal *)op->sc,
                           op->v);
        hipLaunchKernelGGL(k_dnrm2_seq, dim3(1), dim3(64), 0, op->stream, op->n, op->v, &op->sc->pv,
                           (const Scal *)op->sc, 1);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_init_alpha, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, 0, trace_d);
    } else {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(op->part_len, 256, 64)), dim3(256), 0, op->stream,
                           (long long)op->part_len, 0.0, op->part);
        if ((rc = launch_cols(op, op->u, op->v, op->sc, op->part))) return rc;
        hipLaunchKernelGGL(k_s_init_alpha, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G4, trace_d);
    }
    LAUNCH_CHECK(op);
    hipLaunchKernelGGL(k_init_vw, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, op->n,
                       (const Scal *)op->sc, op->v, w);
    LAUNCH_CHECK(op);
    // main loop, timed with events on the handle's stream
    SolveScope scope(op);
    if ((rc = scope.create_events())) return rc;
    HIP_TRY(op, hipEventRecord(scope.e0, op->stream));
    const bool use_graph = env_int("PLSQR_GRAPH", 1) != 0 && itn > 1;
    const int check_every = o->tol > 0.0 ? std::max(1, env_int("PLSQR_CHECK_EVERY", 8)) : 0;
    if (use_graph && (rc = scope.capture([&] { return enqueue_iteration(op, x, w, se, trace_d); }))) return rc;
    for (int it = 0; it < itn; ++it) {
        if (use_graph) {
            HIP_TRY(op, hipGraphLaunch(scope.exec, op->stream));
        } else if ((rc = enqueue_iteration(op, x, w, se, trace_d))) {
            return rc;
        }
        if (check_every && (it + 1) % check_every == 0) {
            // the stop flag through the handle's pinned staging
            HIP_TRY(op, hipMemcpyAsync(&op->hscal->done, &op->sc->done, sizeof(int), hipMemcpyDeviceToHost, op->stream));
            HIP_TRY(op, hipStreamSynchronize(op->stream));
            if (op->hscal->done) break;
        }
    }
    HIP_TRY(op, hipEventRecord(scope.e1, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    float ms = 0.f;
    HIP_TRY(op, hipEventElapsedTime(&ms, scope.e0, scope.e1));
    // the loop leaves the last 1/beta of u pending (lsqr.c:538); apply it so
    // that the resident u is the reference's u after the loop
    hipLaunchKernelGGL(k_scale_pending, dim3(grid_for(op->m, 256, op->cu * 4)), dim3(256), 0, op->stream,
                       (long long)op->m, (const Scal *)op->sc, op->u);
    LAUNCH_CHECK(op);
    if ((rc = read_scal(op, &h))) return rc;
    if (se) {
        // se = rnorm / sqrt(m - n, or m when damped, or 1) * sqrt(var)  (lsqr.c:878-891)
        double t = 1.0;
        if (op->m > op->n) t = (double)(op->m - op->n);
        if (h.damp > 0.0) t = (double)op->m;
        t = h.rnorm / std::sqrt(t);
        hipLaunchKernelGGL(k_se_final, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, op->n, t, se);
        LAUNCH_CHECK(op);
    }
    if (out) {
        memset(out, 0, sizeof(*out));
        int istop = h.istop;
        if (h.damp > 0.0 && istop == 2) istop = 3;  // lsqr.c:895-896
        out->istop = h.init_zero ? 0 : istop;
        out->itn = h.itn;
        out->anorm = h.anorm;
        out->acond = h.acond;
        out->rnorm = h.rnorm;
        out->arnorm = h.test2;  // lsqr.c:913
        out->xnorm = h.xnorm;
        out->bnorm = h.bnorm;
        out->arnorm_true = h.arnorm;
        out->elapsed_ms = ms;
    }
    if (want_trace) {
        const int rows = std::min(h.itn, itn) + 1;
        HIP_TRY(op, hipMemcpy(trace_host, op->trace_d, sizeof(double) * 8 * rows, hipMemcpyDeviceToHost));
        for (int r = rows; r <= itn; ++r)
            for (int k = 0; k < 8; ++k) trace_host[(size_t)r * 8 + k] = NAN;
    }
    return PLSQR_OK;
}

static int ensure_fp32(plsqr_op *op) {
    using namespace plsqr;
    if (op->rv32) return PLSQR_OK;
    if (!op->tiled) return set_err(op, PLSQR_ERR_STATE, "mixed precision needs the tiled operator (not the exact-order one)");
    int rc;
    const long long nr = op->nnz_k + op->nnz_d;
    if ((rc = dalloc(op, &op->rv32, nr)) || (rc = dalloc(op, &op->t_sval32, op->t_nslot)) ||
        (rc = dalloc(op, &op->u32, op->m)) || (rc = dalloc(op, &op->v32, op->n)) || (rc = dalloc(op, &op->dx, op->n)))
        return rc;
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(nr, 256, op->cu * 16)), dim3(256), 0, op->stream, nr, op->rv, op->rv32);
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(op->t_nslot, 256, op->cu * 16)), dim3(256), 0, op->stream,
                       op->t_nslot, op->t_sval, op->t_sval32);
    LAUNCH_CHECK(op);
    if (op->rtiled) {
        if ((rc = dalloc(op, &op->rt_val32, op->rt_nent + 4))) return rc;
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(op->rt_nent, 256, op->cu * 16)), dim3(256), 0, op->stream,
                           op->rt_nent, op->rt_val, op->rt_val32);
        LAUNCH_CHECK(op);
    }
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

static int enqueue_iteration32(plsqr_op *op) {
    using namespace plsqr;
    dim3 g1(op->G1), b(BS);
    if (op->rtiled) {
        launch_pass1<float>(op, op->rt_val32, op->v32, (const Scal *)op->sc, (float *)op->rt_segv);
        launch_combine32(op, op->v32, op->u32, (const Scal *)op->sc);
    } else switch (op->L1) {
        case 8: hipLaunchKernelGGL((k_rows32<8, BS>), g1, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv32, op->v32, op->u32, (const Scal *)op->sc, op->part); break;
        case 32: hipLaunchKernelGGL((k_rows32<32, BS>), g1, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv32, op->v32, op->u32, (const Scal *)op->sc, op->part); break;
        default: hipLaunchKernelGGL((k_rows32<16, BS>), g1, b, 0, op->stream, op->m_k, op->m, op->rp, op->ci, op->rv32, op->v32, op->u32, (const Scal *)op->sc, op->part); break;
    }
    LAUNCH_CHECK(op);
    hipLaunchKernelGGL(k_s_beta, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G1);
    launch_cols_tiled32(op, op->u32, (const Scal *)op->sc);
    hipLaunchKernelGGL((k_cols_combine32<BS>), dim3(op->G4), dim3(BS), 0, op->stream, op->n, op->m_k, op->tS, op->P,
                       dcol_start(op), dcol_row(op), dcol_val(op), op->u32, (const Scal *)op->sc, op->v32, op->part,
                       op->cs_active, (const uint8_t *)op->dlen);
    LAUNCH_CHECK(op);
    hipLaunchKernelGGL(k_s_alpha, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G4);
    hipLaunchKernelGGL((k_update32<BS>), dim3(op->G6), dim3(BS), 0, op->stream, op->n, (const Scal *)op->sc, op->v32,
                       op->w, op->dx, op->part);
    hipLaunchKernelGGL(k_s_est, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G6, (double *)nullptr);
    LAUNCH_CHECK(op);
    return PLSQR_OK;
}

// ||A^T r|| and ||r|| in FP64 for r = [b;0] - A xt (r into op->hy, A^T r
// into op->hx)
static int true_norms(plsqr_op *op, const double *b_dev, const double *xt, double *rn, double *an) {
    using namespace plsqr;
    int rc;
    const int gz = grid_for(std::max(op->m, op->n), 256, op->cu * 4);
    const int gs = std::min(op->part_len / 2, grid_for(op->m, BS, op->cu * 4));
    const int gn = std::min(op->part_len, grid_for(op->n, BS, op->cu * 4));
    if ((rc = launch_residual(op, xt, b_dev, op->hy))) return rc;
    hipLaunchKernelGGL((k_ssq<BS>), dim3(gs), dim3(BS), 0, op->stream, (long long)op->m, (const double *)op->hy,
                       op->part);
    LAUNCH_CHECK(op);
    if ((rc = host_sum(op, op->part, gs, rn))) return rc;
    *rn = std::sqrt(*rn);
    hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, op->hx);
    LAUNCH_CHECK(op);
    if ((rc = aprod2_dev(op, op->hx, op->hy))) return rc;
    hipLaunchKernelGGL((k_ssq<BS>), dim3(gn), dim3(BS), 0, op->stream, (long long)op->n, (const double *)op->hx,
                       op->part);
    LAUNCH_CHECK(op);
    if ((rc = host_sum(op, op->part, gn, an))) return rc;
    *an = std::sqrt(*an);
    return PLSQR_OK;
}

// Column scales d_j = ||A e_j||^-power of [A_k; D] (device column order)
// and ||A||_F.  The column sums of squares are formed once per handle: the
// tiled mode 2 run on the squared values with u = 1 gives every one in a
// fixed order; the scales are remade when the power changes.
static int ensure_colscale(plsqr_op *op, double power) {
    using namespace plsqr;
    if (op->cscale && op->cs_power == power) return PLSQR_OK;
    if (!op->tiled) return set_err(op, PLSQR_ERR_STATE, "column scaling needs the tiled operator");
    if (!(power > 0.0) || !std::isfinite(power)) return set_err(op, PLSQR_ERR_ARG, "precond_power must be > 0");
    int rc;
    if (op->colss && op->cscale) {
        hipLaunchKernelGGL(k_colscale, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, op->n,
                           (const double *)op->colss, power, op->cscale);
        LAUNCH_CHECK(op);
        HIP_TRY(op, hipStreamSynchronize(op->stream));
        op->cs_power = power;
        return PLSQR_OK;
    }
    double *cs = nullptr, *sq_t = nullptr, *sq_d = nullptr, *ones = nullptr;
    if ((!op->colss && (rc = dalloc(op, &op->colss, op->n))) || (rc = dalloc(op, &cs, op->n)) ||
        (rc = dalloc(op, &sq_t, op->t_nslot)) || (rc = dalloc(op, &sq_d, op->nnz_d)) ||
        (rc = dalloc(op, &ones, op->m)))
        return rc;
    const int gz = grid_for(std::max(op->m, op->n), 256, op->cu * 4);
    hipLaunchKernelGGL(k_square, dim3(grid_for(op->t_nslot, 256, op->cu * 16)), dim3(256), 0, op->stream,
                       op->t_nslot, (const double *)op->t_sval, sq_t);
    hipLaunchKernelGGL(k_square, dim3(grid_for(op->nnz_d, 256, op->cu * 16)), dim3(256), 0, op->stream, op->nnz_d,
                       (const double *)op->dcv, sq_d);
    hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->m, 1.0, ones);
    hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, op->colss);
    LAUNCH_CHECK(op);
    // the squared values stand in for the values for one mode 2; the damping
    // columns are read from the CSC for it
    uint8_t *dl = op->dlen;
    op->dlen = nullptr;
    std::swap(op->t_sval, sq_t);
    std::swap(op->dcv, sq_d);
    rc = launch_cols(op, ones, op->colss, nullptr, nullptr);
    std::swap(op->t_sval, sq_t);
    std::swap(op->dcv, sq_d);
    op->dlen = dl;
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    dfree(op, sq_t, op->t_nslot);
    dfree(op, sq_d, op->nnz_d);
    dfree(op, ones, op->m);
    if (rc) return rc;
    double fro2 = 0.0;
    const long long nnz = op->nnz_k + op->nnz_d;
    const int ga = std::min(op->part_len, grid_for(nnz, BS, op->cu * 4));
    hipLaunchKernelGGL((k_ssq<BS>), dim3(ga), dim3(BS), 0, op->stream, nnz, (const double *)op->rv, op->part);
    LAUNCH_CHECK(op);
    if ((rc = host_sum(op, op->part, ga, &fro2))) return rc;
    hipLaunchKernelGGL(k_colscale, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, op->n,
                       (const double *)op->colss, power, cs);
    LAUNCH_CHECK(op);
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    op->fro = std::sqrt(fro2);
    op->cscale = cs;
    op->cs_power = power;
    return PLSQR_OK;
}

// Mixed precision: an inner LSQR in FP32 (values, u, v; FP64-accumulated dot products; FP64
// scalars, w and correction dx) solves min ||A dx - r|| for r = [b;0] - A x.
// Every refine_every inner iterations (refine_adaptive: first after
// refine_every, then when the inner test2 estimate times the last measured
// ratio FP64 / inner reaches tol, at most 4 refine_every apart) the FP64
// residual of xt = x + dx is
// formed and its test2 = ||A^T r|| / (anorm ||r||) checked in FP64 (anorm:
// the largest LSQR estimate of ||A|| so far); the solve stops there when it
// is <= tol.  The inner run is not interrupted by the check: it keeps its
// Krylov basis, and is restarted on the refined residual (x = xt) only when
// its own recurrence stops (its estimate of test2 reached tol) while the FP64
// test2 has not.  itn bounds the inner iterations in total.
// With PLSQR_PRECOND_COLSCALE the inner LSQR runs on A D (D = diag of the
// column scales ||A e_j||^-precond_power) and the FP64 test2 uses anorm_ref,
// or ||A||_F when that is 0, in place of the inner estimate, which is then
// one of ||A D||.
// x_init: NULL for x = 0, else the starting x in the device column order.
static int solve_mixed(plsqr_op *op, const double *b_dev, const plsqr_solve_opts *o, plsqr_solve_out *out,
                       const double *x_init = nullptr) {
    using namespace plsqr;
    int rc;
    // PLSQR_PREC_FP64 here: the same solve with an FP64 inner LSQR (the
    // FP64 operator on A D), which separates what the scaling gains from
    // what FP32 arithmetic gains
    const bool inner64 = o->precision == PLSQR_PREC_FP64;
    if (!inner64 && (rc = ensure_fp32(op))) return rc;
    if (inner64 && !op->tiled) return set_err(op, PLSQR_ERR_STATE, "column scaling needs the tiled operator");
    if (!op->dx && (rc = dalloc(op, &op->dx, op->n))) return rc;
    if (!op->xt && (rc = dalloc(op, &op->xt, op->n))) return rc;
    double *const xt = op->xt;  // the checked iterate x + dx (n doubles, never x_init)
    const bool colscale = o->precond == PLSQR_PRECOND_COLSCALE;
    if (o->precond != PLSQR_PRECOND_NONE && !colscale) return set_err(op, PLSQR_ERR_ARG, "unknown preconditioner");
    if (colscale && (rc = ensure_colscale(op, o->precond_power == 0.0 ? 1.0 : o->precond_power))) return rc;
    SolveScope scope(op);  // events, graph, exec; resets cs_active on every exit
    op->cs_active = colscale ? op->cscale : nullptr;
    const double anorm_fixed = colscale ? (o->anorm_ref > 0.0 ? o->anorm_ref : op->fro) : 0.0;
    const int itn = std::max(0, o->itn);
    const int period = o->refine_every > 0 ? o->refine_every : std::max(itn, 1);
    const int gz = grid_for(std::max(op->m, op->n), 256, op->cu * 4);
    if (x_init)
        HIP_TRY(op, hipMemcpyAsync(op->x, x_init, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
    else
        hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, op->x);
    LAUNCH_CHECK(op);
    if ((rc = scope.create_events())) return rc;
    HIP_TRY(op, hipEventRecord(scope.e0, op->stream));
    const int gs = std::min(op->part_len / 2, grid_for(op->m, BS, op->cu * 4));
    const Scal *hs = op->hscal;  // pinned copy of the inner state for the polls
    const bool adaptive = o->refine_adaptive != 0 && o->tol > 0.0;
    double ratio = -1.0;  // FP64 test2 / inner estimate at the last check
    double anorm_max = 0.0, rn = 0.0, an = 0.0, test2 = INFINITY;
    int total = 0, starts = 0, checks = 0;
    bool converged = false;
    Scal h;
    // r = [b;0] - A x for the first start
    double rn0 = 0.0, an0 = 0.0;
    if ((rc = true_norms(op, b_dev, op->x, &rn0, &an0))) return rc;
    rn = rn0;
    an = an0;
    if (rn0 == 0.0 || itn == 0) {
        test2 = 0.0;
        converged = rn0 == 0.0;
    }
    while (!converged && total < itn) {
        // (re)start the inner FP32 LSQR on r (in op->hy)
        Scal z;
        memset(&z, 0, sizeof(z));
        z.atol = o->atol;
        z.btol = o->btol;
        z.ctol = o->conlim > 0.0 ? 1.0 / o->conlim : 0.0;
        z.tol = o->tol;
        z.itnlim = o->itnlim;
        z.itn_max = itn - total;
        z.cs2 = -1.0;
        HIP_TRY(op, hipMemcpyAsync(op->sc, &z, sizeof(z), hipMemcpyHostToDevice, op->stream));
        if (inner64) {
            HIP_TRY(op, hipMemcpyAsync(op->u, op->hy, sizeof(double) * op->m, hipMemcpyDeviceToDevice, op->stream));
            hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, op->v);
        } else {
            hipLaunchKernelGGL(k_to_f32, dim3(gz), dim3(256), 0, op->stream, (long long)op->m, op->hy, op->u32);
            hipLaunchKernelGGL(k_fill32, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, op->v32);
        }
        hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, op->dx);
        hipLaunchKernelGGL(k_fill, dim3(gz), dim3(256), 0, op->stream, (long long)op->n, 0.0, op->w);
        hipLaunchKernelGGL((k_ssq2<BS>), dim3(gs), dim3(BS), 0, op->stream, op->m_k, op->m, op->hy, op->part);
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_init_beta, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, gs);
        if (inner64) {
            if ((rc = launch_cols(op, op->u, op->v, op->sc, op->part, op->cs_active))) return rc;
        } else {
            launch_cols_tiled32(op, op->u32, (const Scal *)op->sc);
            hipLaunchKernelGGL((k_cols_combine32<BS>), dim3(op->G4), dim3(BS), 0, op->stream, op->n, op->m_k, op->tS,
                               op->P, dcol_start(op), dcol_row(op), dcol_val(op), op->u32, (const Scal *)op->sc,
                               op->v32, op->part, op->cs_active, (const uint8_t *)op->dlen);
        }
        LAUNCH_CHECK(op);
        hipLaunchKernelGGL(k_s_init_alpha, dim3(1), dim3(256), 0, op->stream, op->sc, op->part, op->G4,
                           (double *)nullptr);
        if (inner64)
            hipLaunchKernelGGL(k_init_vw, dim3(gz), dim3(256), 0, op->stream, op->n, (const Scal *)op->sc, op->v,
                               op->w);
        else
            hipLaunchKernelGGL(k_init_vw32, dim3(gz), dim3(256), 0, op->stream, op->n, (const Scal *)op->sc, op->v32,
                               op->w);
        LAUNCH_CHECK(op);
        ++starts;
        if (!scope.exec && (rc = scope.capture([&] {
                return inner64 ? enqueue_iteration(op, op->dx, op->w, nullptr, nullptr, op->cs_active)
                               : enqueue_iteration32(op);
            })))
            return rc;
        int inner_prev = 0;
        for (;;) {
            // a chunk of inner iterations, polling the inner state every 8.
            // Adaptive: the chunk may run to 4 periods and ends early once the
            // inner estimate of test2 times the last measured ratio FP64/inner
            // reaches tol.
            const int k = std::min(adaptive && ratio > 0.0 ? 4 * period : period, itn - total);
            int launched = 0;
            for (int i = 0; i < k; ++i) {
                HIP_TRY(op, hipGraphLaunch(scope.exec, op->stream));
                ++launched;
                if ((i + 1) % 8 == 0 && i + 1 < k) {
                    HIP_TRY(op, hipMemcpyAsync(op->hscal, op->sc, sizeof(Scal), hipMemcpyDeviceToHost, op->stream));
                    HIP_TRY(op, hipStreamSynchronize(op->stream));
                    if (hs->done) break;
                    if (adaptive && ratio > 0.0 && hs->test2 * ratio <= o->tol) break;
                }
            }
            if ((rc = read_scal(op, &h))) return rc;
            const int inner = std::max(h.itn, 0);
            total += inner - inner_prev;
            inner_prev = inner;
            anorm_max = std::max(anorm_max, h.anorm);
            // FP64 check of xt = x + dx
            hipLaunchKernelGGL(k_add3, dim3(gz), dim3(256), 0, op->stream, op->n, (const double *)op->x,
                               (const double *)op->dx, xt);
            LAUNCH_CHECK(op);
            if ((rc = true_norms(op, b_dev, xt, &rn, &an))) return rc;
            ++checks;
            const double anorm_t = colscale ? anorm_fixed : anorm_max;
            test2 = (rn > 0.0 && anorm_t > 0.0) ? an / (anorm_t * rn) : 0.0;
            if (h.test2 > 0.0 && std::isfinite(h.test2)) ratio = test2 / h.test2;
            if (test2 <= o->tol || rn == 0.0) converged = true;
            if (converged || total >= itn || h.done || launched == 0) {
                // x = xt; op->hy now holds its residual for a restart
                HIP_TRY(op, hipMemcpyAsync(op->x, xt, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
                if (inner == 0 && !converged) {  // the inner run could not start: nothing left to gain
                    converged = test2 <= o->tol;
                    total = itn;
                }
                break;
            }
        }
    }
    HIP_TRY(op, hipEventRecord(scope.e1, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    float ms = 0.f;
    HIP_TRY(op, hipEventElapsedTime(&ms, scope.e0, scope.e1));
    if (out) {
        memset(out, 0, sizeof(*out));
        out->itn = total;
        out->inner_total = total;
        out->refinements = checks;
        out->anorm = colscale ? anorm_fixed : anorm_max;
        out->rnorm = rn;
        out->arnorm = test2;
        out->arnorm_true = an;
        out->bnorm = rn0;
        out->elapsed_ms = ms;
        out->istop = test2 <= o->tol ? 2 : 5;
        (void)starts;
    }
    return PLSQR_OK;
}

int plsqr_lsqr_solve_device(plsqr_op *op, const double *b_dev, const plsqr_solve_opts *opts, double *x_dev,
                            double *w_dev, plsqr_solve_out *out, double *trace) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!b_dev || !opts || !x_dev) return set_err(op, PLSQR_ERR_ARG, "b, opts and x are required");
    plsqr_solve_opts o;
    default_opts(opts, &o);
    int rc;
    if ((rc = b_internal(op, b_dev, &b_dev))) return rc;
    if (o.precision == PLSQR_PREC_MIXED || (o.precision == PLSQR_PREC_FP64 && o.precond != PLSQR_PRECOND_NONE)) {
        if (trace) return set_err(op, PLSQR_ERR_ARG, "no trace for the mixed or column-scaled solves");
        rc = solve_mixed(op, b_dev, &o, out);
    } else if (o.precision == PLSQR_PREC_FP64) {
        // the loop works on the handle's x and w (device column order)
        rc = solve_fp64(op, b_dev, &o, op->x, op->w, out, trace);
    } else {
        return set_err(op, PLSQR_ERR_ARG, "unknown precision");
    }
    if (rc) return rc;
    if (op->permuted) {
        if ((rc = to_external(op, op->x, x_dev))) return rc;
        if (w_dev && (rc = to_external(op, op->w, w_dev))) return rc;
    } else {
        if (x_dev != op->x)
            HIP_TRY(op, hipMemcpyAsync(x_dev, op->x, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
        if (w_dev && w_dev != op->w)
            HIP_TRY(op, hipMemcpyAsync(w_dev, op->w, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
    }
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

int plsqr_lsqr_solve(plsqr_op *op, const double *b_host, const plsqr_solve_opts *opts, double *x_host,
                     double *w_host, plsqr_solve_out *out, double *trace) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!b_host || !opts || !x_host) return set_err(op, PLSQR_ERR_ARG, "b, opts and x are required");
    for (int i = 0; i < op->m_k; ++i)
        if (!std::isfinite(b_host[i])) return set_err(op, PLSQR_ERR_NONFINITE, "b holds a NaN or Inf at " + std::to_string(i));
    HIP_TRY(op, hipMemcpyAsync(op->bst, b_host, sizeof(double) * op->m_k, hipMemcpyHostToDevice, op->stream));
    int rc = plsqr_lsqr_solve_device(op, op->bst, opts, op->hx, nullptr, out, trace);
    if (rc) return rc;
    HIP_TRY(op, hipMemcpyAsync(x_host, op->hx, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    if (w_host) {
        if (op->permuted) {
            if ((rc = to_external(op, op->w, op->hx))) return rc;
            HIP_TRY(op, hipMemcpyAsync(w_host, op->hx, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
        } else {
            HIP_TRY(op, hipMemcpyAsync(w_host, op->w, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
        }
        HIP_TRY(op, hipStreamSynchronize(op->stream));
    }
    return PLSQR_OK;
}

int plsqr_lsqr_solve_se(plsqr_op *op, const double *b_host, const plsqr_solve_opts *opts, double *x_host,
                        double *se_host, plsqr_solve_out *out) {
    using namespace plsqr;
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!b_host || !opts || !x_host || !se_host) return set_err(op, PLSQR_ERR_ARG, "b, opts, x and se are required");
    plsqr_solve_opts o;
    default_opts(opts, &o);
    if (o.precision != PLSQR_PREC_FP64) return set_err(op, PLSQR_ERR_ARG, "standard errors come with the FP64 solve");
    for (int i = 0; i < op->m_k; ++i)
        if (!std::isfinite(b_host[i])) return set_err(op, PLSQR_ERR_NONFINITE, "b holds a NaN or Inf at " + std::to_string(i));
    int rc;
    if (!op->se && (rc = dalloc(op, &op->se, op->n))) return rc;
    HIP_TRY(op, hipMemcpyAsync(op->bst, b_host, sizeof(double) * op->m_k, hipMemcpyHostToDevice, op->stream));
    const double *b_dev = op->bst;
    if ((rc = b_internal(op, b_dev, &b_dev))) return rc;
    if ((rc = solve_fp64(op, b_dev, &o, op->x, op->w, out, nullptr, nullptr, op->se))) return rc;
    const double *xs = op->x, *ss = op->se;
    if (op->permuted) {
        if ((rc = to_external(op, op->x, op->hx)) || (rc = to_external(op, op->se, op->hx2))) return rc;
        xs = op->hx;
        ss = op->hx2;
    }
    HIP_TRY(op, hipMemcpyAsync(x_host, xs, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipMemcpyAsync(se_host, ss, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

int plsqr_lsqr_solve_x0_device(plsqr_op *op, const double *b_dev, const double *x0_dev,
                               const plsqr_solve_opts *opts, double *x_dev, plsqr_solve_out *out, int *used_x0) {
    using namespace plsqr;
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!b_dev || !opts || !x_dev) return set_err(op, PLSQR_ERR_ARG, "b, opts and x are required");
    if (op->exact) return set_err(op, PLSQR_ERR_ARG, "warm starts run on the default operator");
    plsqr_solve_opts o;
    default_opts(opts, &o);
    int rc;
    if ((rc = b_internal(op, b_dev, &b_dev))) return rc;
    if (!op->x0i && ((rc = dalloc(op, &op->x0i, op->n)) || (rc = dalloc(op, &op->r0, op->m)))) return rc;
    int use = 0;
    if (x0_dev) {
        // r0 = [b;0] - A x0; the warm start is kept only if ||r0|| <= ||[b;0]||,
        // the residual of the zero start
        if (op->permuted) {
            if ((rc = to_internal(op, x0_dev, op->x0i))) return rc;
        } else {
            HIP_TRY(op, hipMemcpyAsync(op->x0i, x0_dev, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
        }
        if ((rc = launch_residual(op, op->x0i, b_dev, op->r0))) return rc;
        LAUNCH_CHECK(op);
        const int g = std::min(op->part_len, grid_for(op->m, BS, op->cu * 4));
        double rn2 = 0.0, bn2 = 0.0;
        hipLaunchKernelGGL((k_ssq<BS>), dim3(g), dim3(BS), 0, op->stream, (long long)op->m, (const double *)op->r0,
                           op->part);
        LAUNCH_CHECK(op);
        if ((rc = host_sum(op, op->part, g, &rn2))) return rc;
        hipLaunchKernelGGL((k_ssq<BS>), dim3(g), dim3(BS), 0, op->stream, (long long)op->m_k, b_dev, op->part);
        LAUNCH_CHECK(op);
        if ((rc = host_sum(op, op->part, g, &bn2))) return rc;
        use = std::isfinite(rn2) && rn2 <= bn2;
    }
    if (used_x0) *used_x0 = use;
    if (o.precision == PLSQR_PREC_MIXED || (o.precision == PLSQR_PREC_FP64 && o.precond != PLSQR_PRECOND_NONE)) {
        rc = solve_mixed(op, b_dev, &o, out, use ? op->x0i : nullptr);
    } else if (o.precision == PLSQR_PREC_FP64) {
        rc = solve_fp64(op, b_dev, &o, op->x, op->w, out, nullptr, use ? op->r0 : nullptr);
        if (!rc && use) {  // x = x0 + correction
            hipLaunchKernelGGL(k_axpy1, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, op->n,
                               (const double *)op->x0i, op->x);
            LAUNCH_CHECK(op);
        }
    } else {
        return set_err(op, PLSQR_ERR_ARG, "unknown precision");
    }
    if (rc) return rc;
    if (op->permuted) {
        if ((rc = to_external(op, op->x, x_dev))) return rc;
    } else if (x_dev != op->x) {
        HIP_TRY(op, hipMemcpyAsync(x_dev, op->x, sizeof(double) * op->n, hipMemcpyDeviceToDevice, op->stream));
    }
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

int plsqr_lsqr_solve_x0(plsqr_op *op, const double *b_host, const double *x0_host, const plsqr_solve_opts *opts,
                        double *x_host, plsqr_solve_out *out, int *used_x0) {
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!b_host || !opts || !x_host) return set_err(op, PLSQR_ERR_ARG, "b, opts and x are required");
    for (int i = 0; i < op->m_k; ++i)
        if (!std::isfinite(b_host[i])) return set_err(op, PLSQR_ERR_NONFINITE, "b holds a NaN or Inf at " + std::to_string(i));
    if (x0_host)
        for (int i = 0; i < op->n; ++i)
            if (!std::isfinite(x0_host[i]))
                return set_err(op, PLSQR_ERR_NONFINITE, "x0 holds a NaN or Inf at " + std::to_string(i));
    HIP_TRY(op, hipMemcpyAsync(op->bst, b_host, sizeof(double) * op->m_k, hipMemcpyHostToDevice, op->stream));
    if (x0_host) HIP_TRY(op, hipMemcpyAsync(op->hx, x0_host, sizeof(double) * op->n, hipMemcpyHostToDevice, op->stream));
    int rc = plsqr_lsqr_solve_x0_device(op, op->bst, x0_host ? op->hx : nullptr, opts, op->hx, out, used_x0);
    if (rc) return rc;
    HIP_TRY(op, hipMemcpyAsync(x_host, op->hx, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(op, hipStreamSynchronize(op->stream));
    return PLSQR_OK;
}

int plsqr_residual_norms(plsqr_op *op, const double *b_host, const double *x_host, double *rnorm, double *arnorm,
                         double *anorm_f) {
    using namespace plsqr;
    if (!op) return PLSQR_ERR_ARG;
    if (op->allocs.empty()) return set_err(op, PLSQR_ERR_STATE, "handle failed to initialise");
    if (!b_host || !x_host) return set_err(op, PLSQR_ERR_ARG, "b and x are required");
    HIP_TRY(op, hipMemcpyAsync(op->bst, b_host, sizeof(double) * op->m_k, hipMemcpyHostToDevice, op->stream));
    HIP_TRY(op, hipMemcpyAsync(op->hx, x_host, sizeof(double) * op->n, hipMemcpyHostToDevice, op->stream));
    const double *xin = op->hx;
    if (op->permuted) {
        int rc0 = to_internal(op, op->hx, op->hx2);
        if (rc0) return rc0;
        xin = op->hx2;
    }
    const double *bi = op->bst;
    {
        int rcb = b_internal(op, op->bst, &bi);
        if (rcb) return rcb;
    }
    {
        int rcr = launch_residual(op, xin, bi, op->hy);
        if (rcr) return rcr;
    }
    LAUNCH_CHECK(op);
    int rc;
    double rn2 = 0.0, an2 = 0.0, af2 = 0.0;
    const int gs = std::min(op->part_len, grid_for(op->m, BS, op->cu * 4));
    hipLaunchKernelGGL((k_ssq<BS>), dim3(gs), dim3(BS), 0, op->stream, (long long)op->m, op->hy, op->part);
    LAUNCH_CHECK(op);
    if ((rc = host_sum(op, op->part, gs, &rn2))) return rc;
    const double rn = std::sqrt(rn2);
    // A^T r into hx
    hipLaunchKernelGGL(k_fill, dim3(grid_for(op->n, 256, op->cu * 4)), dim3(256), 0, op->stream, (long long)op->n, 0.0,
                       op->hx);
    LAUNCH_CHECK(op);
    if ((rc = aprod2_dev(op, op->hx, op->hy))) return rc;
    const int gn = std::min(op->part_len, grid_for(op->n, BS, op->cu * 4));
    hipLaunchKernelGGL((k_ssq<BS>), dim3(gn), dim3(BS), 0, op->stream, (long long)op->n, op->hx, op->part);
    LAUNCH_CHECK(op);
    if ((rc = host_sum(op, op->part, gn, &an2))) return rc;
    const double an = std::sqrt(an2);
    const long long nnz = op->nnz_k + op->nnz_d;
    const int ga = std::min(op->part_len, grid_for(nnz, BS, op->cu * 4));
    hipLaunchKernelGGL((k_ssq<BS>), dim3(ga), dim3(BS), 0, op->stream, nnz, op->rv, op->part);
    LAUNCH_CHECK(op);
    if ((rc = host_sum(op, op->part, ga, &af2))) return rc;
    if (rnorm) *rnorm = rn;
    if (arnorm) *arnorm = an;
    if (anorm_f) *anorm_f = std::sqrt(af2);
    return PLSQR_OK;
}

}  // extern "C"
