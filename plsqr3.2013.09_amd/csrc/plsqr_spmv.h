// SpMV and vector kernels of the LSQR hot path (gfx950); included once, by plsqr_api.hip.
//
// Operator A = [A_k ; D] (m = m_k + m_d rows, n columns).
//   mode 1  y <- y + A x     (aprod_impl.c:22-267)  on a CSR of [A_k ; D]
//   mode 2  x <- x + A^T y   (aprod_impl.c:276-524) on column-major data
//
// Mode 1 (k_rows):   kernel rows use L lanes per row (strided partial sums,
//                    then a fixed lane butterfly); damping rows (1 or 3
//                    nonzeros) use one lane per row, summed in column order.
// Mode 2 (k_cols_sliced + k_cols_combine): the kernel CSC is split into S row
//                    slices (S = 8, one per XCD).  Workgroup b works on slice
//                    b % S, which the hardware's round-robin dispatch places on
//                    XCD b % 8, so each XCD gathers only its own slice of u and
//                    that slice stays resident in the XCD's 4 MB L2.  The S
//                    per-slice column sums are combined in a fixed order with the
//                    damping column sums (one lane per column, rows ascending).
// Exact-order kernels: one lane per row / column, summed sequentially in the
//                    reference's order (CPUSpMV.c:50-89,162-183,203-245,255-296).
#include "plsqr_kernels.h"

namespace plsqr {

// Streamed matrix data is read once per launch: load it non-temporally so it
// does not evict the gathered vector (x in mode 1, the XCD's slice of u in
// mode 2) from L2.
template <bool NT, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    return *p;
}

// ---------------------------------------------------------------------------
// Mode 1
// ---------------------------------------------------------------------------
// y[r] <- fl(fl(y[r]*s1)*s2) + sum_e val[e]*x[col[e]]   (s1 = s2 = 1: plain aprod)
// In the LSQR loop s1 is the pending 1/beta of u (lsqr.c:538) and s2 = -alpha
// (lsqr.c:502): the reference scales u twice and then adds A v, and so do we.
// part[2b], part[2b+1] receive block b's sums of squares of the new y over the
// kernel rows and the damping rows (the two halves of the split norm,
// mycblas.c:99-133).
template <int L, int BS, bool NT = false>
__global__ __launch_bounds__(BS) void k_rows(int m_k, int m, const int *__restrict__ rp,
                                             const int *__restrict__ ci,
                                             const double *__restrict__ av,
                                             const double *__restrict__ x,
                                             double *__restrict__ y, const Scal *__restrict__ sc,
                                             double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    double s1 = 1.0, s2 = 1.0;
    const bool scaled = sc != nullptr;
    if (scaled) {
        if (sc->done) return;
        s1 = sc->ib;
        s2 = -sc->alpha;
    }
    constexpr int R = BS / L;  // rows per block step
    const int lane = threadIdx.x % L;
    const int slot = threadIdx.x / L;
    double ssq_k = 0.0, ssq_d = 0.0;
    const int groups = (m_k + R - 1) / R;
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int r = g * R + slot;
        double acc = 0.0;
        if (r < m_k) {
            const int b = rp[r], e = rp[r + 1];
            int k = b + lane;
            // four independent gathers in flight per lane
            for (; k + 3 * L < e; k += 4 * L) {
                const int c0 = ld_stream<NT>(ci + k), c1 = ld_stream<NT>(ci + k + L),
                          c2 = ld_stream<NT>(ci + k + 2 * L), c3 = ld_stream<NT>(ci + k + 3 * L);
                const double a0 = ld_stream<NT>(av + k), a1 = ld_stream<NT>(av + k + L),
                             a2 = ld_stream<NT>(av + k + 2 * L), a3 = ld_stream<NT>(av + k + 3 * L);
                const double v0 = a0 * x[c0], v1 = a1 * x[c1], v2 = a2 * x[c2], v3 = a3 * x[c3];
                acc += v0;
                acc += v1;
                acc += v2;
                acc += v3;
            }
            for (; k < e; k += L) acc += ld_stream<NT>(av + k) * x[ld_stream<NT>(ci + k)];
        }
        acc = group_sum<L>(acc);
        if (lane == 0 && r < m_k) {
            double yn = y[r];
            if (scaled) yn = (yn * s1) * s2;
            yn = yn + acc;
            y[r] = yn;
            ssq_k += yn * yn;
        }
    }
    // damping rows: one lane per row, columns ascending (CPUSpMV.c:75-84)
    for (long long r = (long long)m_k + (long long)blockIdx.x * BS + threadIdx.x; r < m;
         r += (long long)gridDim.x * BS) {
        double acc = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) acc += av[k] * x[ci[k]];
        double yn = y[r];
        if (scaled) yn = (yn * s1) * s2;
        yn = yn + acc;
        y[r] = yn;
        ssq_d += yn * yn;
    }
    if (part) {
        const double a = block_sum<BS>(ssq_k, sh);
        const double d = block_sum<BS>(ssq_d, sh);
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = a;
            part[2 * blockIdx.x + 1] = d;
        }
    }
}

// Mode 1 on the bricked layout, pass 1: one workgroup per (tile, row slice)
// item stages the tile of x (at most TW values) in LDS.  The item's
// segments (one ray's entries in the tile) are in ray order, each segment's
// entries stored together in ascending voxel order.  A wavefront takes a group of 64
// segments (one per lane), copies the group's entries into its own LDS
// buffer with coalesced loads, CH at a time, and every lane adds the entries
// of its segment found in the buffer (its first entry from a wavefront
// prefix sum of the lane lengths), in entry order.  No padding is stored.
// The partial goes to segv[seg_start[item] + segment index].  Tiles of more
// than 256 voxels (two or four bricks) store 8-bit voxels and, per segment,
// the number of its entries below voxel 256 (512 voxels: one byte of
// seg_lo; entry k is voxel ent_vox + 256 from k = seg_lo on) or below 256,
// 512, 768 (1024 voxels: the bytes of a 32-bit word; entry k adds 256 for
// each count it has reached).  Inside windows of an item the lanes may take
// the segments sorted by length (seg_dst non-null): each partial still goes
// to the segment's ray-order index, seg_dst.
template <int BS, int CH, int TW = 128>
__global__ __launch_bounds__(BS) void k_rows_tiled_c(int S, const int *__restrict__ tile_start,
                                                     const long long *__restrict__ seg_start,
                                                     const int *__restrict__ grp_base,
                                                     const int *__restrict__ ch_item, const int *__restrict__ ch_g0, int chunk,
                                                     const long long *__restrict__ ent_base,
                                                     const uint8_t *__restrict__ seg_len,
                                                     const uint8_t *__restrict__ seg_lo,
                                                     const uint16_t *__restrict__ seg_dst,
                                                     const uint8_t *__restrict__ ent_vox,
                                                     const double *__restrict__ ent_val, const double *__restrict__ x,
                                                     const Scal *__restrict__ sc, double *__restrict__ segv) {
    constexpr int G = 64;
    constexpr int NWV = BS / G;
    static_assert(G == kWave, "one segment group per wavefront");
    static_assert(TW <= 1024, "8-bit voxels and at most three per-segment counts");
    __shared__ double xl[TW];
    __shared__ __align__(16) double lv[NWV][CH + 4];
    __shared__ __align__(16) uint8_t lx[NWV][CH + 4];
    if (sc && sc->done) return;
    // workgroup b runs groups ch_g0[b] .. + chunk - 1 of item ch_item[b]
    // (ch_item NULL: all groups of item b)
    int it = blockIdx.x;
    int g0, g1;
    if (ch_item) {
        it = ch_item[blockIdx.x];
        if (it < 0) return;  // padding of the chunk list
        g0 = ch_g0[blockIdx.x];
        g1 = min(g0 + chunk, grp_base[it + 1]);
    } else {
        g0 = grp_base[it];
        g1 = grp_base[it + 1];
    }
    if (g0 >= g1) return;  // the same for the whole workgroup
    const int gi0 = grp_base[it];  // the item's first group (segment numbering)
    const int t = it / S;
    const int c0 = tile_start[t], tw = tile_start[t + 1] - c0;
    for (int i = threadIdx.x; i < tw; i += BS) xl[i] = x[c0 + i];
    __syncthreads();
    const int lane = threadIdx.x % G;
    const int wv = threadIdx.x / G;
    double *bv = lv[wv];
    uint8_t *bx = lx[wv];
    const long long sb = seg_start[it];
    for (int g = g0 + wv; g < g1; g += NWV) {
        const int len = seg_len[(size_t)g * G + lane];
        // where the partial goes: its ray-order index in the item
        const int dst = seg_dst ? (int)seg_dst[(size_t)g * G + lane] : (g - gi0) * G + lane;
        // staged entry e (lane's own, beg <= e < end): its voxel in the tile
        // is its 8-bit voxel + 256 for each of h1, h2, h3 it has reached
        // (the first staged indices at voxel 256, 512, 768)
        int h1 = 0, h2 = 0, h3 = 0;
        if constexpr (TW > 512) {
            const uint32_t c = reinterpret_cast<const uint32_t *>(seg_lo)[(size_t)g * G + lane];
            h1 = (int)(c & 255u);
            h2 = (int)((c >> 8) & 255u);
            h3 = (int)((c >> 16) & 255u);
        } else if constexpr (TW > 256) {
            h1 = seg_lo[(size_t)g * G + lane];
        }
        int incl = len;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int up = __shfl_up(incl, o);
            if (lane >= o) incl += up;
        }
        const int tot = __shfl(incl, G - 1);  // entries of the group
        const int beg = incl - len, end = incl;
        h1 += beg;
        h2 += beg;
        h3 += beg;
        auto vox = [&](int sh, int e) -> int {
            if constexpr (TW > 512) return bx[sh + e] + (((e >= h1) + (e >= h2) + (e >= h3)) << 8);
            if constexpr (TW > 256) return bx[sh + e] + (e >= h1 ? 256 : 0);
            return bx[sh + e];
        };
        const long long gb = ent_base[g];
        double acc = 0.0;
        for (int q0 = 0; q0 < tot; q0 += CH) {
            const int cnt = tot - q0 < CH ? tot - q0 : CH;
            const int sh = stage_entries<double, G, CH>(ent_val, ent_vox, gb + q0, cnt, bv, bx, lane) - q0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int e0 = beg > q0 ? beg : q0;
            const int e1 = end < q0 + cnt ? end : q0 + cnt;
            // four, then two entries per step: the products formed before the
            // adds (in entry order), so several LDS gathers are in flight
            int e = e0;
            for (; e + 3 < e1; e += 4) {
                const double p0 = bv[sh + e] * xl[vox(sh, e)];
                const double p1 = bv[sh + e + 1] * xl[vox(sh, e + 1)];
                const double p2 = bv[sh + e + 2] * xl[vox(sh, e + 2)];
                const double p3 = bv[sh + e + 3] * xl[vox(sh, e + 3)];
                acc += p0;
                acc += p1;
                acc += p2;
                acc += p3;
            }
            for (; e + 1 < e1; e += 2) {
                const double p0 = bv[sh + e] * xl[vox(sh, e)];
                const double p1 = bv[sh + e + 1] * xl[vox(sh, e + 1)];
                acc += p0;
                acc += p1;
            }
            if (e < e1) acc += bv[sh + e] * xl[vox(sh, e)];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (len > 0) segv[sb + dst] = acc;
    }
}

// Pass 2: y = (y / beta) * (-alpha) + (the ray's segment partials added in
// tile order, LC lanes per ray with strided partial sums), the damping rows
// as in k_rows, and the ||y||^2 partials.  Workgroup b takes the rays of row
// slice b % S (bounds[s] .. bounds[s+1]), the slice whose items pass 1 ran
// on the same XCD (item t*S + s is workgroup t*S + s there), so the partials
// it reads were written through, and stay in, that XCD's L2.  The grid is a
// multiple of S.
template <int LC, int BS>
__global__ __launch_bounds__(BS) void k_rows_tiled_combine(int m_k, int m, int S, const int *__restrict__ bounds,
                                                           const int *__restrict__ ray_ptr,
                                                           const int *__restrict__ ray_seg,
                                                           const double *__restrict__ segv,
                                                           const int *__restrict__ rp, const int *__restrict__ ci,
                                                           const double *__restrict__ av,
                                                           const double *__restrict__ x, double *__restrict__ y,
                                                           const Scal *__restrict__ sc, double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    double s1 = 1.0, s2 = 1.0;
    const bool scaled = sc != nullptr;
    if (scaled) {
        if (sc->done) return;
        s1 = sc->ib;
        s2 = -sc->alpha;
    }
    constexpr int R = BS / LC;
    const int lane = threadIdx.x % LC;
    const int slot = threadIdx.x / LC;
    double ssq_k = 0.0, ssq_d = 0.0;
    const int q = blockIdx.x % S;
    const int r0 = bounds[q], r1 = bounds[q + 1];
    const int groups = (r1 - r0 + R - 1) / R;
    for (int g = blockIdx.x / S; g < groups; g += gridDim.x / S) {
        const int r = r0 + g * R + slot;
        double acc = 0.0;
        if (r < r1) acc = ray_partials<LC>(ray_seg, segv, ray_ptr[r], ray_ptr[r + 1], lane);
        acc = group_sum<LC>(acc);
        if (lane == 0 && r < r1) {
            double yn = y[r];
            if (scaled) yn = (yn * s1) * s2;
            yn = yn + acc;
            y[r] = yn;
            ssq_k += yn * yn;
        }
    }
    for (long long r = (long long)m_k + (long long)blockIdx.x * BS + threadIdx.x; r < m;
         r += (long long)gridDim.x * BS) {
        const double acc = damp_row_sum(rp[r], rp[r + 1], ci, av, x);
        double yn = y[r];
        if (scaled) yn = (yn * s1) * s2;
        yn = yn + acc;
        y[r] = yn;
        ssq_d += yn * yn;
    }
    if (part) {
        const double a = block_sum<BS>(ssq_k, sh);
        const double d = block_sum<BS>(ssq_d, sh);
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = a;
            part[2 * blockIdx.x + 1] = d;
        }
    }
}

// Exact order: one lane per row, sequential sum from 0 in ascending column
// order.  For a kernel row this is the sequence of partial sums the CSC
// scatter cpuSpMV_CSC (CPUSpMV.c:171-180) builds in tmp[r]; for a damping row
// it is cpuSpMV_CSR_v2 (CPUSpMV.c:75-84).
__global__ __launch_bounds__(256) void k_rows_exact(int m, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ av,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ y,
                                                    const Scal *__restrict__ sc) {
    double s1 = 1.0, s2 = 1.0;
    const bool scaled = sc != nullptr;
    if (scaled) {
        if (sc->done) return;
        s1 = sc->ib;
        s2 = -sc->alpha;
    }
    for (long long r = (long long)blockIdx.x * 256 + threadIdx.x; r < m; r += (long long)gridDim.x * 256) {
        double acc = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) acc += av[k] * x[ci[k]];
        double yn = y[r];
        if (scaled) yn = (yn * s1) * s2;
        y[r] = yn + acc;
    }
}

// ---------------------------------------------------------------------------
// Mode 2
// ---------------------------------------------------------------------------
// P[s*n + c] = sum over slice s of column c of val * fl(u[row]*ib)
// The stored u of the reference is fl(u*(1/beta)) (lsqr.c:538); multiplying
// on the fly gives the same operand.
template <int L, int BS>
__global__ __launch_bounds__(BS) void k_cols_sliced(int n, int nsl, const int *__restrict__ cp,
                                                    const int *__restrict__ ri,
                                                    const double *__restrict__ av,
                                                    const double *__restrict__ u,
                                                    const Scal *__restrict__ sc,
                                                    double *__restrict__ P) {
    double ib = 1.0;
    if (sc) {
        if (sc->done || !sc->beta_pos) return;
        ib = sc->ib;
    }
    const int s = blockIdx.x % nsl;
    const int q = blockIdx.x / nsl;
    const int Q = gridDim.x / nsl;
    const int *cps = cp + (size_t)s * (n + 1);
    constexpr int R = BS / L;
    const int lane = threadIdx.x % L;
    const int slot = threadIdx.x / L;
    const int groups = (n + R - 1) / R;
    for (int g = q; g < groups; g += Q) {
        const int c = g * R + slot;
        double acc = 0.0;
        if (c < n) {
            const int b = cps[c], e = cps[c + 1];
            int k = b + lane;
            for (; k + L < e; k += 2 * L) {
                const double v0 = av[k] * (u[ri[k]] * ib);
                const double v1 = av[k + L] * (u[ri[k + L]] * ib);
                acc += v0;
                acc += v1;
            }
            if (k < e) acc += av[k] * (u[ri[k]] * ib);
        }
        acc = group_sum<L>(acc);
        if (lane == 0 && c < n) P[(size_t)s * n + c] = acc;
    }
}

// Tiled mode 2.  Workgroup = one item (tile t, row slice s), item = t*S + s,
// so consecutive workgroups cover the S slices of a tile and workgroup b runs
// on XCD b % 8: with S = 8 each XCD only gathers rows of its own slice of u,
// which stays in its L2.  Phase 1 gathers fl(u[row]*ib) once per segment (a
// ray crossing the tile) into LDS; phase 2 lets thread (w, v) sum the
// nonzeros j = w, w+NW, ... of voxel v, reading the segment values from LDS;
// the NW partial sums of a voxel are then added in a fixed order.
// Segments beyond the LDS capacity are processed in rounds of CAP.
constexpr int CUT = 1024;  // segments per round cut of the tiled layout (TileLayout::CUT)

// Row addressing of the tiled slots (FORM).  0: thread v is voxel v and row
// j of an item starts at j * TW.  1 and 2: thread v is position v (voxels by
// descending length, pos_vox[item * TW + v] = its voxel), so a row's active
// threads are a prefix; with 1 row j starts at row_off[row_base[item] + j]
// and holds only those (no padding in the slot stream), with 2 it starts at
// j * TW.
struct SlotRows {
    const uint8_t *pos_vox;
    const long long *row_base;
    const int *row_off;
};

template <int TW, int BS, int CAP, bool NT, int FORM>
__device__ __forceinline__ void cols_tiled_item(
    int it, int n, int S, const int *__restrict__ tile_start, const long long *__restrict__ slot_base,
    const long long *__restrict__ seg_start, const uint16_t *__restrict__ vox_len,
    const int *__restrict__ seg_ray, const double *__restrict__ sval,
    const uint16_t *__restrict__ sseg, const long long *__restrict__ cut_base,
    const uint16_t *__restrict__ vox_cut, SlotRows sr, const double *__restrict__ u, double ib,
    double *__restrict__ P, double *g, double (*part)[TW]) {
    static_assert(CAP % CUT == 0, "rounds end on round cuts");
    static_assert(TW % kWave == 0, "the lanes of a wavefront share their row phase w");
    constexpr int NW = BS / TW;
    constexpr int U = 4;  // slots per thread in flight (8 measured no faster, 2 slower)
    const int s = it % S;
    const int t = it / S;
    const int c0 = tile_start[t];
    const int tw = tile_start[t + 1] - c0;
    const int v = threadIdx.x % TW;
    const int w = threadIdx.x / TW;
    const int *ro = nullptr;
    if constexpr (FORM == 1) ro = sr.row_off + sr.row_base[it];
    const long long base = slot_base[it];
    const long long s0 = seg_start[it];
    const int ns = (int)(seg_start[it + 1] - s0);
    const int len = v < tw ? (int)vox_len[(size_t)it * TW + v] : 0;
    // Packed rows: the wavefront visits rows w, w + NW, ... up to its longest
    // position (its first lane: positions descend in length); lane l loads
    // the start of its l-th row here, before the segment gathers, so the slot
    // loop takes row starts from registers instead of waiting on a load.
    int rs = 0, lenw = 0;
    if constexpr (FORM == 1) {
        lenw = __builtin_amdgcn_readfirstlane(len);
        const int r = w + NW * (int)(threadIdx.x % kWave);
        if (r < lenw) rs = ro[r];
    }
    double acc = 0.0;
    int j = w;
    if (ns <= CAP) {
        // one round (the common case): every segment is in LDS, so the slot
        // loop needs no bound check and keeps four slot loads in flight; the
        // sum is still taken in slot order
        // all of a thread's segment rows first, then all its gathers: CAP/BS
        // independent loads in flight instead of one dependent pair at a time
        constexpr int PER = CAP / BS;
        int rr[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * BS;
            rr[q] = i < ns ? ld_stream<NT>(seg_ray + s0 + i) : 0;
        }
        double gv[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * BS;
            gv[q] = i < ns ? u[rr[q]] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * BS;
            if (i < ns) g[i] = gv[q] * ib;
        }
        __syncthreads();
        const double *sv = sval + base + v;
        const uint16_t *ss = sseg + base + v;
        if constexpr (FORM == 1) {
            // a loop the whole wavefront runs together (k counts its rows);
            // each lane adds its own rows, a prefix of them, in row order.
            // The slot loads are not predicated, so all of a step's loads
            // are in flight together: a lane past its length reads a slot
            // of a later row (within TW slots of its row's start; the slot
            // arrays carry TW slots of slack at the end) and drops it.
            const int wu = __builtin_amdgcn_readfirstlane(w);
            const int kend = lenw > wu ? (lenw - wu + NW - 1) / NW : 0;
            auto row = [&](int k) -> long long {
                return k < kWave ? (long long)__builtin_amdgcn_readlane(rs, k) : (long long)ro[wu + NW * k];
            };
            int k = 0;
            for (; k + U <= kend; k += U) {
                long long o[U];
#pragma unroll
                for (int q = 0; q < U; ++q) o[q] = row(k + q);
                int gi[U];
                double a[U];
#pragma unroll
                for (int q = 0; q < U; ++q) gi[q] = ld_stream<NT>(ss + o[q]);
#pragma unroll
                for (int q = 0; q < U; ++q) a[q] = ld_stream<NT>(sv + o[q]);
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const bool on = wu + NW * (k + q) < len;
                    const double pr = a[q] * g[on ? gi[q] : 0];
                    acc = on ? acc + pr : acc;
                }
            }
            for (; k < kend; ++k) {
                const long long o = row(k);
                const bool on = wu + NW * k < len;
                const double pr = ld_stream<NT>(sv + o) * g[on ? (int)ld_stream<NT>(ss + o) : 0];
                acc = on ? acc + pr : acc;
            }
        } else {
            for (; j + (U - 1) * NW < len; j += U * NW) {
                const long long o = (long long)j * TW;
                int gi[U];
                double a[U];
#pragma unroll
                for (int q = 0; q < U; ++q) gi[q] = ld_stream<NT>(ss + o + q * NW * TW);
#pragma unroll
                for (int q = 0; q < U; ++q) a[q] = ld_stream<NT>(sv + o + q * NW * TW);
#pragma unroll
                for (int q = 0; q < U; ++q) acc += a[q] * g[gi[q]];
            }
            for (; j < len; j += NW) {
                const long long o = (long long)j * TW;
                acc += ld_stream<NT>(sv + o) * g[ld_stream<NT>(ss + o)];
            }
        }
    } else {
        // rounds of CAP segments; the layout's round cuts give each voxel's
        // slot range per round, so every round is a plain unrolled loop
        for (int r0 = 0; r0 < ns; r0 += CAP) {
            const int cnt = ns - r0 < CAP ? ns - r0 : CAP;
            {
                constexpr int PER = CAP / BS;
                int rr[PER];
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int i = threadIdx.x + q * BS;
                    rr[q] = i < cnt ? ld_stream<NT>(seg_ray + s0 + r0 + i) : 0;
                }
                double gv[PER];
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int i = threadIdx.x + q * BS;
                    gv[q] = i < cnt ? u[rr[q]] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int i = threadIdx.x + q * BS;
                    if (i < cnt) g[i] = gv[q] * ib;
                }
            }
            __syncthreads();
            const int hi = r0 + cnt;
            int jend = len;
            if (hi < ns && v < tw) jend = (int)vox_cut[(size_t)(cut_base[it] + hi / CUT - 1) * TW + v];
            const double *sv = sval + base + v;
            const uint16_t *ss = sseg + base + v;
            for (; j + (U - 1) * NW < jend; j += U * NW) {
                long long o[U];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    if constexpr (FORM == 1) o[q] = ro[j + q * NW];
                    else o[q] = (long long)j * TW + q * NW * TW;
                }
                int gi[U];
                double a[U];
#pragma unroll
                for (int q = 0; q < U; ++q) gi[q] = ld_stream<NT>(ss + o[q]) - r0;
#pragma unroll
                for (int q = 0; q < U; ++q) a[q] = ld_stream<NT>(sv + o[q]);
#pragma unroll
                for (int q = 0; q < U; ++q) acc += a[q] * g[gi[q]];
            }
            for (; j < jend; j += NW) {
                long long o;
                if constexpr (FORM == 1) o = ro[j];
                else o = (long long)j * TW;
                acc += ld_stream<NT>(sv + o) * g[ld_stream<NT>(ss + o) - r0];
            }
            __syncthreads();
        }
    }
    part[w][v] = acc;
    __syncthreads();
    if (w == 0 && v < tw) {
        double tot = part[0][v];
#pragma unroll
        for (int k = 1; k < NW; ++k) tot += part[k][v];
        int vx = v;
        if constexpr (FORM != 0) vx = sr.pos_vox[(size_t)it * TW + v];
        P[(size_t)s * n + c0 + vx] = tot;
    }
    __syncthreads();  // g and part are reused by the next item
}

template <int TW, int BS, int CAP, bool NT, int FORM>
__global__ __launch_bounds__(BS) void k_cols_tiled(
    int n, int S, const int *__restrict__ tile_start, const long long *__restrict__ slot_base,
    const long long *__restrict__ seg_start, const uint16_t *__restrict__ vox_len,
    const int *__restrict__ seg_ray, const double *__restrict__ sval,
    const uint16_t *__restrict__ sseg, const long long *__restrict__ cut_base,
    const uint16_t *__restrict__ vox_cut, SlotRows sr, const double *__restrict__ u, const Scal *__restrict__ sc,
    double *__restrict__ P) {
    __shared__ double g[CAP];
    __shared__ double part[BS / TW][TW];
    double ib = 1.0;
    if (sc) {
        if (sc->done || !sc->beta_pos) return;
        ib = sc->ib;
    }
    cols_tiled_item<TW, BS, CAP, NT, FORM>(blockIdx.x, n, S, tile_start, slot_base, seg_start, vox_len, seg_ray,
                                          sval, sseg, cut_base, vox_cut, sr, u, ib, P, g, part);
}

// dst[i] = src[idx[i]]  (column renumbering at the API boundary)
__global__ void k_gather(int n, const int *__restrict__ idx, const double *__restrict__ src,
                         double *__restrict__ dst) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) dst[i] = src[idx[i]];
}

__global__ void k_gather32(int n, const int *__restrict__ idx, const float *__restrict__ src,
                           float *__restrict__ dst) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) dst[i] = src[idx[i]];
}

// a[i] = map[a[i]]
__global__ void k_remap(long long len, const int *__restrict__ map, int *__restrict__ a) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
        a[i] = map[a[i]];
}

// Interleaved damping column: its len entries at q, q + 64, ..., four at a
// time (their indices and values, then the four u gathers in flight
// together), added in entry order.
template <class U>
__device__ __forceinline__ double damp_col_sum_il(int len, int q, const int *__restrict__ dri,
                                                  const double *__restrict__ dav, const U *__restrict__ ud, double ib) {
    double td = 0.0;
    for (int k0 = 0; k0 < len; k0 += 4, q += 4 * kWave) {
        int i[4];
        double a[4], g[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool on = k0 + j < len;
            i[j] = on ? dri[q + j * kWave] : 0;
            a[j] = on ? dav[q + j * kWave] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) g[j] = k0 + j < len ? (double)ud[i[j]] : 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k0 + j < len) td += a[j] * (g[j] * ib);
    }
    return td;
}

// v[c] <- fl(v[c]*vscale) + (tk + td) with tk the slice partials summed in
// slice order and td the damping column sum in row order
// (cpuSpMVTranspose_CSC_v2, CPUSpMV.c:281-291); x += (tk + td) groups as
// aprod_impl.c:487-489.  part[b] = block b's sum of squares of the new v.
// With column scales cs (LSQR on A D, D = diag(cs)): the solver keeps
// v^ = D v, so v^ <- v^*vscale + cs^2 (tk + td) and part holds ||v^ / cs||^2.
// The damping columns come as a CSC (dlen NULL: column c's entries at
// dcp[c] .. dcp[c+1]-1) or interleaved by groups of 64 columns (dlen[c]
// entries at dcp[c], dcp[c] + 64, ...: lane c % 64 of a wavefront reads
// the address next to lane c % 64 - 1's), entries in the same order either
// way.
__device__ __forceinline__ double damp_col_sum(int c, const int *__restrict__ dcp, const uint8_t *__restrict__ dlen,
                                               const int *__restrict__ dri, const double *__restrict__ dav,
                                               const double *__restrict__ ud, double ib) {
    double td = 0.0;
    if (dlen) {
        td = damp_col_sum_il(dlen[c], dcp[c], dri, dav, ud, ib);
    } else {
        for (int k = dcp[c]; k < dcp[c + 1]; ++k) td += dav[k] * (ud[dri[k]] * ib);
    }
    return td;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_cols_combine(int n, int m_k, int nsl,
                                                     const double *__restrict__ P,
                                                     const int *__restrict__ dcp,
                                                     const int *__restrict__ dri,
                                                     const double *__restrict__ dav,
                                                     const double *__restrict__ u,
                                                     const Scal *__restrict__ sc,
                                                     double *__restrict__ v,
                                                     double *__restrict__ part,
                                                     const double *__restrict__ cs = nullptr,
                                                     const uint8_t *__restrict__ dlen = nullptr) {
    __shared__ double sh[BS / kWave];
    double ib = 1.0, vs = 1.0;
    const bool scaled = sc != nullptr;
    if (scaled) {
        if (sc->done || !sc->beta_pos) return;
        ib = sc->ib;
        vs = sc->vscale;
    }
    const double *ud = u + m_k;
    double ssq = 0.0;
    for (int c = blockIdx.x * BS + threadIdx.x; c < n; c += gridDim.x * BS) {
        double tk = P[c];
        for (int s = 1; s < nsl; ++s) tk += P[(size_t)s * n + c];
        const double td = damp_col_sum(c, dcp, dlen, dri, dav, ud, ib);
        double vn = v[c];
        if (scaled) vn = vn * vs;
        if (cs) {
            const double d = cs[c];
            vn = vn + d * d * (tk + td);
            v[c] = vn;
            const double q = vn / d;
            ssq += q * q;
        } else {
            vn = vn + (tk + td);
            v[c] = vn;
            ssq += vn * vn;
        }
    }
    if (part) {
        const double a = block_sum<BS>(ssq, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = a;
    }
}

// Exact order (cpuSpMVTranspose_CSC and _v2, CPUSpMV.c:230-240,281-291, and
// aprod_impl.c:487-489): one lane per column.
__global__ __launch_bounds__(256) void k_cols_exact(int n, int m_k, const int *__restrict__ cp,
                                                    const int *__restrict__ ri,
                                                    const double *__restrict__ av,
                                                    const int *__restrict__ dcp,
                                                    const int *__restrict__ dri,
                                                    const double *__restrict__ dav,
                                                    const double *__restrict__ u,
                                                    const Scal *__restrict__ sc,
                                                    double *__restrict__ v) {
    double ib = 1.0, vs = 1.0;
    const bool scaled = sc != nullptr;
    if (scaled) {
        if (sc->done || !sc->beta_pos) return;
        ib = sc->ib;
        vs = sc->vscale;
    }
    const double *ud = u + m_k;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
        double tk = 0.0;
        for (int k = cp[c]; k < cp[c + 1]; ++k) tk += av[k] * (u[ri[k]] * ib);
        double td = 0.0;
        for (int k = dcp[c]; k < dcp[c + 1]; ++k) td += dav[k] * (ud[dri[k]] * ib);
        double vn = v[c];
        if (scaled) vn = vn * vs;
        v[c] = vn + (tk + td);
    }
}

// ---------------------------------------------------------------------------
// Norms
// ---------------------------------------------------------------------------
// Block partial sums of squares of y[0:len) (used for ||b|| at start).
template <int BS>
__global__ __launch_bounds__(BS) void k_ssq(long long len, const double *__restrict__ y,
                                            double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < len; i += (long long)gridDim.x * BS)
        s += y[i] * y[i];
    const double a = block_sum<BS>(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// Block partial sums of squares in the pair layout of k_rows: part[2b] over
// y[0:m_k) (kernel rows), part[2b+1] over y[m_k:m) (damping rows).
template <int BS>
__global__ __launch_bounds__(BS) void k_ssq2(int m_k, int m, const double *__restrict__ y,
                                             double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    double a = 0.0, d = 0.0;
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < m; i += (long long)gridDim.x * BS) {
        const double q = y[i] * y[i];
        if (i < m_k)
            a += q;
        else
            d += q;
    }
    a = block_sum<BS>(a, sh);
    d = block_sum<BS>(d, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = d;
    }
}

// u <- u * ib for the 1/beta scale still pending when the loop ends
__global__ void k_scale_pending(long long len, const Scal *__restrict__ sc, double *__restrict__ u) {
    const double ib = sc->ib;
    if (!sc->beta_pos || ib == 1.0) return;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
        u[i] = u[i] * ib;
}

// cblas_dnrm2 (mycblas.c:194-219) in its own sequential order, one lane.
// Writes the norm to *out.  Exact-order mode only.
__global__ void k_dnrm2_seq(int len, const double *__restrict__ v, double *out, const Scal *sc,
                            int check_beta) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (sc && (sc->done || (check_beta && !sc->beta_pos))) return;
    double scale = 0.0, ssq = 1.0;
    if (len <= 0) {
        *out = 0.0;
        return;
    }
    if (len == 1) {
        *out = fabs(v[0]);
        return;
    }
    for (int i = 0; i < len; ++i) {
        const double x = v[i];
        if (x != 0.0) {
            const double ax = fabs(x);
            if (scale < ax) {
                const double q = scale / ax;
                ssq = 1.0 + ssq * q * q;
                scale = ax;
            } else {
                const double q = ax / scale;
                ssq += q * q;
            }
        }
    }
    *out = scale * sqrt(ssq);
}

// dknorm of lsqr.c:685-690 in the reference's order, one lane: reads the
// w of before the update.  Exact-order mode only.
__global__ void k_dknorm_seq(int n, const double *__restrict__ w, Scal *sc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (sc->done) return;
    const double t3 = sc->t3;
    double dk = 0.0;
    for (int i = 0; i < n; ++i) {
        const double t = w[i];
        dk = (t3 * t) * (t3 * t) + dk;
    }
    sc->dk_seq = dk;
}

// ---------------------------------------------------------------------------
// x / w update (lsqr.c:590-597 then 685-690), fused with the v scale
// ---------------------------------------------------------------------------
template <int BS>
__global__ __launch_bounds__(BS) void k_update(int n, const Scal *__restrict__ sc,
                                               double *__restrict__ v, double *__restrict__ w,
                                               double *__restrict__ x, double *__restrict__ se,
                                               double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    if (sc->done) return;
    const bool scl = sc->scale_v != 0;
    const double ia = sc->inv_alpha, t1 = sc->t1, t2 = sc->t2, t3 = sc->t3;
    double dk = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        double vi = v[i];
        if (scl) {
            vi = vi * ia;
            v[i] = vi;
        }
        const double t = w[i];
        x[i] = t1 * t + x[i];
        w[i] = t2 * t + vi;
        const double q = t3 * t;
        const double qq = q * q;
        if (se) se[i] = qq + se[i];
        dk += qq;
    }
    if (part) {
        const double a = block_sum<BS>(dk, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = a;
    }
}

// v <- v * inv_alpha (if alpha > 0), w <- v   (lsqr.c:392-395)
__global__ void k_init_vw(int n, const Scal *__restrict__ sc, double *__restrict__ v,
                          double *__restrict__ w) {
    if (sc->done) return;
    const bool scl = sc->scale_v != 0;
    const double ia = sc->inv_alpha;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double vi = v[i];
        if (scl) {
            vi = vi * ia;
            v[i] = vi;
            w[i] = vi;
        }
    }
}

// final standard errors, lsqr.c:878-891
__global__ void k_se_final(int n, double t, double *__restrict__ se) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) se[i] = t * sqrt(se[i]);
}

// ---------------------------------------------------------------------------
// Scalar recurrences (one block; thread 0 runs the device functions of
// plsqr_kernels.h on the block-reduced sums)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_s_init_beta(Scal *sc, const double *part, int count) {
    __shared__ double sh[4];
    const double sk = reduce_partials<256>(part, count, 2, sh);
    const double sd = reduce_partials<256>(part + 1, count, 2, sh);
    if (threadIdx.x == 0) s_init_beta(*sc, sk, sd);
}

__global__ __launch_bounds__(256) void k_s_init_alpha(Scal *sc, const double *part, int count, double *trace) {
    __shared__ double sh[4];
    const double sv = reduce_partials<256>(part, count, 1, sh);
    if (threadIdx.x == 0) s_init_alpha(*sc, sv, trace);
}

__global__ __launch_bounds__(256) void k_s_beta(Scal *sc, const double *part, int count) {
    __shared__ double sh[4];
    if (sc->done) return;  // uniform across the block
    // the (kernel rows, damping rows) pairs in one pass of 16-byte loads:
    // thread t adds pairs t, t + 256, ... of each in that order, as
    // reduce_partials does for each half, then the same block tree
    double a = 0.0, d = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < count; i += 256) {
        const double2 p = reinterpret_cast<const double2 *>(part)[i];
        a += p.x;
        d += p.y;
    }
    const double sk = block_sum<256>(a, sh);
    const double sd = block_sum<256>(d, sh);
    if (threadIdx.x == 0) s_beta(*sc, sk, sd);
}

__global__ __launch_bounds__(256) void k_s_alpha(Scal *sc, const double *part, int count) {
    __shared__ double sh[4];
    if (sc->done) return;
    const double sv = reduce_partials<256>(part, count, 1, sh);
    if (threadIdx.x == 0) s_alpha(*sc, sv);
}

__global__ __launch_bounds__(256) void k_s_est(Scal *sc, const double *part, int count, double *trace) {
    __shared__ double sh[4];
    if (sc->done) return;
    const double sk = reduce_partials<256>(part, count, 1, sh);
    if (threadIdx.x == 0) s_est(*sc, sk, trace);
}

// ---------------------------------------------------------------------------
// Re-layout helpers
// ---------------------------------------------------------------------------
__global__ void k_expand_ptr(int n, const int *__restrict__ ptr, int *__restrict__ idx) {
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256)
        for (int e = ptr[c]; e < ptr[c + 1]; ++e) idx[e] = c;
}

__global__ void k_iota(long long len, int *__restrict__ a) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
        a[i] = (int)i;
}

__global__ void k_count(long long len, const int *__restrict__ key, int *__restrict__ cnt) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
        atomicAdd(&cnt[key[i]], 1);
}

// dst_idx[p] = src_idx[perm[p]] (+ offset), dst_val[p] = src_val[perm[p]]
__global__ void k_permute(long long len, const int *__restrict__ perm, const int *__restrict__ src_idx,
                          const double *__restrict__ src_val, int *__restrict__ dst_idx,
                          double *__restrict__ dst_val) {
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < len; p += (long long)gridDim.x * 256) {
        const int q = perm[p];
        dst_idx[p] = src_idx[q];
        dst_val[p] = src_val[q];
    }
}

// key[e] = slice(row_idx[e]) * n + col[e]
__global__ void k_slice_key(long long len, int n, int nsl, const int *__restrict__ bounds,
                            const int *__restrict__ row, const int *__restrict__ col,
                            int *__restrict__ key) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < len; e += (long long)gridDim.x * 256) {
        const int r = row[e];
        int s = 0;
        while (s + 1 < nsl && r >= bounds[s + 1]) ++s;
        key[e] = s * n + col[e];
    }
}

// sliced column pointers from the exclusive scan of per-(slice, column) counts
__global__ void k_slice_ptr(int n, int nsl, const int *__restrict__ off, long long total,
                            int *__restrict__ cp) {
    const long long len = (long long)nsl * (n + 1);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256) {
        const int s = (int)(i / (n + 1));
        const int c = (int)(i % (n + 1));
        const long long key = (long long)s * n + c;
        cp[i] = key < (long long)nsl * n ? off[key] : (int)total;
    }
}

// stacked CSR: damping rows appended after the kernel rows
__global__ void k_append_rows(int m_k, int m_d, long long nnz_k, long long nnz_d,
                              const int *__restrict__ drp, const int *__restrict__ dci,
                              const double *__restrict__ dav, int *__restrict__ rp,
                              int *__restrict__ ci, double *__restrict__ av) {
    const long long tid = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = tid; i <= m_d; i += stride) rp[m_k + i] = (int)(nnz_k + drp[i]);
    for (long long e = tid; e < nnz_d; e += stride) {
        ci[nnz_k + e] = dci[e];
        av[nnz_k + e] = dav[e];
    }
}

__global__ void k_fill(long long len, double v, double *__restrict__ a) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256) a[i] = v;
}

__global__ void k_copy(long long len, const double *__restrict__ src, double *__restrict__ dst) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

// r = [b;0] - A x on the stacked CSR (rows), FP64
// r = [b;0] - A x with the kernel rows' A x from the per-brick segment
// partials of k_rows_tiled_c (added in tile order, LC lanes per ray) and the
// damping rows from the CSR; rows of slice b % S per workgroup as in
// k_rows_tiled_combine.
template <int LC, int BS>
__global__ __launch_bounds__(BS) void k_rows_tiled_residual(int m_k, int m, int S, const int *__restrict__ bounds,
                                                            const int *__restrict__ ray_ptr,
                                                            const int *__restrict__ ray_seg,
                                                            const double *__restrict__ segv,
                                                            const int *__restrict__ rp, const int *__restrict__ ci,
                                                            const double *__restrict__ av,
                                                            const double *__restrict__ x,
                                                            const double *__restrict__ b, double *__restrict__ r) {
    constexpr int R = BS / LC;
    const int lane = threadIdx.x % LC;
    const int slot = threadIdx.x / LC;
    const int q = blockIdx.x % S;
    const int r0 = bounds[q], r1 = bounds[q + 1];
    const int groups = (r1 - r0 + R - 1) / R;
    for (int g = blockIdx.x / S; g < groups; g += gridDim.x / S) {
        const int i = r0 + g * R + slot;
        double acc = 0.0;
        if (i < r1) acc = ray_partials<LC>(ray_seg, segv, ray_ptr[i], ray_ptr[i + 1], lane);
        acc = group_sum<LC>(acc);
        if (lane == 0 && i < r1) r[i] = b[i] - acc;
    }
    for (long long i = (long long)m_k + (long long)blockIdx.x * BS + threadIdx.x; i < m;
         i += (long long)gridDim.x * BS) {
        double acc = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) acc += av[k] * x[ci[k]];
        r[i] = 0.0 - acc;
    }
}

__global__ void k_residual(int m_k, int m, const int *__restrict__ rp, const int *__restrict__ ci,
                           const double *__restrict__ av, const double *__restrict__ x,
                           const double *__restrict__ b, double *__restrict__ r) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long long)gridDim.x * 256) {
        double acc = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) acc += av[k] * x[ci[k]];
        r[i] = (i < m_k ? b[i] : 0.0) - acc;
    }
}

}  // namespace plsqr
