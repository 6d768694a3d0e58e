// Device kernels of batched solves (included once, by plsqr_batch.hip).
//
// One launch covers every problem of the batch and every right-hand side:
// the operator is block diagonal, vectors hold K right-hand sides interleaved
// (element (i, k) at i*K + k), and each workgroup works on a piece of a single
// problem (row chunk, tile or column chunk; batch_build.h).  Scalar states
// are indexed p*K + k; a piece whose problem has no live state returns at
// once, so converged problems stop costing bandwidth, and a right-hand side
// that has converged keeps its vectors untouched while the others go on.
//
// T = double: FP64 values and vectors, plain sums (the FP64 path).
// T = float : FP32 values and vectors, each dot product accumulated in FP64
//             from exact products (dot32_add); scalars, x and w stay FP64.
#pragma once

#include <type_traits>

#include "plsqr_kernels.h"

namespace plsqr {
namespace batch {

// accumulator of one dot product
template <class T>
struct Acc;

template <>
struct Acc<double> {
    double s;
    __device__ __forceinline__ void zero() { s = 0.0; }
    __device__ __forceinline__ void add(double a, double x) { s += a * x; }
    template <int W>
    __device__ __forceinline__ void group() {
        s = group_sum<W>(s);
    }
    __device__ __forceinline__ void merge(const Acc &o) { s += o.s; }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ double sum64() const { return s; }
};

template <>
struct Acc<float> {
    double s;  // exact FP64 products of the FP32 operands, summed in FP64
    __device__ __forceinline__ void zero() { s = 0.0; }
    __device__ __forceinline__ void add(float a, float x) { dot32_add(s, a, x); }
    template <int W>
    __device__ __forceinline__ void group() {
        s = group_sum<W>(s);
    }
    __device__ __forceinline__ void merge(const Acc &o) { s += o.s; }
    __device__ __forceinline__ float value() const { return (float)s; }
    __device__ __forceinline__ double sum64() const { return s; }
};

// K consecutive values (one index of an interleaved vector)
template <class T, int K>
struct VK {
    T a[K];
};

template <class T, int K>
__device__ __forceinline__ VK<T, K> ldk(const T *__restrict__ p, size_t i) {
    VK<T, K> r;
    if constexpr (K == 4 && std::is_same<T, float>::value) {
        const float4 q = *reinterpret_cast<const float4 *>(p + i * 4);
        r.a[0] = q.x;
        r.a[1] = q.y;
        r.a[2] = q.z;
        r.a[3] = q.w;
    } else if constexpr (K == 4) {
        const double2 q0 = reinterpret_cast<const double2 *>(p + i * 4)[0];
        const double2 q1 = reinterpret_cast<const double2 *>(p + i * 4)[1];
        r.a[0] = q0.x;
        r.a[1] = q0.y;
        r.a[2] = q1.x;
        r.a[3] = q1.y;
    } else if constexpr (K == 2 && std::is_same<T, float>::value) {
        const float2 q = *reinterpret_cast<const float2 *>(p + i * 2);
        r.a[0] = q.x;
        r.a[1] = q.y;
    } else if constexpr (K == 2) {
        const double2 q = *reinterpret_cast<const double2 *>(p + i * 2);
        r.a[0] = q.x;
        r.a[1] = q.y;
    } else {
        r.a[0] = p[i];
    }
    return r;
}

template <class T, int K>
__device__ __forceinline__ void stk(T *__restrict__ p, size_t i, const VK<T, K> &r) {
    if constexpr (K == 4 && std::is_same<T, float>::value) {
        *reinterpret_cast<float4 *>(p + i * 4) = make_float4(r.a[0], r.a[1], r.a[2], r.a[3]);
    } else if constexpr (K == 4) {
        reinterpret_cast<double2 *>(p + i * 4)[0] = make_double2(r.a[0], r.a[1]);
        reinterpret_cast<double2 *>(p + i * 4)[1] = make_double2(r.a[2], r.a[3]);
    } else if constexpr (K == 2 && std::is_same<T, float>::value) {
        *reinterpret_cast<float2 *>(p + i * 2) = make_float2(r.a[0], r.a[1]);
    } else if constexpr (K == 2) {
        *reinterpret_cast<double2 *>(p + i * 2) = make_double2(r.a[0], r.a[1]);
    } else {
        p[i] = r.a[0];
    }
}

// K block sums at once; the results are valid in thread 0.
template <int BS, int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K], double (*sh)[BS / kWave]) {
    const int w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double t = group_sum<kWave>(v[k]);
        if ((threadIdx.x & (kWave - 1)) == 0) sh[k][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double r = sh[k][0];
#pragma unroll
            for (int i = 1; i < BS / kWave; ++i) r += sh[k][i];
            v[k] = r;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Mode 1: y <- fl(fl(y*ib)*(-alpha)) + A x on row chunks.  Kernel-row chunks
// use L lanes per row (64/L rows per wavefront) with U nonzeros per lane in
// flight; damping-row chunks one lane per row.  part[ch*K + k] = the chunk's
// sum of squares of the new y of right-hand side k.
// ---------------------------------------------------------------------------
template <class T, int K, int L, int BS, int U>
__global__ __launch_bounds__(BS) void kb_rows(int nch, const int *__restrict__ ch_row, const int *__restrict__ ch_prob,
                                              const uint8_t *__restrict__ ch_kind, const int *__restrict__ rp,
                                              const int *__restrict__ ci, const T *__restrict__ av,
                                              const T *__restrict__ x, T *__restrict__ y,
                                              const Scal *__restrict__ sc, double *__restrict__ part) {
    __shared__ double sh[K][BS / kWave];
    constexpr int R = BS / L;
    const int lane = threadIdx.x % L;
    const int slot = threadIdx.x / L;
    for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const int p = ch_prob[ch];
        T s1[K], s2[K];
        bool live[K];
        bool any = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const Scal &s = sc[p * K + k];
            live[k] = !s.done;
            s1[k] = (T)s.ib;
            s2[k] = (T)(-s.alpha);
            any = any || live[k];
        }
        if (!any) continue;  // uniform across the block
        const int r0 = ch_row[ch], r1 = ch_row[ch + 1];
        double ssq[K];
#pragma unroll
        for (int k = 0; k < K; ++k) ssq[k] = 0.0;
        if (ch_kind[ch] == 0) {
            for (int rb = r0; rb < r1; rb += R) {
                const int r = rb + slot;
                Acc<T> acc[K];
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k].zero();
                if (r < r1) {
                    const int e = rp[r + 1];
                    int j = rp[r] + lane;
                    for (; j + (U - 1) * L < e; j += U * L) {
                        int c[U];
                        T a[U];
#pragma unroll
                        for (int q = 0; q < U; ++q) {
                            c[q] = ci[j + q * L];
                            a[q] = av[j + q * L];
                        }
                        VK<T, K> xv[U];
#pragma unroll
                        for (int q = 0; q < U; ++q) xv[q] = ldk<T, K>(x, (size_t)c[q]);
#pragma unroll
                        for (int q = 0; q < U; ++q)
#pragma unroll
                            for (int k = 0; k < K; ++k) acc[k].add(a[q], xv[q].a[k]);
                    }
                    for (; j < e; j += L) {
                        const VK<T, K> xv = ldk<T, K>(x, (size_t)ci[j]);
                        const T a = av[j];
#pragma unroll
                        for (int k = 0; k < K; ++k) acc[k].add(a, xv.a[k]);
                    }
                }
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k].template group<L>();
                if (lane == 0 && r < r1) {
                    VK<T, K> yv = ldk<T, K>(y, (size_t)r);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const T yn = (yv.a[k] * s1[k]) * s2[k] + acc[k].value();
                        if (live[k]) {
                            yv.a[k] = yn;
                            ssq[k] += (double)yn * (double)yn;
                        }
                    }
                    stk<T, K>(y, (size_t)r, yv);
                }
            }
        } else {
            for (int r = r0 + threadIdx.x; r < r1; r += BS) {
                Acc<T> acc[K];
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k].zero();
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const VK<T, K> xv = ldk<T, K>(x, (size_t)ci[j]);
                    const T a = av[j];
#pragma unroll
                    for (int k = 0; k < K; ++k) acc[k].add(a, xv.a[k]);
                }
                VK<T, K> yv = ldk<T, K>(y, (size_t)r);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const T yn = (yv.a[k] * s1[k]) * s2[k] + acc[k].value();
                    if (live[k]) {
                        yv.a[k] = yn;
                        ssq[k] += (double)yn * (double)yn;
                    }
                }
                stk<T, K>(y, (size_t)r, yv);
            }
        }
        block_sum_k<BS, K>(ssq, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) part[(size_t)ch * K + k] = ssq[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Mode 1 on (tile, row slice) items, pass 1: workgroup b runs item
// blk_item[b] (-1: none), the same queue as mode 2.  The brick's x values
// (all K) are staged in LDS; a wavefront takes a group of 64 of the item's
// ray segments (one per lane, entries stored together in voxel order),
// copies the group's entries into its LDS buffer CH at a time, and each lane
// sums its segment for the K right-hand sides.  segv[seg*K + k] = the
// partial, stored in T (the FP32 path rounds its FP64 sum once).
// ---------------------------------------------------------------------------
template <class T, int K, int BS, int CH>
__global__ __launch_bounds__(BS) void kb_seg(const int *__restrict__ blk_item, const int *__restrict__ i_prob,
                                             const int *__restrict__ i_c0, const int *__restrict__ i_c1,
                                             const long long *__restrict__ seg_start,
                                             const int *__restrict__ grp_base,
                                             const long long *__restrict__ ent_base,
                                             const uint8_t *__restrict__ seg_len,
                                             const uint8_t *__restrict__ ent_vox, const T *__restrict__ ent_val,
                                             const T *__restrict__ x, const Scal *__restrict__ sc,
                                             T *__restrict__ segv) {
    constexpr int G = 64;
    constexpr int NWV = BS / G;
    static_assert(G == kWave, "one segment group per wavefront");
    __shared__ T xl[128 * K];
    __shared__ __align__(16) T lv[NWV][CH + 8];
    __shared__ __align__(16) uint8_t lx[NWV][CH + 8];
    const int it = blk_item[blockIdx.x];
    if (it < 0) return;
    const int p = i_prob[it];
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; ++k) any = any || !sc[p * K + k].done;
    if (!any) return;  // the same for the whole workgroup
    const int g0 = grp_base[it], g1 = grp_base[it + 1];
    if (g0 == g1) return;
    const int c0 = i_c0[it], tw = i_c1[it] - c0;
    for (int i = threadIdx.x; i < tw * K; i += BS) xl[i] = x[(size_t)c0 * K + i];
    __syncthreads();
    const int lane = threadIdx.x % G;
    const int wv = threadIdx.x / G;
    T *bv = lv[wv];
    uint8_t *bx = lx[wv];
    const long long sb = seg_start[it];
    for (int g = g0 + wv; g < g1; g += NWV) {
        const int len = seg_len[(size_t)g * G + lane];
        int incl = len;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int up = __shfl_up(incl, o);
            if (lane >= o) incl += up;
        }
        const int tot = __shfl(incl, G - 1);
        const int beg = incl - len, end = incl;
        const long long gb = ent_base[g];
        Acc<T> acc[K];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k].zero();
        for (int q0 = 0; q0 < tot; q0 += CH) {
            const int cnt = tot - q0 < CH ? tot - q0 : CH;
            const int sh = stage_entries<T, G, CH>(ent_val, ent_vox, gb + q0, cnt, bv, bx, lane) - q0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int e0 = beg > q0 ? beg : q0;
            const int e1 = end < q0 + cnt ? end : q0 + cnt;
            for (int e = e0; e < e1; ++e) {
                const T a = bv[sh + e];
                const T *xv = xl + (int)bx[sh + e] * K;
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k].add(a, xv[k]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (len > 0) {
            T *o = segv + (size_t)(sb + (long long)(g - g0) * G + lane) * K;
#pragma unroll
            for (int k = 0; k < K; ++k) o[k] = (T)acc[k].sum64();
        }
    }
}

// Pass 2 on row chunks: kernel rows y <- fl(fl(y*ib)*(-alpha)) + (the ray's
// segment partials added in tile order in FP64, LC lanes per ray); damping
// rows as kb_rows.  part[ch*K + k] as kb_rows.
template <class T, int K, int LC, int BS>
__global__ __launch_bounds__(BS) void kb_rows_comb(const int *__restrict__ blk_ch, const int *__restrict__ ch_row,
                                                   const int *__restrict__ ch_prob, const uint8_t *__restrict__ ch_kind,
                                                   const int *__restrict__ ray_ptr, const int *__restrict__ ray_seg,
                                                   const T *__restrict__ segv, const int *__restrict__ rp,
                                                   const int *__restrict__ ci, const T *__restrict__ av,
                                                   const T *__restrict__ x, T *__restrict__ y,
                                                   const Scal *__restrict__ sc, double *__restrict__ part) {
    __shared__ double sh[K][BS / kWave];
    constexpr int R = BS / LC;
    const int lane = threadIdx.x % LC;
    const int slot = threadIdx.x / LC;
    // workgroup b takes row chunk blk_ch[b] (-1: none): a slice's kernel-row
    // chunks run on the XCD whose L2 holds the slice's segment partials
    {
        const int ch = blk_ch[blockIdx.x];
        if (ch < 0) return;
        const int p = ch_prob[ch];
        T s1[K], s2[K];
        bool live[K];
        bool any = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const Scal &s = sc[p * K + k];
            live[k] = !s.done;
            s1[k] = (T)s.ib;
            s2[k] = (T)(-s.alpha);
            any = any || live[k];
        }
        if (!any) return;  // uniform across the block
        const int r0 = ch_row[ch], r1 = ch_row[ch + 1];
        double ssq[K];
#pragma unroll
        for (int k = 0; k < K; ++k) ssq[k] = 0.0;
        if (ch_kind[ch] == 0) {
            for (int rb = r0; rb < r1; rb += R) {
                const int r = rb + slot;
                double acc[K];
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k] = 0.0;
                if (r < r1) {
                    const int e = ray_ptr[r + 1];
                    for (int j = ray_ptr[r] + lane; j < e; j += LC) {
                        const T *sv = segv + (size_t)ray_seg[j] * K;
#pragma unroll
                        for (int k = 0; k < K; ++k) acc[k] += (double)sv[k];
                    }
                }
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k] = group_sum<LC>(acc[k]);
                if (lane == 0 && r < r1) {
                    VK<T, K> yv = ldk<T, K>(y, (size_t)r);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const T yn = (yv.a[k] * s1[k]) * s2[k] + (T)acc[k];
                        if (live[k]) {
                            yv.a[k] = yn;
                            ssq[k] += (double)yn * (double)yn;
                        }
                    }
                    stk<T, K>(y, (size_t)r, yv);
                }
            }
        } else {
            for (int r = r0 + threadIdx.x; r < r1; r += BS) {
                Acc<T> acc[K];
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k].zero();
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const VK<T, K> xv = ldk<T, K>(x, (size_t)ci[j]);
                    const T a = av[j];
#pragma unroll
                    for (int k = 0; k < K; ++k) acc[k].add(a, xv.a[k]);
                }
                VK<T, K> yv = ldk<T, K>(y, (size_t)r);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const T yn = (yv.a[k] * s1[k]) * s2[k] + acc[k].value();
                    if (live[k]) {
                        yv.a[k] = yn;
                        ssq[k] += (double)yn * (double)yn;
                    }
                }
                stk<T, K>(y, (size_t)r, yv);
            }
        }
        block_sum_k<BS, K>(ssq, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) part[(size_t)ch * K + k] = ssq[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Mode 2 on (tile, row slice) items: P[(pbase + c)*K + k] = sum over the
// slice's rows r of a(r, c) * u'[r][k] for the columns c of one brick, u' =
// u*ib (the pending 1/beta).  Workgroup b runs item blk_item[b] (-1: none):
// the rays of the slice crossing the brick are gathered once each (all K
// values) into LDS, CAP at a time, and reused for every voxel they cross; the
// round cuts of the layout give each voxel's slot range per round.
// kb_combine then adds the slices and the damping columns into v.
// ---------------------------------------------------------------------------
template <class T, int K, int TW, int BS, int CAP, int CUT>
__global__ __launch_bounds__(BS) void kb_cols(
    const int *__restrict__ blk_item, const int *__restrict__ i_prob, const int *__restrict__ i_c0,
    const int *__restrict__ i_c1, const long long *__restrict__ i_pbase, const long long *__restrict__ slot_base,
    const long long *__restrict__ seg_start, const uint16_t *__restrict__ vox_len, const int *__restrict__ seg_ray,
    const T *__restrict__ sval, const uint16_t *__restrict__ sseg, const long long *__restrict__ cut_base,
    const uint16_t *__restrict__ vox_cut, const T *__restrict__ u, const Scal *__restrict__ sc,
    double *__restrict__ P) {
    static_assert(BS % TW == 0 && CAP % BS == 0 && CAP % CUT == 0, "tile shape");
    constexpr int NW = BS / TW;
    constexpr int B = (CAP / BS) < 8 ? (CAP / BS) : 8;
    __shared__ T g[CAP * K];
    __shared__ Acc<T> red[NW][K][TW];
    const int it = blk_item[blockIdx.x];
    if (it < 0) return;
    const int p = i_prob[it];
    T ibT[K];
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const Scal &s = sc[p * K + k];
        ibT[k] = (T)s.ib;
        any = any || (!s.done && s.beta_pos);
    }
    if (!any) return;  // uniform across the block, before any barrier
    const int c0 = i_c0[it];
    const int tw = i_c1[it] - c0;
    const int vx = threadIdx.x % TW;
    const int w = threadIdx.x / TW;
    const long long base = slot_base[it];
    const long long s0 = seg_start[it];
    const int ns = (int)(seg_start[it + 1] - s0);
    const int len = vx < tw ? (int)vox_len[(size_t)it * TW + vx] : 0;
    Acc<T> acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k].zero();
    int j = w;
    for (int r0 = 0; r0 < ns; r0 += CAP) {
        const int cnt = ns - r0 < CAP ? ns - r0 : CAP;
        for (int i0 = 0; i0 < cnt; i0 += BS * B) {
            int rr[B];
#pragma unroll
            for (int q = 0; q < B; ++q) {
                const int i = i0 + threadIdx.x + q * BS;
                rr[q] = i < cnt ? seg_ray[s0 + r0 + i] : 0;
            }
            VK<T, K> gv[B];
#pragma unroll
            for (int q = 0; q < B; ++q) {
                const int i = i0 + threadIdx.x + q * BS;
                if (i < cnt) {
                    gv[q] = ldk<T, K>(u, (size_t)rr[q]);
                } else {
#pragma unroll
                    for (int k = 0; k < K; ++k) gv[q].a[k] = T(0);
                }
            }
#pragma unroll
            for (int q = 0; q < B; ++q) {
                const int i = i0 + threadIdx.x + q * BS;
                if (i < cnt) {
#pragma unroll
                    for (int k = 0; k < K; ++k) g[i * K + k] = gv[q].a[k] * ibT[k];
                }
            }
        }
        __syncthreads();
        // this round's slots of voxel vx: [j, jend), segments r0 .. hi-1
        const int hi = r0 + cnt;
        int jend = len;
        if (hi < ns && vx < tw) jend = (int)vox_cut[(size_t)(cut_base[it] + hi / CUT - 1) * TW + vx];
        const T *sv = sval + base + vx;
        const uint16_t *ss = sseg + base + vx;
        for (; j + 3 * NW < jend; j += 4 * NW) {
            const long long o0 = (long long)j * TW, o1 = o0 + NW * TW, o2 = o1 + NW * TW, o3 = o2 + NW * TW;
            const int g0 = ss[o0] - r0, g1 = ss[o1] - r0, g2 = ss[o2] - r0, g3 = ss[o3] - r0;
            const T a0 = sv[o0], a1 = sv[o1], a2 = sv[o2], a3 = sv[o3];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                acc[k].add(a0, g[g0 * K + k]);
                acc[k].add(a1, g[g1 * K + k]);
                acc[k].add(a2, g[g2 * K + k]);
                acc[k].add(a3, g[g3 * K + k]);
            }
        }
        for (; j < jend; j += NW) {
            const long long o = (long long)j * TW;
            const int g0 = ss[o] - r0;
            const T a0 = sv[o];
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k].add(a0, g[g0 * K + k]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) red[w][k][vx] = acc[k];
    __syncthreads();
    if (w == 0 && vx < tw) {
        VK<double, K> out;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            Acc<T> tot = red[0][k][vx];
#pragma unroll
            for (int q = 1; q < NW; ++q) tot.merge(red[q][k][vx]);
            out.a[k] = tot.sum64();
        }
        stk<double, K>(P, (size_t)(i_pbase[it] + c0 + vx), out);
    }
}

// v[c] <- fl(v*vscale) + (sum_s P_s[c] + D^T u'[c]) over column chunks, for
// the right-hand sides with beta > 0 (lsqr.c:530, 577): the slices summed in
// slice order, the damping column in row order (x += (tk + td),
// aprod_impl.c:487-489).  part[ch*K + k] = the chunk's sum of squares of the
// new v.
template <class T, int K, int BS>
__global__ __launch_bounds__(BS) void kb_combine(int ncc, const int *__restrict__ cc_col,
                                                 const int *__restrict__ cc_prob, const long long *__restrict__ pofs,
                                                 const int *__restrict__ nslice, const long long *__restrict__ col0,
                                                 const int *__restrict__ ncol, const double *__restrict__ P,
                                                 const int *__restrict__ dcp, const int *__restrict__ dri,
                                                 const double *__restrict__ dcv, const T *__restrict__ u,
                                                 T *__restrict__ v, const Scal *__restrict__ sc,
                                                 double *__restrict__ part, const double *__restrict__ cs) {
    // cs (may be NULL): column scales d of LSQR on A D; v holds v^ = D v, so
    // v^ = d^2 A^T u - beta v^ and the norm partials are those of v^ / d
    __shared__ double sh[K][BS / kWave];
    for (int ch = blockIdx.x; ch < ncc; ch += gridDim.x) {
        const int p = cc_prob[ch];
        bool live[K];
        double ib[K];
        T vsT[K];
        bool any = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const Scal &s = sc[p * K + k];
            live[k] = !s.done && s.beta_pos;
            ib[k] = s.ib;
            vsT[k] = (T)s.vscale;
            any = any || live[k];
        }
        if (!any) continue;  // uniform across the block
        const int S = nslice[p], n = ncol[p];
        const long long pb = pofs[p], cb = col0[p];
        double ssq[K];
#pragma unroll
        for (int k = 0; k < K; ++k) ssq[k] = 0.0;
        for (int c = cc_col[ch] + threadIdx.x; c < cc_col[ch + 1]; c += BS) {
            const long long cl = c - cb;
            VK<double, K> tk = ldk<double, K>(P, (size_t)(pb + cl));
            for (int q = 1; q < S; ++q) {
                const VK<double, K> t2 = ldk<double, K>(P, (size_t)(pb + (long long)q * n + cl));
#pragma unroll
                for (int k = 0; k < K; ++k) tk.a[k] += t2.a[k];
            }
            double td[K];
#pragma unroll
            for (int k = 0; k < K; ++k) td[k] = 0.0;
            for (int e = dcp[c]; e < dcp[c + 1]; ++e) {
                const VK<T, K> q = ldk<T, K>(u, (size_t)dri[e]);
                const double d = dcv[e];
#pragma unroll
                for (int k = 0; k < K; ++k) td[k] += d * ((double)q.a[k] * ib[k]);
            }
            VK<T, K> vv = ldk<T, K>(v, (size_t)c);
            if (cs) {
                const double dd = cs[c];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (live[k]) {
                        const T vn = vv.a[k] * vsT[k] + (T)(dd * dd * (tk.a[k] + td[k]));
                        vv.a[k] = vn;
                        const double q = (double)vn / dd;
                        ssq[k] += q * q;
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (live[k]) {
                        const T vn = vv.a[k] * vsT[k] + (T)(tk.a[k] + td[k]);
                        vv.a[k] = vn;
                        ssq[k] += (double)vn * (double)vn;
                    }
                }
            }
            stk<T, K>(v, (size_t)c, vv);
        }
        block_sum_k<BS, K>(ssq, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) part[(size_t)ch * K + k] = ssq[k];
        }
    }
}

// ---------------------------------------------------------------------------
// x/w update (lsqr.c:590-597, 685-690) on column chunks; part[ch*K + k] =
// the chunk's sum of (t3*w)^2.
// ---------------------------------------------------------------------------
template <class T, int K, int BS>
__global__ __launch_bounds__(BS) void kb_update(int ncc, const int *__restrict__ cc_col,
                                                const int *__restrict__ cc_prob, const Scal *__restrict__ sc,
                                                T *__restrict__ v, double *__restrict__ w, double *__restrict__ x,
                                                double *__restrict__ part) {
    __shared__ double sh[K][BS / kWave];
    for (int ch = blockIdx.x; ch < ncc; ch += gridDim.x) {
        const int p = cc_prob[ch];
        bool live[K], scl[K];
        T ia[K];
        double t1[K], t2[K], t3[K];
        bool any = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const Scal &s = sc[p * K + k];
            live[k] = !s.done;
            scl[k] = s.scale_v != 0;
            ia[k] = (T)s.inv_alpha;
            t1[k] = s.t1;
            t2[k] = s.t2;
            t3[k] = s.t3;
            any = any || live[k];
        }
        if (!any) continue;
        double dk[K];
#pragma unroll
        for (int k = 0; k < K; ++k) dk[k] = 0.0;
        for (int c = cc_col[ch] + threadIdx.x; c < cc_col[ch + 1]; c += BS) {
            VK<T, K> vv = ldk<T, K>(v, (size_t)c);
            VK<double, K> ww = ldk<double, K>(w, (size_t)c);
            VK<double, K> xx = ldk<double, K>(x, (size_t)c);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (!live[k]) continue;
                T vi = vv.a[k];
                if (scl[k]) {
                    vi = vi * ia[k];
                    vv.a[k] = vi;
                }
                const double t = ww.a[k];
                xx.a[k] = t1[k] * t + xx.a[k];
                ww.a[k] = t2[k] * t + (double)vi;
                const double q = t3[k] * t;
                dk[k] += q * q;
            }
            stk<T, K>(v, (size_t)c, vv);
            stk<double, K>(w, (size_t)c, ww);
            stk<double, K>(x, (size_t)c, xx);
        }
        block_sum_k<BS, K>(dk, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) part[(size_t)ch * K + k] = dk[k];
        }
    }
}

// v <- v/alpha (alpha > 0), w <- v, for live states (lsqr.c:392-395)
template <class T, int K>
__global__ void kb_init_vw(int ncc, const int *__restrict__ cc_col, const int *__restrict__ cc_prob,
                           const Scal *__restrict__ sc, T *__restrict__ v, double *__restrict__ w) {
    for (int ch = blockIdx.x; ch < ncc; ch += gridDim.x) {
        const int p = cc_prob[ch];
        for (int c = cc_col[ch] + threadIdx.x; c < cc_col[ch + 1]; c += blockDim.x) {
            VK<T, K> vv = ldk<T, K>(v, (size_t)c);
            VK<double, K> ww = ldk<double, K>(w, (size_t)c);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const Scal &s = sc[p * K + k];
                if (s.done) continue;
                T vi = vv.a[k];
                if (s.scale_v) {
                    vi = vi * (T)s.inv_alpha;
                    vv.a[k] = vi;
                }
                ww.a[k] = (double)vi;
            }
            stk<T, K>(v, (size_t)c, vv);
            stk<double, K>(w, (size_t)c, ww);
        }
    }
}

// Sums of squares of u over row chunks (the start: ||[b;0]||).
template <class T, int K, int BS>
__global__ __launch_bounds__(BS) void kb_ssq_rows(int nch, const int *__restrict__ ch_row, const T *__restrict__ y,
                                                  double *__restrict__ part) {
    __shared__ double sh[K][BS / kWave];
    for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        double s[K];
#pragma unroll
        for (int k = 0; k < K; ++k) s[k] = 0.0;
        for (int r = ch_row[ch] + threadIdx.x; r < ch_row[ch + 1]; r += BS) {
            const VK<T, K> q = ldk<T, K>(y, (size_t)r);
#pragma unroll
            for (int k = 0; k < K; ++k) s[k] += (double)q.a[k] * (double)q.a[k];
        }
        block_sum_k<BS, K>(s, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) part[(size_t)ch * K + k] = s[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Scalar recurrences, one workgroup per state (p, k): reduce the partials of
// the problem's pieces [rng[p], rng[p+1]) in a fixed order, then thread 0
// runs the recurrence of plsqr_kernels.h.
// ---------------------------------------------------------------------------
enum Phase { PH_INIT_BETA = 0, PH_INIT_ALPHA = 1, PH_BETA = 2, PH_ALPHA = 3, PH_EST = 4 };

template <int K, int PH>
__global__ __launch_bounds__(256) void kb_scalar(const int *__restrict__ rng, const double *__restrict__ part,
                                                 Scal *__restrict__ sc) {
    __shared__ double sh[4];
    const int st = blockIdx.x;
    const int p = st / K, k = st % K;
    Scal *s = sc + st;
    if (s->done) return;  // uniform across the block
    const int b = rng[p], e = rng[p + 1];
    const double sum = reduce_partials<256>(part + (size_t)b * K + k, e - b, K, sh);
    if (threadIdx.x != 0) return;
    if (PH == PH_INIT_BETA) {
        s_init_beta(*s, sum, 0.0);
    } else if (PH == PH_INIT_ALPHA) {
        s_init_alpha(*s, sum, nullptr);
    } else if (PH == PH_BETA) {
        s_beta(*s, sum, 0.0);
    } else if (PH == PH_ALPHA) {
        s_alpha(*s, sum);
    } else {
        s_est(*s, sum, nullptr);
    }
}

// number of live states -> *out
__global__ void kb_count_live(int n_states, const Scal *__restrict__ sc, int *out) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int c = 0;
    for (int i = threadIdx.x; i < n_states; i += blockDim.x) c += sc[i].done ? 0 : 1;
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) *out = cnt;
}

// out[0] = states still live, out[1] = live states whose inner test2
// estimate times ratio[i] (the FP64 / inner ratio measured at the last FP64
// check, <= 0 when unknown) has reached tol: the adaptive check schedule of
// the mixed solve checks as soon as one of them may have converged.
__global__ void kb_count_pred(int n_states, const Scal *__restrict__ sc, const double *__restrict__ ratio, double tol,
                              int *out) {
    __shared__ int cnt[2];
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    int live = 0, pred = 0;
    for (int i = threadIdx.x; i < n_states; i += blockDim.x) {
        if (sc[i].done) continue;
        ++live;
        const double r = ratio[i];
        if (r > 0.0 && sc[i].test2 * r <= tol) ++pred;
    }
    atomicAdd(&cnt[0], live);
    atomicAdd(&cnt[1], pred);
    __syncthreads();
    if (threadIdx.x < 2) out[threadIdx.x] = cnt[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Layout conversion between the caller's arrays (problem after problem, and
// inside a problem right-hand side after right-hand side: b of length m_k, x
// of length n in the original column order, y of length m_k + m_d) and the
// interleaved device vectors.  Kr = the caller's right-hand sides (<= K);
// lanes Kr..K-1 are zero.
// ---------------------------------------------------------------------------
struct ExtMap {
    const int *row_prob, *row_old, *col_prob, *col_old;  // row_old: the caller's row of a device row
    const long long *row0, *col0, *xb, *yb, *bb;
    const int *mk, *mrow, *ncol;
};

// u <- [b ; 0] (from the caller's b in this layout)
template <class T, int K>
__global__ void kb_load_b(long long rows, int Kr, ExtMap m, const double *__restrict__ b, T *__restrict__ u) {
    for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (long long)gridDim.x * blockDim.x) {
        const int p = m.row_prob[r];
        const long long i = m.row_old[r];
        VK<T, K> q;
#pragma unroll
        for (int k = 0; k < K; ++k)
            q.a[k] = (k < Kr && i < m.mk[p]) ? (T)b[m.bb[p] + (long long)k * m.mk[p] + i] : T(0);
        stk<T, K>(u, (size_t)r, q);
    }
}

// device y <- caller's y (dir 0) or caller's y <- device y (dir 1)
template <class T, int K>
__global__ void kb_conv_rows(long long rows, int Kr, ExtMap m, int dir, double *__restrict__ ext, T *__restrict__ y) {
    for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (long long)gridDim.x * blockDim.x) {
        const int p = m.row_prob[r];
        const long long i = m.row_old[r];
        const long long base = m.yb[p] + i;
        if (dir == 0) {
            VK<T, K> q;
#pragma unroll
            for (int k = 0; k < K; ++k) q.a[k] = k < Kr ? (T)ext[base + (long long)k * m.mrow[p]] : T(0);
            stk<T, K>(y, (size_t)r, q);
        } else {
            const VK<T, K> q = ldk<T, K>(y, (size_t)r);
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < Kr) ext[base + (long long)k * m.mrow[p]] = (double)q.a[k];
        }
    }
}

template <class T, int K>
__global__ void kb_conv_cols(long long cols, int Kr, ExtMap m, int dir, double *__restrict__ ext, T *__restrict__ x) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += (long long)gridDim.x * blockDim.x) {
        const int p = m.col_prob[c];
        const long long base = m.xb[p] + m.col_old[c];
        if (dir == 0) {
            VK<T, K> q;
#pragma unroll
            for (int k = 0; k < K; ++k) q.a[k] = k < Kr ? (T)ext[base + (long long)k * m.ncol[p]] : T(0);
            stk<T, K>(x, (size_t)c, q);
        } else {
            const VK<T, K> q = ldk<T, K>(x, (size_t)c);
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < Kr) ext[base + (long long)k * m.ncol[p]] = (double)q.a[k];
        }
    }
}

// r <- [b ; 0] - A x in FP64 (CSR, one lane per row) with the chunk sums of
// squares of r; states of done problems are still computed (cheap, rare).
template <int K, int BS>
__global__ __launch_bounds__(BS) void kb_residual(int nch, const int *__restrict__ ch_row, const int *__restrict__ rp,
                                                  const int *__restrict__ ci, const double *__restrict__ av,
                                                  const double *__restrict__ x, const double *__restrict__ b0,
                                                  double *__restrict__ r, double *__restrict__ part) {
    __shared__ double sh[K][BS / kWave];
    for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        double s[K];
#pragma unroll
        for (int k = 0; k < K; ++k) s[k] = 0.0;
        for (int i = ch_row[ch] + threadIdx.x; i < ch_row[ch + 1]; i += BS) {
            double acc[K];
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] = 0.0;
            for (int j = rp[i]; j < rp[i + 1]; ++j) {
                const VK<double, K> xv = ldk<double, K>(x, (size_t)ci[j]);
                const double a = av[j];
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k] += a * xv.a[k];
            }
            VK<double, K> q = ldk<double, K>(b0, (size_t)i);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                q.a[k] = q.a[k] - acc[k];
                s[k] += q.a[k] * q.a[k];
            }
            stk<double, K>(r, (size_t)i, q);
        }
        block_sum_k<BS, K>(s, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) part[(size_t)ch * K + k] = s[k];
        }
    }
}

// per-state sums of a partial array: out[p*K + k] = sum over [rng[p], rng[p+1])
template <int K>
__global__ __launch_bounds__(256) void kb_state_sums(const int *__restrict__ rng, const double *__restrict__ part,
                                                     double *__restrict__ out) {
    __shared__ double sh[4];
    const int st = blockIdx.x;
    const int p = st / K, k = st % K;
    const int b = rng[p], e = rng[p + 1];
    const double s = reduce_partials<256>(part + (size_t)b * K + k, e - b, K, sh);
    if (threadIdx.x == 0) out[st] = s;
}

// a <- (T) b over len elements
template <class T>
__global__ void kb_convert(long long len, const double *__restrict__ b, T *__restrict__ a) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
        a[i] = (T)b[i];
}

template <class T>
__global__ void kb_fill(long long len, T val, T *__restrict__ a) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
        a[i] = val;
}

// column sums of squares from the first of K interleaved mode-2 results
// (colss[c] = v[c*K]) and column scales cs = colss^(-power/2), 1 for an
// empty column
__global__ void kb_take_first(long long cols, int K, const double *__restrict__ v, double *__restrict__ colss) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += (long long)gridDim.x * blockDim.x)
        colss[c] = v[c * K];
}

__global__ void kb_colscale(long long cols, const double *__restrict__ colss, double power, double *__restrict__ cs) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += (long long)gridDim.x * blockDim.x) {
        const double q = colss[c];
        cs[c] = q > 0.0 ? (power == 1.0 ? 1.0 / sqrt(q) : pow(q, -0.5 * power)) : 1.0;
    }
}

__global__ void kb_square(long long len, const double *__restrict__ a, double *__restrict__ b) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i] * a[i];
}

// xt = x + dx
__global__ void kb_add3(long long len, const double *__restrict__ x, const double *__restrict__ dx,
                        double *__restrict__ xt) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
        xt[i] = x[i] + dx[i];
}

// x += dx
__global__ void kb_axpy(long long len, const double *__restrict__ dx, double *__restrict__ x) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
        x[i] = x[i] + dx[i];
}

}  // namespace batch
}  // namespace plsqr
