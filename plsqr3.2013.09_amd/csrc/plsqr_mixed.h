// Mixed-precision kernels (included once, by plsqr_api.hip).
//
// Matrix values and the Golub-Kahan vectors u, v are stored in FP32; every
// row / column dot product accumulates the exact FP64 products of its FP32
// operands in FP64 (dot32_add), so each dot product is as accurate as if
// computed in twice the working precision and is rounded to FP32 once.
// alpha, beta, the rotations and the estimates stay FP64 (the Scal kernels),
// and so do x, w and the outer residual r = [b;0] - A x.
#pragma once

#include "plsqr_kernels.h"

namespace plsqr {

// Mode 1 in FP32: y <- fl32(fl32(y*s1)*s2) + fl32(dot(row, x)); FP64 partial sums
// of squares of the new y, pair layout (kernel rows, damping rows).
template <int L, int BS>
__global__ __launch_bounds__(BS) void k_rows32(int m_k, int m, const int *__restrict__ rp,
                                               const int *__restrict__ ci, const float *__restrict__ av,
                                               const float *__restrict__ x, float *__restrict__ y,
                                               const Scal *__restrict__ sc, double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    if (sc->done) return;
    const float s1 = (float)sc->ib, s2 = (float)(-sc->alpha);
    constexpr int R = BS / L;
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
            for (; k + 3 * L < e; k += 4 * L) {
                const int c0 = ci[k], c1 = ci[k + L], c2 = ci[k + 2 * L], c3 = ci[k + 3 * L];
                const float a0 = av[k], a1 = av[k + L], a2 = av[k + 2 * L], a3 = av[k + 3 * L];
                const float x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
                dot32_add(acc, a0, x0);
                dot32_add(acc, a1, x1);
                dot32_add(acc, a2, x2);
                dot32_add(acc, a3, x3);
            }
            for (; k < e; k += L) dot32_add(acc, av[k], x[ci[k]]);
        }
        acc = group_sum<L>(acc);
        if (lane == 0 && r < m_k) {
            const float yn = (y[r] * s1) * s2 + (float)acc;
            y[r] = yn;
            ssq_k += (double)yn * (double)yn;
        }
    }
    for (long long r = (long long)m_k + (long long)blockIdx.x * BS + threadIdx.x; r < m;
         r += (long long)gridDim.x * BS) {
        double acc = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) dot32_add(acc, av[k], x[ci[k]]);
        const float yn = (y[r] * s1) * s2 + (float)acc;
        y[r] = yn;
        ssq_d += (double)yn * (double)yn;
    }
    const double a = block_sum<BS>(ssq_k, sh);
    const double d = block_sum<BS>(ssq_d, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = d;
    }
}

// Mode 1 in FP32 on the bricked layout, pass 1 (the structure of
// k_rows_tiled_c): FP32 entries and x tile, FP64 sum per segment, the
// partial s + c rounded once to FP32.  Tiles of 512 and 1024 voxels as in
// k_rows_tiled_c (seg_lo).
template <int BS, int CH, int TW = 128>
__global__ __launch_bounds__(BS) void k_rows_tiled_c32(int S, const int *__restrict__ tile_start,
                                                       const long long *__restrict__ seg_start,
                                                       const int *__restrict__ grp_base,
                                                     const int *__restrict__ ch_item, const int *__restrict__ ch_g0, int chunk,
                                                       const long long *__restrict__ ent_base,
                                                       const uint8_t *__restrict__ seg_len,
                                                       const uint8_t *__restrict__ seg_lo,
                                                       const uint16_t *__restrict__ seg_dst,
                                                       const uint8_t *__restrict__ ent_vox,
                                                       const float *__restrict__ ent_val, const float *__restrict__ x,
                                                       const Scal *__restrict__ sc, float *__restrict__ segv) {
    constexpr int G = 64;
    constexpr int NWV = BS / G;
    static_assert(G == kWave, "one segment group per wavefront");
    static_assert(TW <= 1024, "8-bit voxels and at most three per-segment counts");
    __shared__ double xl[TW];  // the tile of x, widened once per item
    __shared__ __align__(16) float lv[NWV][CH + 8];
    __shared__ __align__(16) uint8_t lx[NWV][CH + 8];
    if (sc->done) return;
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
    for (int i = threadIdx.x; i < tw; i += BS) xl[i] = (double)x[c0 + i];
    __syncthreads();
    const int lane = threadIdx.x % G;
    const int wv = threadIdx.x / G;
    float *bv = lv[wv];
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
        const int tot = __shfl(incl, G - 1);
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
            const int sh = stage_entries<float, G, CH>(ent_val, ent_vox, gb + q0, cnt, bv, bx, lane) - q0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int e0 = beg > q0 ? beg : q0;
            const int e1 = end < q0 + cnt ? end : q0 + cnt;
            // four, then two entries per step: the operands of a step read
            // first (several LDS gathers in flight), then the exact products
            // added in entry order
            int e = e0;
            for (; e + 3 < e1; e += 4) {
                const float a0 = bv[sh + e], a1 = bv[sh + e + 1], a2 = bv[sh + e + 2], a3 = bv[sh + e + 3];
                const double x0 = xl[vox(sh, e)], x1 = xl[vox(sh, e + 1)], x2 = xl[vox(sh, e + 2)],
                             x3 = xl[vox(sh, e + 3)];
                dot32_add(acc, a0, x0);
                dot32_add(acc, a1, x1);
                dot32_add(acc, a2, x2);
                dot32_add(acc, a3, x3);
            }
            for (; e + 1 < e1; e += 2) {
                const float a0 = bv[sh + e], a1 = bv[sh + e + 1];
                const double x0 = xl[vox(sh, e)], x1 = xl[vox(sh, e + 1)];
                dot32_add(acc, a0, x0);
                dot32_add(acc, a1, x1);
            }
            if (e < e1) dot32_add(acc, bv[sh + e], xl[vox(sh, e)]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (len > 0) segv[sb + dst] = (float)acc;
    }
}

// Pass 2 in FP32: y = fl32(fl32(y*s1)*s2) + fl32(the ray's FP32 partials
// added in tile order in FP64); damping rows and the FP64 ||y||^2 partials as in
// k_rows32.  Workgroup b takes the rays of row slice b % S, as in
// k_rows_tiled_combine.
template <int LC, int BS>
__global__ __launch_bounds__(BS) void k_rows_tiled_combine32(int m_k, int m, int S, const int *__restrict__ bounds,
                                                             const int *__restrict__ ray_ptr,
                                                             const int *__restrict__ ray_seg,
                                                             const float *__restrict__ segv,
                                                             const int *__restrict__ rp, const int *__restrict__ ci,
                                                             const float *__restrict__ av,
                                                             const float *__restrict__ x, float *__restrict__ y,
                                                             const Scal *__restrict__ sc, double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    if (sc->done) return;
    const float s1 = (float)sc->ib, s2 = (float)(-sc->alpha);
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
            const float yn = (y[r] * s1) * s2 + (float)acc;
            y[r] = yn;
            ssq_k += (double)yn * (double)yn;
        }
    }
    for (long long r = (long long)m_k + (long long)blockIdx.x * BS + threadIdx.x; r < m;
         r += (long long)gridDim.x * BS) {
        const double acc = damp_row_sum(rp[r], rp[r + 1], ci, av, x);
        const float yn = (y[r] * s1) * s2 + (float)acc;
        y[r] = yn;
        ssq_d += (double)yn * (double)yn;
    }
    const double a = block_sum<BS>(ssq_k, sh);
    const double d = block_sum<BS>(ssq_d, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = d;
    }
}

// Mode 2 in FP32 on the tiled layout (same item structure and slot forms as
// k_cols_tiled).
template <int TW, int BS, int CAP, int FORM>
__global__ __launch_bounds__(BS, 8) void k_cols_tiled32(
    int n, int S, const int *__restrict__ tile_start, const long long *__restrict__ slot_base,
    const long long *__restrict__ seg_start, const uint16_t *__restrict__ vox_len,
    const int *__restrict__ seg_ray, const float *__restrict__ sval, const uint16_t *__restrict__ sseg,
    const long long *__restrict__ cut_base, const uint16_t *__restrict__ vox_cut, SlotRows sr,
    const float *__restrict__ u, const Scal *__restrict__ sc, double *__restrict__ P) {
    static_assert(TW % kWave == 0, "the lanes of a wavefront share their row phase w");
    constexpr int NW = BS / TW;
    __shared__ float g[CAP];  // fl32(u * ib) per segment
    __shared__ double part[NW][TW];
    if (sc->done || !sc->beta_pos) return;
    const float ib = (float)sc->ib;
    const int it = blockIdx.x;
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
    // packed rows: row starts of the wavefront's rows in its lanes, as in
    // k_cols_tiled
    int rs = 0, lenw = 0;
    if constexpr (FORM == 1) {
        lenw = __builtin_amdgcn_readfirstlane(len);
        const int r = w + NW * (int)(threadIdx.x % kWave);
        if (r < lenw) rs = ro[r];
    }
    double acc = 0.0;
    int j = w;
    for (int r0 = 0; r0 < ns; r0 += CAP) {
        const int cnt = ns - r0 < CAP ? ns - r0 : CAP;
        // the round's segments, at most 8 per thread in flight at a time
        constexpr int PER = CAP / BS;
        constexpr int PB = PER < 8 ? PER : 8;
        static_assert(PER % PB == 0, "whole batches");
#pragma unroll
        for (int b0 = 0; b0 < PER; b0 += PB) {
            if (b0 * BS < cnt) {
                int rr[PB];
#pragma unroll
                for (int q = 0; q < PB; ++q) {
                    const int i = threadIdx.x + (b0 + q) * BS;
                    rr[q] = i < cnt ? seg_ray[s0 + r0 + i] : 0;
                }
                float gv[PB];
#pragma unroll
                for (int q = 0; q < PB; ++q) {
                    const int i = threadIdx.x + (b0 + q) * BS;
                    gv[q] = i < cnt ? u[rr[q]] : 0.f;
                }
#pragma unroll
                for (int q = 0; q < PB; ++q) {
                    const int i = threadIdx.x + (b0 + q) * BS;
                    if (i < cnt) g[i] = gv[q] * ib;
                }
            }
        }
        __syncthreads();
        const int hi = r0 + cnt;
        if (ns <= CAP) {
            const float *sv = sval + base + v;
            const uint16_t *ss = sseg + base + v;
            if constexpr (FORM == 1) {
                const int wu = __builtin_amdgcn_readfirstlane(w);
                const int kend = lenw > wu ? (lenw - wu + NW - 1) / NW : 0;
                auto row = [&](int k) -> long long {
                    return k < kWave ? (long long)__builtin_amdgcn_readlane(rs, k) : (long long)ro[wu + NW * k];
                };
                // loads not predicated (lanes past their length read slack
                // slots and drop them), as in k_cols_tiled
                int k = 0;
                for (; k + 4 <= kend; k += 4) {
                    const long long o0 = row(k), o1 = row(k + 1), o2 = row(k + 2), o3 = row(k + 3);
                    const bool n0 = wu + NW * k < len, n1 = wu + NW * (k + 1) < len, n2 = wu + NW * (k + 2) < len,
                               n3 = wu + NW * (k + 3) < len;
                    const int g0 = ss[o0], g1 = ss[o1], g2 = ss[o2], g3 = ss[o3];
                    const float a0 = sv[o0], a1 = sv[o1], a2 = sv[o2], a3 = sv[o3];
                    const double s0 = fma((double)a0, g[n0 ? g0 : 0], acc);
                    acc = n0 ? s0 : acc;
                    const double s1 = fma((double)a1, g[n1 ? g1 : 0], acc);
                    acc = n1 ? s1 : acc;
                    const double s2 = fma((double)a2, g[n2 ? g2 : 0], acc);
                    acc = n2 ? s2 : acc;
                    const double s3 = fma((double)a3, g[n3 ? g3 : 0], acc);
                    acc = n3 ? s3 : acc;
                }
                for (; k < kend; ++k) {
                    const long long o = row(k);
                    const bool on = wu + NW * k < len;
                    const double s0 = fma((double)sv[o], g[on ? (int)ss[o] : 0], acc);
                    acc = on ? s0 : acc;
                }
            } else {
                for (; j + 3 * NW < len; j += 4 * NW) {
                    const long long o0 = (long long)j * TW, o1 = o0 + NW * TW, o2 = o1 + NW * TW, o3 = o2 + NW * TW;
                    const int g0 = ss[o0], g1 = ss[o1], g2 = ss[o2], g3 = ss[o3];
                    const float a0 = sv[o0], a1 = sv[o1], a2 = sv[o2], a3 = sv[o3];
                    dot32_add(acc, a0, g[g0]);
                    dot32_add(acc, a1, g[g1]);
                    dot32_add(acc, a2, g[g2]);
                    dot32_add(acc, a3, g[g3]);
                }
                for (; j < len; j += NW) {
                    const long long o = (long long)j * TW;
                    dot32_add(acc, sv[o], g[ss[o]]);
                }
            }
        } else {
            int jend = len;
            if (hi < ns && v < tw) jend = (int)vox_cut[(size_t)(cut_base[it] + hi / CUT - 1) * TW + v];
            const float *sv = sval + base + v;
            const uint16_t *ss = sseg + base + v;
            auto roff = [&](int jj) -> long long {
                if constexpr (FORM == 1) return ro[jj];
                return (long long)jj * TW;
            };
            for (; j + 3 * NW < jend; j += 4 * NW) {
                long long o0, o1, o2, o3;
                if constexpr (FORM == 1) {
                    o0 = ro[j], o1 = ro[j + NW], o2 = ro[j + 2 * NW], o3 = ro[j + 3 * NW];
                } else {
                    o0 = (long long)j * TW, o1 = o0 + NW * TW, o2 = o1 + NW * TW, o3 = o2 + NW * TW;
                }
                const int g0 = ss[o0] - r0, g1 = ss[o1] - r0, g2 = ss[o2] - r0, g3 = ss[o3] - r0;
                const float a0 = sv[o0], a1 = sv[o1], a2 = sv[o2], a3 = sv[o3];
                dot32_add(acc, a0, g[g0]);
                dot32_add(acc, a1, g[g1]);
                dot32_add(acc, a2, g[g2]);
                dot32_add(acc, a3, g[g3]);
            }
            for (; j < jend; j += NW) {
                const long long o = roff(j);
                dot32_add(acc, sv[o], g[ss[o] - r0]);
            }
        }
        __syncthreads();
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
}

// v32[c] <- fl32(v32*vscale) + (tk + td): tk the FP64 slice partials, td the
// damping column (FP64 values, FP32 u); FP64 partial sums of squares.
template <int BS>
//
// cs (may be NULL): the column scales d of LSQR on A D.  The solver then
// keeps v^ = D v (and w^ = D w, so x += t1 w^ directly): mode 1 of A D on v
// is mode 1 of A on v^, and v^ = D^2 A^T u - beta v^; the norm partials are
// those of v = v^ / d.
__global__ __launch_bounds__(BS) void k_cols_combine32(int n, int m_k, int nsl, const double *__restrict__ P,
                                                       const int *__restrict__ dcp, const int *__restrict__ dri,
                                                       const double *__restrict__ dav, const float *__restrict__ u,
                                                       const Scal *__restrict__ sc, float *__restrict__ v,
                                                       double *__restrict__ part,
                                                       const double *__restrict__ cs = nullptr,
                                                       const uint8_t *__restrict__ dlen = nullptr) {
    __shared__ double sh[BS / kWave];
    if (sc->done || !sc->beta_pos) return;
    const double ib = sc->ib;
    const float vs = (float)sc->vscale;
    const float *ud = u + m_k;
    double ssq = 0.0;
    for (int c = blockIdx.x * BS + threadIdx.x; c < n; c += gridDim.x * BS) {
        double tk = P[c];
        for (int s = 1; s < nsl; ++s) tk += P[(size_t)s * n + c];
        double td = 0.0;
        if (dlen) {  // damping columns interleaved by groups of 64 (k_cols_combine)
            td = damp_col_sum_il(dlen[c], dcp[c], dri, dav, ud, ib);
        } else {
            for (int k = dcp[c]; k < dcp[c + 1]; ++k) td += dav[k] * ((double)ud[dri[k]] * ib);
        }
        if (cs) {
            const double d = cs[c];
            const float vn = v[c] * vs + (float)(d * d * (tk + td));
            v[c] = vn;
            const double q = (double)vn / d;
            ssq += q * q;
        } else {
            const float vn = v[c] * vs + (float)(tk + td);
            v[c] = vn;
            ssq += (double)vn * (double)vn;
        }
    }
    const double a = block_sum<BS>(ssq, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// x/w update with FP32 v and FP64 x, w (lsqr.c:590-597, 685-690)
template <int BS>
__global__ __launch_bounds__(BS) void k_update32(int n, const Scal *__restrict__ sc, float *__restrict__ v,
                                                 double *__restrict__ w, double *__restrict__ x,
                                                 double *__restrict__ part) {
    __shared__ double sh[BS / kWave];
    if (sc->done) return;
    const bool scl = sc->scale_v != 0;
    const float ia = (float)sc->inv_alpha;
    const double t1 = sc->t1, t2 = sc->t2, t3 = sc->t3;
    double dk = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        float vi = v[i];
        if (scl) {
            vi = vi * ia;
            v[i] = vi;
        }
        const double t = w[i];
        x[i] = t1 * t + x[i];
        w[i] = t2 * t + (double)vi;
        const double q = t3 * t;
        dk += q * q;
    }
    const double a = block_sum<BS>(dk, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void k_init_vw32(int n, const Scal *__restrict__ sc, float *__restrict__ v, double *__restrict__ w) {
    if (sc->done) return;
    const bool scl = sc->scale_v != 0;
    const float ia = (float)sc->inv_alpha;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float vi = v[i];
        if (scl) {
            vi = vi * ia;
            v[i] = vi;
        }
        w[i] = (double)vi;
    }
}

// d = (sum of squares of the column)^(-power/2), 1 for an empty column
__global__ void k_colscale(int n, const double *__restrict__ colss, double power, double *__restrict__ d) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double q = colss[i];
        d[i] = q > 0.0 ? (power == 1.0 ? 1.0 / sqrt(q) : pow(q, -0.5 * power)) : 1.0;
    }
}

__global__ void k_square(long long len, const double *__restrict__ a, double *__restrict__ b) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
        b[i] = a[i] * a[i];
}

__global__ void k_to_f32(long long len, const double *__restrict__ a, float *__restrict__ b) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
        b[i] = (float)a[i];
}

__global__ void k_fill32(long long len, float *__restrict__ a) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long long)gridDim.x * 256) a[i] = 0.f;
}

// xt = x + dx
__global__ void k_add3(int n, const double *__restrict__ x, const double *__restrict__ dx, double *__restrict__ xt) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) xt[i] = x[i] + dx[i];
}

// x += dx
__global__ void k_axpy1(int n, const double *__restrict__ dx, double *__restrict__ x) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) x[i] = x[i] + dx[i];
}

}  // namespace plsqr
