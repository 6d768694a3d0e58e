// Device kernels of the LSQR hot path for gfx950 (64-wide wavefronts).
//
// All floating-point code in this library is compiled with -ffp-contract=off:
// every a*b+c below is a rounded product followed by a rounded sum, the
// arithmetic of the reference's x86-64 build (no FMA), so that the exact-order
// kernels reproduce the reference bit for bit and the wavefront-parallel
// kernels differ from it only by summation order.
//
// Reductions never use atomics on floating-point data: every sum is formed in
// a fixed order (lane butterflies, fixed block trees, fixed partial slots), so
// results are bitwise reproducible from run to run.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plsqr {

constexpr int kWave = 64;

// Device-resident LSQR state.  One thread of a one-block kernel owns the
// scalar recurrences of lsqr.c:425-775; the vector kernels read the
// coefficients they need from here, so no scalar crosses PCIe inside the loop.
struct Scal {
    // parameters
    double damp, atol, btol, ctol, tol;
    int itnlim, itn_max, exact, wantse;
    // bidiagonalisation
    double alpha, beta;
    double ib;          // pending 1/beta scale of u (1 when beta == 0)
    double vscale;      // factor applied to v before v += A^T u  (1 or -beta)
    double inv_alpha;   // 1/alpha for the v scale (lsqr.c:590-597)
    int beta_pos;       // beta > 0 in this iteration (lsqr.c:530)
    int scale_v;        // v must be scaled by inv_alpha in the update kernel
    // norms from the sequential (exact-order) norm kernels
    double pk, pd, pv, dk_seq;
    // LSQR recurrences
    double anorm, acond, dnorm, dxmax, res2, psi, xnorm, xnorm1, cs2, sn2, z;
    double rnorm, bnorm, rhobar, phibar, arnorm, test1, test2, test3;
    double phi, rho, theta, t1, t2, t3;
    int itn, istop, nstop, maxdx, done, init_zero;
};

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ double group_sum(double v) {
    // butterfly inside aligned groups of W lanes; every lane ends with the sum
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// One lane's share of a ray's segment partials: segments b + lane,
// b + lane + LC, ... below e, added in that order, two gathers in flight.
// (Groups of 8 index loads ahead of the partial loads measured slower:
// FP32 mode 1 0.413 -> 0.430 ms at 67^3/2M rays.)
template <int LC, class T>
__device__ __forceinline__ double ray_partials(const int *__restrict__ ray_seg, const T *__restrict__ segv, int b,
                                               int e, int lane) {
    double acc = 0.0;
    int k = b + lane;
    for (; k + LC < e; k += 2 * LC) {
        const int i0 = ray_seg[k], i1 = ray_seg[k + LC];
        const double p0 = (double)segv[i0], p1 = (double)segv[i1];
        acc += p0;
        acc += p1;
    }
    if (k < e) acc += (double)segv[ray_seg[k]];
    return acc;
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = group_sum<kWave>(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = sh[0];
#pragma unroll
        for (int i = 1; i < BS / kWave; ++i) r += sh[i];
    }
    __syncthreads();
    return r;  // valid in thread 0
}

// Sum of `count` partials in a fixed order by one block: thread t takes the
// strided slots t, t+BS, ...; then the fixed block tree.
template <int BS>
__device__ double reduce_partials(const double *p, int count, int stride, double *sh) {
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += BS) s += p[(size_t)i * stride];
    return block_sum<BS>(s, sh);
}

// FP32 dot products with FP64 accumulation: the product of two FP32 values
// is exact in FP64 (24 + 24 significant bits <= 53), so acc += (double)a *
// (double)x rounds once per term, in FP64.  Each FP32 dot product is thus as
// accurate as if it were computed in twice the working precision.  Since the
// product is exact, the fused multiply-add rounds exactly as the separate
// multiply and add would: one FP64 operation per term, the same bits.
__device__ __forceinline__ void dot32_add(double &acc, float a, float x) { acc = fma((double)a, (double)x, acc); }
// the same with x already widened (an FP32 value held in FP64)
__device__ __forceinline__ void dot32_add(double &acc, float a, double x) { acc = fma((double)a, x, acc); }

// A damping row's entries b .. e-1 (1 to 3 for the grid stencils), four at a
// time: the column indices and values, then the x gathers in flight
// together, added in entry order as acc += a * x (FP64 values) or as
// dot32_add (FP32 values, the exact product).
template <class A, class X>
__device__ __forceinline__ double damp_row_sum(int b, int e, const int *__restrict__ ci, const A *__restrict__ av,
                                               const X *__restrict__ x) {
    double acc = 0.0;
    for (int k0 = b; k0 < e; k0 += 4) {
        int i[4];
        A a[4];
        X g[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool on = k0 + j < e;
            i[j] = on ? ci[k0 + j] : 0;
            a[j] = on ? av[k0 + j] : A(0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) g[j] = k0 + j < e ? x[i[j]] : X(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (k0 + j < e) {
                if constexpr (std::is_same<A, double>::value) acc += a[j] * g[j];
                else dot32_add(acc, a[j], g[j]);
            }
        }
    }
    return acc;
}

// Copies the cnt <= CH (value, voxel) entries at global index base into a
// wavefront's LDS buffers with 16-byte loads (two FP64 or four FP32 values,
// and the matching voxel bytes per lane and step).  The copy starts at the
// 16-byte boundary at or below base, so the caller finds entry i at
// bv[shift + i], shift = the return value < V; the buffers hold cnt + 2V - 2
// entries, and the global arrays carry V - 1 entries of slack past their end.
// A lane issues all of its (at most K) loads before its first LDS store, so
// the copy waits for memory once.
template <class T, int G, int CH>
__device__ __forceinline__ int stage_entries(const T *__restrict__ val, const uint8_t *__restrict__ vox, long long base,
                                             int cnt, T *bv, uint8_t *bx, int lane) {
    constexpr int V = 16 / (int)sizeof(T);
    constexpr int K = (CH / V + 1 + G - 1) / G;  // 16-byte steps per lane: cnt <= CH spans <= CH/V + 1
    using VoxV = typename std::conditional<V == 2, uint16_t, uint32_t>::type;
    const long long a0 = base & ~(long long)(V - 1);
    const int sh = (int)(base - a0);
    const int nv = (sh + cnt + V - 1) / V;
    uint4 rv[K];
    VoxV rx[K];
    // steps past nv (cnt >= 1, so nv >= 1) repeat the last step: the same
    // bytes loaded and stored again, so no step needs a branch
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = lane + k * G < nv ? lane + k * G : nv - 1;
        rv[k] = *reinterpret_cast<const uint4 *>(val + a0 + (long long)i * V);
        rx[k] = *reinterpret_cast<const VoxV *>(vox + a0 + (long long)i * V);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = lane + k * G < nv ? lane + k * G : nv - 1;
        *reinterpret_cast<uint4 *>(bv + i * V) = rv[k];
        *reinterpret_cast<VoxV *>(bx + i * V) = rx[k];
    }
    return sh;
}

// cblas.h:53-63
__device__ __forceinline__ double d2norm(double a, double b) {
    const double scale = fabs(a) + fabs(b);
    if (scale == 0.0) return 0.0;
    const double as = a / scale, bs = b / scale;
    return scale * sqrt(as * as + bs * bs);
}

// ---------------------------------------------------------------------------
// Scalar recurrences of lsqr.c on one state.  The inputs are the sums of
// squares the vector kernels produced (kernel rows sk, damping rows sd, v, the
// dk terms); in exact-order mode the sequential norms held in the state are
// used instead.
// ---------------------------------------------------------------------------
__device__ inline void trace_row(double *trace, const Scal &s, double alpha) {
    if (!trace) return;
    double *t = trace + (size_t)s.itn * 8;
    t[0] = alpha;
    t[1] = s.beta;
    t[2] = s.rnorm;
    t[3] = s.test1;
    t[4] = s.test2;
    t[5] = s.anorm;
    t[6] = s.acond;
    t[7] = s.xnorm;
}

// Start: beta = ||[b;0]|| (lsqr.c:343); u is scaled lazily (ib).
__device__ inline void s_init_beta(Scal &s, double sk, double sd) {
    double beta;
    if (s.exact) {
        const double pk = s.pk, pd = sqrt(s.pd * s.pd);
        beta = sqrt(pk * pk + pd * pd);
    } else {
        beta = sqrt(sk + sd);
    }
    s.beta = beta;
    s.alpha = 0.0;
    if (beta > 0.0) {
        s.ib = 1.0 / beta;
        s.beta_pos = 1;
    } else {
        s.ib = 1.0;
        s.beta_pos = 0;
    }
    s.vscale = 1.0;  // first v += A^T u starts from v = 0 (lsqr.c:311,366)
}

// After the first v += A^T u: alpha, w = v / alpha, initial estimates
// (lsqr.c:375-445).
__device__ inline void s_init_alpha(Scal &s, double sv, double *trace) {
    double alpha = 0.0;
    if (s.beta_pos) alpha = s.exact ? sqrt(s.pv * s.pv) : sqrt(sv);
    s.alpha = alpha;
    if (alpha > 0.0) {
        s.inv_alpha = 1.0 / alpha;
        s.scale_v = 1;
    } else {
        s.inv_alpha = 1.0;
        s.scale_v = 0;
    }
    s.arnorm = alpha * s.beta;
    if (s.arnorm == 0.0) {  // goto_800
        s.done = 1;
        s.init_zero = 1;
        s.itn = 0;
        s.rnorm = 0.0;
        s.test1 = 0.0;
        s.test2 = 0.0;
        trace_row(trace, s, alpha);
        return;
    }
    s.rhobar = alpha;
    s.phibar = s.beta;
    s.bnorm = s.beta;
    s.rnorm = s.beta;
    s.test1 = 1.0;
    s.test2 = alpha / s.beta;
    s.itn = 0;
    trace_row(trace, s, alpha);
    if (s.itn_max <= 0) s.done = 1;
}

// After u <- u*(-alpha) + A v: beta, anorm (lsqr.c:487, 516-528, 530-538, 558).
__device__ inline void s_beta(Scal &s, double sk, double sd) {
    s.itn = s.itn + 1;
    double beta;
    if (s.exact) {
        const double pk = s.pk, pd = sqrt(s.pd * s.pd);
        beta = sqrt(pk * pk + pd * pd);
    } else {
        beta = sqrt(sk + sd);
    }
    s.beta = beta;
    double temp = d2norm(s.alpha, beta);
    temp = d2norm(temp, s.damp);
    s.anorm = d2norm(s.anorm, temp);
    if (beta > 0.0) {
        s.ib = 1.0 / beta;
        s.vscale = -beta;
        s.beta_pos = 1;
    } else {
        s.ib = 1.0;
        s.beta_pos = 0;
    }
}

// After v <- v*(-beta) + A^T u: alpha and the two plane rotations
// (lsqr.c:577-637).
__device__ inline void s_alpha(Scal &s, double sv) {
    s.scale_v = 0;
    if (s.beta_pos) {
        const double alpha = s.exact ? sqrt(s.pv * s.pv) : sqrt(sv);
        s.alpha = alpha;
        if (alpha > 0.0) {
            s.inv_alpha = 1.0 / alpha;
            s.scale_v = 1;
        }
    }
    const double alpha = s.alpha, beta = s.beta;
    double rhbar1 = s.rhobar;
    if (s.damp > 0.0) {
        rhbar1 = d2norm(s.rhobar, s.damp);
        const double cs1 = s.rhobar / rhbar1;
        const double sn1 = s.damp / rhbar1;
        s.psi = sn1 * s.phibar;
        s.phibar = cs1 * s.phibar;
    }
    const double rho = d2norm(rhbar1, beta);
    const double cs = rhbar1 / rho;
    const double sn = beta / rho;
    const double theta = sn * alpha;
    s.rhobar = -cs * alpha;
    const double phi = cs * s.phibar;
    s.phibar = sn * s.phibar;
    const double tau = sn * phi;
    s.phi = phi;
    s.rho = rho;
    s.t1 = phi / rho;
    s.t2 = -theta / rho;
    s.t3 = 1.0 / rho;
    s.theta = theta;
    s.arnorm = alpha * fabs(tau);  // lsqr.c:735 (alpha and tau are final here)
}

// After the x/w update: dknorm, norm and condition estimates, test1..3,
// istop and the stop decision (lsqr.c:703-775, 820-834).
__device__ inline void s_est(Scal &s, double dksum, double *trace) {
    const double theta = s.theta;
    double dknorm = s.exact ? s.dk_seq : dksum;
    dknorm = sqrt(dknorm);
    s.dnorm = d2norm(s.dnorm, dknorm);
    const double dxk = fabs(s.phi * dknorm);
    if (s.dxmax < dxk) {
        s.dxmax = dxk;
        s.maxdx = s.itn;
    }
    const double rho = s.rho;
    const double delta = s.sn2 * rho;
    const double gambar = -s.cs2 * rho;
    const double rhs = s.phi - delta * s.z;
    const double zbar = rhs / gambar;
    s.xnorm = d2norm(s.xnorm1, zbar);
    const double gamma = d2norm(gambar, theta);
    s.cs2 = gambar / gamma;
    s.sn2 = theta / gamma;
    s.z = rhs / gamma;
    s.xnorm1 = d2norm(s.xnorm1, s.z);
    s.acond = s.anorm * s.dnorm;
    s.res2 = d2norm(s.res2, s.psi);
    s.rnorm = d2norm(s.res2, s.phibar);
    s.test1 = s.rnorm / s.bnorm;
    s.test2 = 0.0;
    if (s.rnorm > 0.0) s.test2 = s.arnorm / (s.anorm * s.rnorm);
    s.test3 = 1.0 / s.acond;
    double t1 = s.test1 / (1.0 + s.anorm * s.xnorm / s.bnorm);
    const double rtol = s.btol + s.atol * s.anorm * s.xnorm / s.bnorm;
    const double t3 = 1.0 + s.test3;
    const double t2 = 1.0 + s.test2;
    t1 = 1.0 + t1;
    if (s.itn >= s.itnlim) s.istop = 5;
    if (t3 <= 1.0) s.istop = 4;
    if (t2 <= 1.0) s.istop = 2;
    if (t1 <= 1.0) s.istop = 1;
    if (s.test3 <= s.ctol) s.istop = 4;
    if (s.test2 <= s.atol) s.istop = 2;
    if (s.test1 <= rtol) s.istop = 1;
    if (s.istop == 0) {
        s.nstop = 0;
    } else {
        s.nstop = s.nstop + 1;
        if (s.nstop < 1 && s.itn < s.itnlim) s.istop = 0;
    }
    trace_row(trace, s, s.alpha);
    if (s.itn >= s.itn_max) s.done = 1;
    if (s.tol > 0.0 && s.test2 <= s.tol) s.done = 1;
}

}  // namespace plsqr
