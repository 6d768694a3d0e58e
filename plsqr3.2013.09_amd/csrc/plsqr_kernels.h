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

// cblas.h:53-63
__device__ __forceinline__ double d2norm(double a, double b) {
    const double scale = fabs(a) + fabs(b);
    if (scale == 0.0) return 0.0;
    const double as = a / scale, bs = b / scale;
    return scale * sqrt(as * as + bs * bs);
}

}  // namespace plsqr
