// Device kernels of the warm-start network (included once, by plsqr_cnn.hip).
//
// A 3-D convolutional network maps the back-projection A^T b on the voxel
// grid to a starting point x0 for LSQR.  Every activation has C = 32 channels
// stored channels-last as bf16, element (g, c) at g*C + c with g = b*V + v the
// (sample, voxel) index; the network's one input and one output channel are
// channel 0 of a 32-channel tensor whose other channels are zero (input) or
// ignored (output), so that every layer is the same 32 -> 32 convolution.
//
// A 3x3x3 convolution with zero padding is an implicit GEMM: M = voxels,
// N = 32 output channels, K = 27 taps x 32 input channels.  One wavefront
// owns 32 consecutive voxels x 32 output channels and issues 108
// v_mfma_f32_32x32x8_bf16 (bf16 operands, FP32 accumulation):
//   A (32 x 8):  lane l holds row l%32, k = 4*(l/32) .. +3
//   B (8 x 32):  lane l holds column l%32, k = 4*(l/32) .. +3
//   D (32 x 32): lane l, register r holds row 8*(r/4) + 4*(l/32) + r%4,
//                column l%32
// Weights are bf16 in [tap][out channel][in channel] order so that a lane's
// four k values are contiguous.  Tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) reads
// the neighbour at (x+dx, y+dy, z+dz); tap 26-t is its mirror.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plsqr {
namespace cnn {

constexpr int C = 32;
constexpr int TAPS = 27;

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// float -> bf16 bits, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// voxel coordinates of the flattened (sample, voxel) index g
struct Vox {
    int x, y, z;
};

__device__ __forceinline__ Vox vox_of(long long g, int nx, int ny, int V) {
    const int v = (int)(g % V);
    Vox r;
    r.x = v % nx;
    r.y = (v / nx) % ny;
    r.z = v / (nx * ny);
    return r;
}

// ---------------------------------------------------------------------------
// out[g][n] = f(bias[n] + sum_t sum_k in[g + off(t)][k] * w[t][n][k]) for the
// `total` (sample, voxel) rows; neighbours outside the grid read as zero.
// f: ReLU when relu != 0; then, when mask is given, out = 0 where
// mask[g][n] <= 0 (the ReLU derivative of the layer below, for the data
// gradient).  Writes bf16 (out_bf) and/or FP32 (out_f), each may be NULL.
// Workgroups of 4 wavefronts; wavefront w of block b owns rows
// [(4b + w) * 32, +32).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kc_conv_fwd(int nx, int ny, int nz, long long total,
                                                   const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                                   const float *__restrict__ bias, const uint16_t *__restrict__ mask,
                                                   int relu, uint16_t *__restrict__ out_bf, float *__restrict__ out_f) {
    const int lane = threadIdx.x & 63;
    const long long g0 = ((long long)blockIdx.x * 4 + threadIdx.x / 64) * 32;
    if (g0 >= total) return;  // uniform per wavefront; no block barrier below
    const int V = nx * ny * nz;
    const int i = lane & 31, kq = lane >> 5;
    const long long g = g0 + i;
    const bool row_ok = g < total;
    Vox p{0, 0, 0};
    if (row_ok) p = vox_of(g, nx, ny, V);
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            const bool yz_ok = row_ok && (unsigned)(p.z + dz) < (unsigned)nz && (unsigned)(p.y + dy) < (unsigned)ny;
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                const bool ok = yz_ok && (unsigned)(p.x + dx) < (unsigned)nx;
                const long long src = ok ? (g + (long long)dz * nx * ny + (long long)dy * nx + dx) * C : 0;
                const uint16_t *wt = w + ((size_t)t * C + i) * C + kq * 4;
                const uint16_t *at = in + src + kq * 4;
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    short4_t a = {0, 0, 0, 0};
                    if (ok) a = *reinterpret_cast<const short4_t *>(at + kc * 8);
                    const short4_t b = *reinterpret_cast<const short4_t *>(wt + kc * 8);
                    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
                }
            }
        }
    const int n = i;
    const float bn = bias ? bias[n] : 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const long long gg = g0 + 8 * (r / 4) + 4 * kq + (r % 4);
        if (gg >= total) continue;
        float val = acc[r] + bn;
        if (relu) val = fmaxf(val, 0.f);
        if (mask && !(bf2f(mask[gg * C + n]) > 0.f)) val = 0.f;
        if (out_bf) out_bf[gg * C + n] = f2bf(val);
        if (out_f) out_f[gg * C + n] = val;
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: part[(t * gridDim.x + chunk)][n][k] = sum over the chunk's
// rows g of dout[g][n] * in[g + off(t)][k] (neighbours outside the grid are
// zero).  blockIdx.x = chunk of chunk_len rows, blockIdx.y = tap; the 4
// wavefronts take 8-row steps (MFMA reduction dimension = rows) and their
// 32 x 32 results are summed in wavefront order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kc_conv_wgrad(int nx, int ny, int nz, long long total, int chunk_len,
                                                     const uint16_t *__restrict__ in,
                                                     const uint16_t *__restrict__ dout, float *__restrict__ part) {
    __shared__ float red[4][C * C];
    const int t = blockIdx.y;
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const long long off = (long long)dz * nx * ny + (long long)dy * nx + dx;
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    const int i = lane & 31, kq = lane >> 5;
    const int V = nx * ny * nz;
    const long long c0 = (long long)blockIdx.x * chunk_len;
    const long long c1 = c0 + chunk_len < total ? c0 + chunk_len : total;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    for (long long gs = c0 + wave * 8; gs < c1; gs += 32) {
        short4_t a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long g = gs + kq * 4 + j;
            short av = 0, bv = 0;
            if (g < c1) {
                av = (short)dout[g * C + i];
                const Vox p = vox_of(g, nx, ny, V);
                if ((unsigned)(p.x + dx) < (unsigned)nx && (unsigned)(p.y + dy) < (unsigned)ny &&
                    (unsigned)(p.z + dz) < (unsigned)nz)
                    bv = (short)in[(g + off) * C + i];
            }
            a[j] = av;
            b[j] = bv;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    // D[row n][column k]: column = lane%32, rows per register
#pragma unroll
    for (int r = 0; r < 16; ++r) red[wave][(8 * (r / 4) + 4 * kq + (r % 4)) * C + i] = acc[r];
    __syncthreads();
    float *dst = part + ((size_t)t * gridDim.x + blockIdx.x) * (C * C);
    for (int e = threadIdx.x; e < C * C; e += 256) dst[e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

// grad[t][e] = sum over chunks (in order) of part[t][chunk][e]
__global__ void kc_wgrad_reduce(int nchunk, const float *__restrict__ part, float *__restrict__ grad) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= TAPS * C * C) return;
    const int t = idx / (C * C), e = idx % (C * C);
    const float *p = part + (size_t)t * nchunk * (C * C) + e;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += p[(size_t)c * (C * C)];
    grad[idx] = s;
}

// bias gradient partials: part[block][n] = sum over the block's rows of dout[g][n]
__global__ __launch_bounds__(256) void kc_bgrad(long long total, int chunk_len, const uint16_t *__restrict__ dout,
                                                float *__restrict__ part) {
    __shared__ float red[8][C];
    const int n = threadIdx.x % C, q = threadIdx.x / C;
    const long long c0 = (long long)blockIdx.x * chunk_len;
    const long long c1 = c0 + chunk_len < total ? c0 + chunk_len : total;
    float s = 0.f;
    for (long long g = c0 + q; g < c1; g += 8) s += bf2f(dout[g * C + n]);
    red[q][n] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float r = 0.f;
        for (int k = 0; k < 8; ++k) r += red[k][threadIdx.x];
        part[(size_t)blockIdx.x * C + threadIdx.x] = r;
    }
}

__global__ void kc_bgrad_reduce(int nblk, const float *__restrict__ part, float *__restrict__ grad) {
    const int n = threadIdx.x;
    if (n >= C) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * C + n];
    grad[n] = s;
}

// bf16 weights of one layer: forward order wf[t][n][k] = W[t][n][k] and the
// data-gradient order wb[t][k][n] = W[26 - t][n][k]
__global__ void kc_prep_weights(const float *__restrict__ W, uint16_t *__restrict__ wf, uint16_t *__restrict__ wb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= TAPS * C * C) return;
    const int t = idx / (C * C), n = (idx / C) % C, k = idx % C;
    const uint16_t h = f2bf(W[idx]);
    wf[idx] = h;
    wb[((size_t)(TAPS - 1 - t) * C + k) * C + n] = h;
}

// a0[g][c] = bf16(x[g]) for c = 0, 0 otherwise (x: one value per row)
__global__ void kc_expand_input(long long total, const float *__restrict__ x, uint16_t *__restrict__ a0) {
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (long long)gridDim.x * blockDim.x) {
        const uint16_t h = f2bf(x[g]);
        uint16_t *row = a0 + g * C;
        row[0] = h;
        for (int c = 1; c < C; ++c) row[c] = 0;
    }
}

// round an FP32 tensor to bf16
__global__ void kc_to_bf16(long long len, const float *__restrict__ a, uint16_t *__restrict__ b) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
        b[i] = f2bf(a[i]);
}

// ---------------------------------------------------------------------------
// Loss and its gradient.  Sample s of the batch has prediction z[g][0]
// (g in its V rows) and target tgt[s*V + v] with tn[s] = ||tgt_s||^2:
//   loss = (1/B) sum_s ||z_s - tgt_s||^2 / tn[s]
//   dz[g][0] = bf16(2 (z - tgt) / (B tn[s])), other channels 0.
// part[s * nblk + block] = the block's share of ||z_s - tgt_s||^2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kc_loss(int V, int B, int nblk, const float *__restrict__ z,
                                               const float *__restrict__ tgt, const float *__restrict__ tn,
                                               uint16_t *__restrict__ dz, float *__restrict__ part) {
    __shared__ float red[256];
    const int s = blockIdx.y;
    float acc = 0.f;
    const float scale = 2.f / ((float)B * tn[s]);
    for (int v = blockIdx.x * 256 + threadIdx.x; v < V; v += nblk * 256) {
        const long long g = (long long)s * V + v;
        const float d = z[g * C] - tgt[g];
        acc += d * d;
        uint16_t *row = dz + g * C;
        row[0] = f2bf(scale * d);
        for (int c = 1; c < C; ++c) row[c] = 0;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)s * nblk + blockIdx.x] = red[0];
}

// Adam (Kingma and Ba, 2015) on FP32 master parameters; mask[i] == 0 keeps a
// padded parameter at zero.
__global__ void kc_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                        float *__restrict__ v, const uint8_t *__restrict__ mask, float lr, float b1, float b2,
                        float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !mask[i]) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace cnn
}  // namespace plsqr
