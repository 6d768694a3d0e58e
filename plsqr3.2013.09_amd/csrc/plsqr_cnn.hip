// C-ABI of the warm-start network (include/plsqr_hip.h, plsqr_cnn_*): a
// 4-layer 3x3x3 convolutional network, 32 channels, bf16 weights and
// activations with FP32 accumulation on the MFMA units, trained on the GPU
// with Adam on FP32 master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "plsqr_cnn.h"
#include "plsqr_hip.h"

using namespace plsqr::cnn;

namespace {

constexpr int LAYERS = 4;
constexpr int WSZ = TAPS * C * C;      // weights of one layer
constexpr int LSZ = WSZ + C;           // weights + bias
constexpr int NPARAM = LAYERS * LSZ + 1;  // + the output gain (last entry, not trained)
constexpr int GAIN = NPARAM - 1;
constexpr int WCHUNK = 4096;           // rows per weight-gradient workgroup
constexpr int BCHUNK = 4096;           // rows per bias-gradient workgroup
constexpr int LOSS_BLK = 64;           // workgroups per sample in the loss

thread_local std::string g_cnn_err;

}  // namespace

struct plsqr_cnn {
    int nx = 0, ny = 0, nz = 0, V = 0, max_batch = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint8_t *mask = nullptr;
    uint16_t *wf[LAYERS] = {}, *wb[LAYERS] = {};
    bool weights_dirty = true;
    uint16_t *act[LAYERS] = {};
    float *zout = nullptr;
    uint16_t *dz[2] = {};
    float *wpart = nullptr, *bpart = nullptr, *lpart = nullptr;
    float *xin = nullptr, *tgt = nullptr, *tn = nullptr, *x0 = nullptr;
    int count = 0;
    float *data_in = nullptr, *data_t = nullptr, *data_tn = nullptr;
    size_t data_bytes = 0;
    long long adam_t = 0;
    float gain = 1.f;  // host copy of params[GAIN]
    std::vector<void *> allocs;
    size_t bytes = 0;
    std::string err;
};

namespace {

int set_err(plsqr_cnn *h, int code, const std::string &msg) {
    if (h) {
        if (h->err.empty()) h->err = msg;
    } else {
        g_cnn_err = msg;
    }
    return code;
}

#define CN_TRY(h, expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_err(h, PLSQR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CN_LAUNCH(h) CN_TRY(h, hipGetLastError())

template <class T>
int calloc_dev(plsqr_cnn *h, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(h, PLSQR_ERR_NOMEM,
                       "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed: " + hipGetErrorString(e));
    }
    h->allocs.push_back((void *)*p);
    h->bytes += sizeof(T) * count;
    return PLSQR_OK;
}

void free_dev(plsqr_cnn *h, void *p) {
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) {
        h->allocs.erase(it);
        (void)hipFree(p);
    }
}

int grid_for(long long work, int per, int cap) {
    long long g = (work + per - 1) / per;
    return (int)std::max(1LL, std::min<long long>(g, cap));
}

// SplitMix64 (Steele, Lea and Flood, 2014) and Box-Muller normals
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// zero for the padded parameters: first layer input channels 1..31, last
// layer output channels 1..31 (weights and bias)
bool param_live(int i) {
    if (i == GAIN) return false;
    const int l = i / LSZ, j = i % LSZ;
    if (j < WSZ) {
        const int n = (j / C) % C, k = j % C;
        if (l == 0 && k != 0) return false;
        if (l == LAYERS - 1 && n != 0) return false;
        return true;
    }
    const int n = j - WSZ;
    return !(l == LAYERS - 1 && n != 0);
}

int prep_weights(plsqr_cnn *h) {
    if (!h->weights_dirty) return PLSQR_OK;
    for (int l = 0; l < LAYERS; ++l)
        hipLaunchKernelGGL(kc_prep_weights, dim3((WSZ + 255) / 256), dim3(256), 0, h->stream, h->params + (size_t)l * LSZ,
                           h->wf[l], h->wb[l]);
    CN_LAUNCH(h);
    h->weights_dirty = false;
    return PLSQR_OK;
}

int conv_fwd(plsqr_cnn *h, long long total, const uint16_t *in, const uint16_t *w, const float *bias,
             const uint16_t *mask, int relu, uint16_t *out_bf, float *out_f, hipStream_t st) {
    const long long waves = (total + 31) / 32;
    const long long blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(kc_conv_fwd, dim3((unsigned)blocks), dim3(256), 0, st, h ? h->nx : 0, h ? h->ny : 0,
                       h ? h->nz : 0, total, in, w, bias, mask, relu, out_bf, out_f);
    return PLSQR_OK;
}

// the network on h->xin (batch samples): activations act[0..3], zout
int forward(plsqr_cnn *h, int batch) {
    int rc;
    if ((rc = prep_weights(h))) return rc;
    const long long total = (long long)batch * h->V;
    hipLaunchKernelGGL(kc_expand_input, dim3(grid_for(total, 256, 4096)), dim3(256), 0, h->stream, total, h->xin,
                       h->act[0]);
    CN_LAUNCH(h);
    for (int l = 0; l < LAYERS; ++l) {
        const float *bias = h->params + (size_t)l * LSZ + WSZ;
        const bool last = l == LAYERS - 1;
        const long long blocks = ((total + 31) / 32 + 3) / 4;
        hipLaunchKernelGGL(kc_conv_fwd, dim3((unsigned)blocks), dim3(256), 0, h->stream, h->nx, h->ny, h->nz, total,
                           h->act[l], h->wf[l], bias, (const uint16_t *)nullptr, last ? 0 : 1,
                           last ? (uint16_t *)nullptr : h->act[l + 1], last ? h->zout : (float *)nullptr);
        CN_LAUNCH(h);
    }
    return PLSQR_OK;
}

int wgrad(plsqr_cnn *h, int nx, int ny, int nz, long long total, const uint16_t *in, const uint16_t *dout, float *gw,
          float *gb, float *wpart, float *bpart, hipStream_t st) {
    const int nch = (int)((total + WCHUNK - 1) / WCHUNK);
    hipLaunchKernelGGL(kc_conv_wgrad, dim3(nch, TAPS), dim3(256), 0, st, nx, ny, nz, total, WCHUNK, in, dout, wpart);
    hipLaunchKernelGGL(kc_wgrad_reduce, dim3((WSZ + 255) / 256), dim3(256), 0, st, nch, (const float *)wpart, gw);
    const int nb = (int)((total + BCHUNK - 1) / BCHUNK);
    hipLaunchKernelGGL(kc_bgrad, dim3(nb), dim3(256), 0, st, total, BCHUNK, dout, bpart);
    hipLaunchKernelGGL(kc_bgrad_reduce, dim3(1), dim3(C), 0, st, nb, (const float *)bpart, gb);
    return hipGetLastError() == hipSuccess ? PLSQR_OK : set_err(h, PLSQR_ERR_HIP, "gradient kernels failed to launch");
}

// loss of zout against tgt/tn (batch samples), dz[0] = its gradient; *loss on the host
int loss_and_grad(plsqr_cnn *h, int batch, double *loss) {
    hipLaunchKernelGGL(kc_loss, dim3(LOSS_BLK, batch), dim3(256), 0, h->stream, h->V, batch, LOSS_BLK,
                       (const float *)h->zout, (const float *)h->tgt, (const float *)h->tn, h->dz[0], h->lpart);
    CN_LAUNCH(h);
    std::vector<float> part((size_t)batch * LOSS_BLK), tn((size_t)batch);
    CN_TRY(h, hipMemcpyAsync(part.data(), h->lpart, sizeof(float) * part.size(), hipMemcpyDeviceToHost, h->stream));
    CN_TRY(h, hipMemcpyAsync(tn.data(), h->tn, sizeof(float) * tn.size(), hipMemcpyDeviceToHost, h->stream));
    CN_TRY(h, hipStreamSynchronize(h->stream));
    double L = 0.0;
    for (int s = 0; s < batch; ++s) {
        double e = 0.0;
        for (int b = 0; b < LOSS_BLK; ++b) e += part[(size_t)s * LOSS_BLK + b];
        L += e / tn[(size_t)s];
    }
    *loss = L / batch;
    return PLSQR_OK;
}

int backward(plsqr_cnn *h, int batch) {
    const long long total = (long long)batch * h->V;
    int cur = 0, rc;
    for (int l = LAYERS - 1; l >= 0; --l) {
        float *gw = h->grads + (size_t)l * LSZ;
        if ((rc = wgrad(h, h->nx, h->ny, h->nz, total, h->act[l], h->dz[cur], gw, gw + WSZ, h->wpart, h->bpart,
                        h->stream)))
            return rc;
        if (l > 0) {  // gradient of the layer's input, through the ReLU of layer l-1
            const long long blocks = ((total + 31) / 32 + 3) / 4;
            hipLaunchKernelGGL(kc_conv_fwd, dim3((unsigned)blocks), dim3(256), 0, h->stream, h->nx, h->ny, h->nz, total,
                               h->dz[cur], h->wb[l], (const float *)nullptr, h->act[l], 0, h->dz[1 - cur],
                               (float *)nullptr);
            CN_LAUNCH(h);
            cur = 1 - cur;
        }
    }
    return PLSQR_OK;
}

// normalised copies of caller samples: in = atb / s, t = x / s, s = rms(atb)
void normalise(int V, const double *atb, const double *xs, double gain, std::vector<float> &in, std::vector<float> &t,
               float *tn, double *scale) {
    double ss = 0.0;
    for (int v = 0; v < V; ++v) ss += atb[v] * atb[v];
    double s = std::sqrt(ss / V);
    if (!(s > 0.0) || !std::isfinite(s)) s = 1.0;
    *scale = s;
    double tt = 0.0;
    for (int v = 0; v < V; ++v) {
        in[(size_t)v] = (float)(atb[v] / s);
        if (xs) {
            const double q = xs[v] / s / gain;
            t[(size_t)v] = (float)q;
            tt += q * q;
        }
    }
    if (tn) *tn = (float)(tt > 0.0 ? tt : 1.0);
}

int check(plsqr_cnn *h) {
    if (!h) return PLSQR_ERR_ARG;
    if (h->allocs.empty()) return set_err(h, PLSQR_ERR_STATE, "network failed to initialise");
    return PLSQR_OK;
}

// load `batch` caller samples into xin/tgt/tn
int load_samples(plsqr_cnn *h, int batch, const double *atb, const double *xs, std::vector<double> *scales) {
    std::vector<float> in((size_t)h->V), t((size_t)h->V);
    std::vector<float> tn((size_t)batch, 1.f);
    if (scales) scales->assign((size_t)batch, 1.0);
    for (int s = 0; s < batch; ++s) {
        double sc;
        normalise(h->V, atb + (size_t)s * h->V, xs ? xs + (size_t)s * h->V : nullptr, h->gain, in, t, &tn[(size_t)s],
                  &sc);
        if (scales) (*scales)[(size_t)s] = sc;
        CN_TRY(h, hipMemcpyAsync(h->xin + (size_t)s * h->V, in.data(), sizeof(float) * h->V, hipMemcpyHostToDevice,
                                 h->stream));
        if (xs)
            CN_TRY(h, hipMemcpyAsync(h->tgt + (size_t)s * h->V, t.data(), sizeof(float) * h->V, hipMemcpyHostToDevice,
                                     h->stream));
        CN_TRY(h, hipStreamSynchronize(h->stream));  // the host vectors are reused
    }
    CN_TRY(h, hipMemcpyAsync(h->tn, tn.data(), sizeof(float) * batch, hipMemcpyHostToDevice, h->stream));
    CN_TRY(h, hipStreamSynchronize(h->stream));
    return PLSQR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free variant of plsqr_cnn_conv_forward for callers that cannot
// allocate device memory (HIP graph capture, serving loops).  Its scratch --
// the bf16 copy of the input and the two bf16 weight layouts -- lives in one
// file-scope device array, fixed in size; each device holds its own copy,
// found once per device with hipGetSymbolAddress.  Every element the kernels
// read is written earlier in the same call, so no clearing is needed; calls
// that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
constexpr long long NOALLOC_MAX_ELEMS = 1LL << 22;  // batch * voxels * 32 channels of input
constexpr size_t NOALLOC_BYTES = sizeof(uint16_t) * ((size_t)NOALLOC_MAX_ELEMS + 2 * (size_t)WSZ);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// The current device's address of `scratch` (looked up on the first call on
// that device, with that device current).
int scratch_ptr(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_err(nullptr, PLSQR_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return set_err(nullptr, PLSQR_ERR_STATE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(nullptr, PLSQR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PLSQR_OK;
}
}  // namespace

extern "C" {

const char *plsqr_cnn_last_error(const plsqr_cnn *h) { return h ? h->err.c_str() : g_cnn_err.c_str(); }

int plsqr_cnn_param_count(void) { return NPARAM; }

void plsqr_cnn_destroy(plsqr_cnn *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int plsqr_cnn_create(int nx, int ny, int nz, int max_batch, unsigned long long seed, plsqr_cnn **out) {
    g_cnn_err.clear();
    if (!out) return PLSQR_ERR_ARG;
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1 || max_batch < 1 || (long long)nx * ny * nz > (1LL << 26) ||
        (long long)nx * ny * nz * max_batch * C > (1LL << 33))
        return set_err(nullptr, PLSQR_ERR_ARG, "grid or batch out of range");
    plsqr_cnn *h = new plsqr_cnn();
    h->nx = nx;
    h->ny = ny;
    h->nz = nz;
    h->V = nx * ny * nz;
    h->max_batch = max_batch;
    auto fail = [&](int code) {
        g_cnn_err = h->err;
        plsqr_cnn_destroy(h);
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(set_err(h, PLSQR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    const long long total = (long long)max_batch * h->V;
    const int nch = (int)((total + WCHUNK - 1) / WCHUNK), nb = (int)((total + BCHUNK - 1) / BCHUNK);
    int rc;
    if ((rc = calloc_dev(h, &h->params, NPARAM)) || (rc = calloc_dev(h, &h->grads, NPARAM)) ||
        (rc = calloc_dev(h, &h->adam_m, NPARAM)) || (rc = calloc_dev(h, &h->adam_v, NPARAM)) ||
        (rc = calloc_dev(h, &h->mask, NPARAM)) || (rc = calloc_dev(h, &h->zout, (size_t)total * C)) ||
        (rc = calloc_dev(h, &h->dz[0], (size_t)total * C)) || (rc = calloc_dev(h, &h->dz[1], (size_t)total * C)) ||
        (rc = calloc_dev(h, &h->wpart, (size_t)nch * TAPS * C * C)) || (rc = calloc_dev(h, &h->bpart, (size_t)nb * C)) ||
        (rc = calloc_dev(h, &h->lpart, (size_t)max_batch * LOSS_BLK)) || (rc = calloc_dev(h, &h->xin, (size_t)total)) ||
        (rc = calloc_dev(h, &h->tgt, (size_t)total)) || (rc = calloc_dev(h, &h->tn, (size_t)max_batch)) ||
        (rc = calloc_dev(h, &h->x0, (size_t)total)))
        return fail(rc);
    for (int l = 0; l < LAYERS; ++l)
        if ((rc = calloc_dev(h, &h->wf[l], WSZ)) || (rc = calloc_dev(h, &h->wb[l], WSZ)) ||
            (rc = calloc_dev(h, &h->act[l], (size_t)total * C)))
            return fail(rc);
    // He initialisation (fan-in = 27 x real input channels), zero biases
    std::vector<float> p((size_t)NPARAM, 0.f);
    std::vector<uint8_t> m((size_t)NPARAM, 0);
    Rng rng{seed};
    for (int i = 0; i < NPARAM; ++i) {
        m[(size_t)i] = param_live(i) ? 1 : 0;
        const int l = i / LSZ, j = i % LSZ;
        if (!m[(size_t)i] || j >= WSZ) continue;
        const double fan_in = TAPS * (l == 0 ? 1.0 : (double)C);
        const double sd = std::sqrt((l == LAYERS - 1 ? 1.0 : 2.0) / fan_in);
        p[(size_t)i] = (float)(sd * rng.normal());
    }
    p[GAIN] = 1.f;
    e = hipMemcpy(h->params, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->mask, m.data(), NPARAM, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->adam_m, 0, sizeof(float) * NPARAM);
    if (e == hipSuccess) e = hipMemset(h->adam_v, 0, sizeof(float) * NPARAM);
    if (e == hipSuccess) e = hipMemset(h->grads, 0, sizeof(float) * NPARAM);
    if (e != hipSuccess) return fail(set_err(h, PLSQR_ERR_HIP, std::string("initialise: ") + hipGetErrorString(e)));
    *out = h;
    return PLSQR_OK;
}

size_t plsqr_cnn_device_bytes(const plsqr_cnn *h) { return h ? h->bytes : 0; }

int plsqr_cnn_get_params(plsqr_cnn *h, float *params) {
    int rc;
    if ((rc = check(h))) return rc;
    if (!params) return set_err(h, PLSQR_ERR_ARG, "params is NULL");
    CN_TRY(h, hipMemcpyAsync(params, h->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, h->stream));
    CN_TRY(h, hipStreamSynchronize(h->stream));
    return PLSQR_OK;
}

int plsqr_cnn_set_params(plsqr_cnn *h, const float *params) {
    int rc;
    if ((rc = check(h))) return rc;
    if (!params) return set_err(h, PLSQR_ERR_ARG, "params is NULL");
    for (int i = 0; i < NPARAM; ++i) {
        if (!std::isfinite(params[i])) return set_err(h, PLSQR_ERR_NONFINITE, "parameter " + std::to_string(i) + " is not finite");
        if (i != GAIN && !param_live(i) && params[i] != 0.f)
            return set_err(h, PLSQR_ERR_ARG, "parameter " + std::to_string(i) + " is a padded entry and must be 0");
    }
    if (!(params[GAIN] > 0.f)) return set_err(h, PLSQR_ERR_ARG, "the output gain must be positive");
    h->gain = params[GAIN];
    CN_TRY(h, hipMemcpyAsync(h->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, h->stream));
    CN_TRY(h, hipStreamSynchronize(h->stream));
    h->weights_dirty = true;
    return PLSQR_OK;
}

int plsqr_cnn_set_data(plsqr_cnn *h, int count, const double *atb, const double *xstar) {
    int rc;
    if ((rc = check(h))) return rc;
    if (count < 1 || !atb || !xstar) return set_err(h, PLSQR_ERR_ARG, "count, atb and xstar are required");
    for (float **p : {&h->data_in, &h->data_t, &h->data_tn})
        if (*p) {
            free_dev(h, *p);
            *p = nullptr;
        }
    h->count = 0;
    if ((rc = calloc_dev(h, &h->data_in, (size_t)count * h->V)) || (rc = calloc_dev(h, &h->data_t, (size_t)count * h->V)) ||
        (rc = calloc_dev(h, &h->data_tn, (size_t)count)))
        return rc;
    std::vector<float> in((size_t)h->V), t((size_t)h->V), tn((size_t)count);
    // output gain: the RMS of the scaled targets x*/s over the set, so that
    // the network's output channel is of order one
    double gsum = 0.0;
    for (int s = 0; s < count; ++s) {
        const double *a = atb + (size_t)s * h->V, *x = xstar + (size_t)s * h->V;
        double aa = 0.0, xx = 0.0;
        for (int v = 0; v < h->V; ++v) {
            if (!std::isfinite(a[v]) || !std::isfinite(x[v]))
                return set_err(h, PLSQR_ERR_NONFINITE, "sample " + std::to_string(s) + " holds a NaN or Inf");
            aa += a[v] * a[v];
            xx += x[v] * x[v];
        }
        if (aa > 0.0) gsum += xx / aa;  // mean((x/s)^2) with s^2 = aa / V
    }
    const double g = std::sqrt(gsum / count);
    h->gain = (float)(g > 0.0 && std::isfinite(g) ? g : 1.0);
    CN_TRY(h, hipMemcpyAsync(h->params + GAIN, &h->gain, sizeof(float), hipMemcpyHostToDevice, h->stream));
    for (int s = 0; s < count; ++s) {
        double sc;
        const double *a = atb + (size_t)s * h->V, *x = xstar + (size_t)s * h->V;
        normalise(h->V, a, x, h->gain, in, t, &tn[(size_t)s], &sc);
        CN_TRY(h, hipMemcpyAsync(h->data_in + (size_t)s * h->V, in.data(), sizeof(float) * h->V, hipMemcpyHostToDevice,
                                 h->stream));
        CN_TRY(h, hipMemcpyAsync(h->data_t + (size_t)s * h->V, t.data(), sizeof(float) * h->V, hipMemcpyHostToDevice,
                                 h->stream));
        CN_TRY(h, hipStreamSynchronize(h->stream));
    }
    CN_TRY(h, hipMemcpyAsync(h->data_tn, tn.data(), sizeof(float) * count, hipMemcpyHostToDevice, h->stream));
    CN_TRY(h, hipStreamSynchronize(h->stream));
    h->count = count;
    return PLSQR_OK;
}

int plsqr_cnn_train(plsqr_cnn *h, int steps, int batch, float lr, unsigned long long seed, double *loss_out) {
    int rc;
    if ((rc = check(h))) return rc;
    if (h->count < 1) return set_err(h, PLSQR_ERR_STATE, "no training data (plsqr_cnn_set_data)");
    if (batch < 1 || batch > h->max_batch || batch > h->count || steps < 0 || !(lr > 0.f))
        return set_err(h, PLSQR_ERR_ARG, "bad steps, batch or learning rate");
    std::vector<int> order((size_t)h->count);
    Rng rng{seed};
    size_t pos = order.size();
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < steps; ++it) {
        std::vector<int> pick((size_t)batch);
        for (int s = 0; s < batch; ++s) {
            if (pos >= order.size()) {  // a new epoch: Fisher-Yates shuffle
                for (int i = 0; i < h->count; ++i) order[(size_t)i] = i;
                for (int i = h->count - 1; i > 0; --i) std::swap(order[(size_t)i], order[(size_t)(rng.next() % (uint64_t)(i + 1))]);
                pos = 0;
            }
            pick[(size_t)s] = order[pos++];
        }
        for (int s = 0; s < batch; ++s) {
            const size_t src = (size_t)pick[(size_t)s] * h->V;
            CN_TRY(h, hipMemcpyAsync(h->xin + (size_t)s * h->V, h->data_in + src, sizeof(float) * h->V,
                                     hipMemcpyDeviceToDevice, h->stream));
            CN_TRY(h, hipMemcpyAsync(h->tgt + (size_t)s * h->V, h->data_t + src, sizeof(float) * h->V,
                                     hipMemcpyDeviceToDevice, h->stream));
            CN_TRY(h, hipMemcpyAsync(h->tn + s, h->data_tn + pick[(size_t)s], sizeof(float), hipMemcpyDeviceToDevice,
                                     h->stream));
        }
        double loss;
        if ((rc = forward(h, batch)) || (rc = loss_and_grad(h, batch, &loss)) || (rc = backward(h, batch))) return rc;
        ++h->adam_t;
        const float c1 = 1.f - std::pow(b1, (float)h->adam_t), c2 = 1.f - std::pow(b2, (float)h->adam_t);
        hipLaunchKernelGGL(kc_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, h->stream, NPARAM, h->params,
                           (const float *)h->grads, h->adam_m, h->adam_v, (const uint8_t *)h->mask, lr, b1, b2, eps, c1, c2);
        CN_LAUNCH(h);
        h->weights_dirty = true;
        if (loss_out) loss_out[it] = loss;
    }
    CN_TRY(h, hipStreamSynchronize(h->stream));
    return PLSQR_OK;
}

int plsqr_cnn_gradient(plsqr_cnn *h, int batch, const double *atb, const double *xstar, float *grads, double *loss) {
    int rc;
    if ((rc = check(h))) return rc;
    if (batch < 1 || batch > h->max_batch || !atb || !xstar || !grads || !loss)
        return set_err(h, PLSQR_ERR_ARG, "bad gradient arguments");
    if ((rc = load_samples(h, batch, atb, xstar, nullptr))) return rc;
    if ((rc = forward(h, batch)) || (rc = loss_and_grad(h, batch, loss)) || (rc = backward(h, batch))) return rc;
    CN_TRY(h, hipMemcpyAsync(grads, h->grads, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, h->stream));
    CN_TRY(h, hipStreamSynchronize(h->stream));
    return PLSQR_OK;
}

int plsqr_cnn_predict(plsqr_cnn *h, int count, const double *atb, double *x0) {
    int rc;
    if ((rc = check(h))) return rc;
    if (count < 0 || (count > 0 && (!atb || !x0))) return set_err(h, PLSQR_ERR_ARG, "atb and x0 are required");
    std::vector<float> z((size_t)h->V * C);
    for (int s0 = 0; s0 < count; s0 += h->max_batch) {
        const int nb = std::min(h->max_batch, count - s0);
        std::vector<double> scales;
        if ((rc = load_samples(h, nb, atb + (size_t)s0 * h->V, nullptr, &scales)) || (rc = forward(h, nb))) return rc;
        for (int s = 0; s < nb; ++s) {
            CN_TRY(h, hipMemcpyAsync(z.data(), h->zout + (size_t)s * h->V * C, sizeof(float) * z.size(),
                                     hipMemcpyDeviceToHost, h->stream));
            CN_TRY(h, hipStreamSynchronize(h->stream));
            double *o = x0 + (size_t)(s0 + s) * h->V;
            const double f = scales[(size_t)s] * (double)h->gain;
            for (int v = 0; v < h->V; ++v) o[v] = f * (double)z[(size_t)v * C];
        }
    }
    return PLSQR_OK;
}

// One 32 -> 32 convolution layer through the MFMA kernels, host arrays
// (operands rounded to bf16 inside; results in FP32).
int plsqr_cnn_conv_forward(int nx, int ny, int nz, int batch, const float *in, const float *w, const float *bias,
                           int relu, float *out) {
    g_cnn_err.clear();
    if (nx < 1 || ny < 1 || nz < 1 || batch < 1 || !in || !w || !out) return set_err(nullptr, PLSQR_ERR_ARG, "bad arguments");
    const long long total = (long long)batch * nx * ny * nz;
    plsqr_cnn tmp;
    tmp.nx = nx;
    tmp.ny = ny;
    tmp.nz = nz;
    plsqr_cnn *h = &tmp;
    float *din = nullptr, *dw = nullptr, *db = nullptr, *dout = nullptr;
    uint16_t *bin = nullptr, *wf = nullptr, *wb = nullptr;
    int rc = PLSQR_OK;
    auto cleanup = [&]() {
        for (void *p : tmp.allocs) (void)hipFree(p);
        tmp.allocs.clear();
    };
    if ((rc = calloc_dev(h, &din, (size_t)total * C)) || (rc = calloc_dev(h, &dw, WSZ)) || (rc = calloc_dev(h, &db, C)) ||
        (rc = calloc_dev(h, &dout, (size_t)total * C)) || (rc = calloc_dev(h, &bin, (size_t)total * C)) ||
        (rc = calloc_dev(h, &wf, WSZ)) || (rc = calloc_dev(h, &wb, WSZ))) {
        g_cnn_err = tmp.err;
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(din, in, sizeof(float) * total * C, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dw, w, sizeof(float) * WSZ, hipMemcpyHostToDevice);
    if (e == hipSuccess && bias) e = hipMemcpy(db, bias, sizeof(float) * C, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(kc_to_bf16, dim3(grid_for(total * C, 256, 4096)), dim3(256), 0, 0, total * C,
                           (const float *)din, bin);
        hipLaunchKernelGGL(kc_prep_weights, dim3((WSZ + 255) / 256), dim3(256), 0, 0, (const float *)dw, wf, wb);
        conv_fwd(h, total, bin, wf, bias ? db : nullptr, nullptr, relu, nullptr, dout, 0);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(float) * total * C, hipMemcpyDeviceToHost);
    }
    cleanup();
    if (e != hipSuccess) return set_err(nullptr, PLSQR_ERR_HIP, std::string("conv forward: ") + hipGetErrorString(e));
    return PLSQR_OK;
}

// plsqr_cnn_conv_forward on device pointers with the scratch above: no allocation.
int plsqr_cnn_conv_forward_noalloc(int nx, int ny, int nz, int batch, const float *in_dev, const float *w_dev,
                                   const float *bias_dev, int relu, float *out_dev, void *stream) {
    g_cnn_err.clear();
    if (nx < 1 || ny < 1 || nz < 1 || batch < 1 || !in_dev || !w_dev || !out_dev)
        return set_err(nullptr, PLSQR_ERR_ARG, "bad arguments");
    const long long total = (long long)batch * nx * ny * nz;
    if (total > NOALLOC_MAX_ELEMS / C)
        return set_err(nullptr, PLSQR_ERR_ARG,
                       "batch * nx * ny * nz = " + std::to_string(total) + " exceeds the allocation-free limit of " +
                           std::to_string(NOALLOC_MAX_ELEMS / C) + " voxels");
    void *base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    uint16_t *bin = (uint16_t *)base;
    uint16_t *wf = bin + NOALLOC_MAX_ELEMS, *wb = wf + WSZ;
    const hipStream_t st = (hipStream_t)stream;
    plsqr_cnn tmp;  // grid only: no buffers
    tmp.nx = nx;
    tmp.ny = ny;
    tmp.nz = nz;
    hipLaunchKernelGGL(kc_to_bf16, dim3(grid_for(total * C, 256, 4096)), dim3(256), 0, st, total * C, in_dev, bin);
    hipLaunchKernelGGL(kc_prep_weights, dim3((WSZ + 255) / 256), dim3(256), 0, st, w_dev, wf, wb);
    conv_fwd(&tmp, total, bin, wf, bias_dev, nullptr, relu, nullptr, out_dev, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(nullptr, PLSQR_ERR_HIP, std::string("conv forward: ") + hipGetErrorString(e));
    return PLSQR_OK;
}

// Gradients of one layer: din = data gradient (no ReLU mask), dw, dbias.
int plsqr_cnn_conv_backward(int nx, int ny, int nz, int batch, const float *in, const float *w, const float *dout,
                            float *din, float *dw, float *dbias) {
    g_cnn_err.clear();
    if (nx < 1 || ny < 1 || nz < 1 || batch < 1 || !in || !w || !dout || !din || !dw || !dbias)
        return set_err(nullptr, PLSQR_ERR_ARG, "bad arguments");
    const long long total = (long long)batch * nx * ny * nz;
    plsqr_cnn tmp;
    tmp.nx = nx;
    tmp.ny = ny;
    tmp.nz = nz;
    plsqr_cnn *h = &tmp;
    const int nch = (int)((total + WCHUNK - 1) / WCHUNK), nb = (int)((total + BCHUNK - 1) / BCHUNK);
    float *fin = nullptr, *fw = nullptr, *fdo = nullptr, *fdi = nullptr, *gw = nullptr, *wpart = nullptr, *bpart = nullptr;
    uint16_t *bin = nullptr, *bdo = nullptr, *wf = nullptr, *wb = nullptr;
    int rc;
    auto cleanup = [&]() {
        for (void *p : tmp.allocs) (void)hipFree(p);
        tmp.allocs.clear();
    };
    if ((rc = calloc_dev(h, &fin, (size_t)total * C)) || (rc = calloc_dev(h, &fw, WSZ)) ||
        (rc = calloc_dev(h, &fdo, (size_t)total * C)) || (rc = calloc_dev(h, &fdi, (size_t)total * C)) ||
        (rc = calloc_dev(h, &gw, LSZ)) || (rc = calloc_dev(h, &wpart, (size_t)nch * TAPS * C * C)) ||
        (rc = calloc_dev(h, &bpart, (size_t)nb * C)) || (rc = calloc_dev(h, &bin, (size_t)total * C)) ||
        (rc = calloc_dev(h, &bdo, (size_t)total * C)) || (rc = calloc_dev(h, &wf, WSZ)) || (rc = calloc_dev(h, &wb, WSZ))) {
        g_cnn_err = tmp.err;
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(fin, in, sizeof(float) * total * C, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(fw, w, sizeof(float) * WSZ, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(fdo, dout, sizeof(float) * total * C, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int g = grid_for(total * C, 256, 4096);
        hipLaunchKernelGGL(kc_to_bf16, dim3(g), dim3(256), 0, 0, total * C, (const float *)fin, bin);
        hipLaunchKernelGGL(kc_to_bf16, dim3(g), dim3(256), 0, 0, total * C, (const float *)fdo, bdo);
        hipLaunchKernelGGL(kc_prep_weights, dim3((WSZ + 255) / 256), dim3(256), 0, 0, (const float *)fw, wf, wb);
        conv_fwd(h, total, bdo, wb, nullptr, nullptr, 0, nullptr, fdi, 0);
        rc = wgrad(h, nx, ny, nz, total, bin, bdo, gw, gw + WSZ, wpart, bpart, 0);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(din, fdi, sizeof(float) * total * C, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(dw, gw, sizeof(float) * WSZ, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(dbias, gw + WSZ, sizeof(float) * C, hipMemcpyDeviceToHost);
    }
    cleanup();
    if (e != hipSuccess) return set_err(nullptr, PLSQR_ERR_HIP, std::string("conv backward: ") + hipGetErrorString(e));
    return rc;
}

}  // extern "C"
