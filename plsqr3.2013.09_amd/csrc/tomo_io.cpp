// Synthetic straight-ray tomography systems and the reference's file formats.
//
// Host-only C++ (no HIP calls), linked into libplsqr_hip.so.  Everything here
// is deterministic for a given seed, independent of the number of threads:
// every ray, every noise sample and every blob draws from its own counter-based
// random stream.
//
// Formats follow the reference loaders and tools:
//   kernel .bin/.info   Ker2PLSQR3.c:505-511,595-606  read by plsqr3.c:1379-1409,1515-1626
//   damping row/col     damping_binary.py:84-289,523-551 read by plsqr3.c:1008-1194,503-528
//   b                   read with "%le" at plsqr3.c:580-589
//   x                   mpiio_write_result plsqr3.c:1986-2013
#include "plsqr_hip.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------------------
// Counter-based random streams (SplitMix64 finaliser over a keyed counter).
// ---------------------------------------------------------------------------
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t state;
    Stream(uint64_t seed, uint64_t stream, uint64_t index) {
        state = mix64(seed + 0x9E3779B97F4A7C15ULL * (stream + 1)) ^
                mix64(index * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL);
    }
    uint64_t next() {
        state += 0x9E3779B97F4A7C15ULL;
        return mix64(state);
    }
    // uniform in [0, 1)
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform in (0, 1]
    double uniform_open0() { return ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller, one sample per call
        const double u1 = uniform_open0();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
    }
    int below(int k) { return (int)((next() >> 33) % (uint64_t)k); }
};

enum : uint64_t { STREAM_RAY = 1, STREAM_NOISE = 2, STREAM_BLOB = 3 };

template <class F>
void parallel_for(long long n, F &&f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("PLSQR_HOST_THREADS")) nt = (unsigned)std::max(1, atoi(e));
    if (nt == 0) nt = 1;
    nt = std::min<unsigned>(nt, 64);
    if (n < 4096 || nt == 1) {
        f(0LL, n, 0u);
        return;
    }
    std::vector<std::thread> th;
    const long long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const long long b = (long long)t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&f, b, e, t] { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}

unsigned n_workers(long long n) {
    unsigned nt = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("PLSQR_HOST_THREADS")) nt = (unsigned)std::max(1, atoi(e));
    if (nt == 0) nt = 1;
    nt = std::min<unsigned>(nt, 64);
    if (n < 4096) nt = 1;
    return nt;
}

// ---------------------------------------------------------------------------
// Ray geometry and Amanatides-Woo voxel traversal.
// ---------------------------------------------------------------------------
struct Grid {
    int n[3];  // nx, ny, nz
};

void ray_endpoints(const Grid &g, int geometry, Stream &s, double p0[3], double p1[3]) {
    const double N[3] = {(double)g.n[0], (double)g.n[1], (double)g.n[2]};
    if (geometry == PLSQR_GEOM_OPPOSITE_FACES) {
        const int a = s.below(3);
        for (int k = 0; k < 3; ++k) {
            p0[k] = s.uniform() * N[k];
            p1[k] = s.uniform() * N[k];
        }
        p0[a] = 0.0;
        p1[a] = N[a];
    } else {
        // source on the top face z = nz
        p0[0] = s.uniform() * N[0];
        p0[1] = s.uniform() * N[1];
        p0[2] = N[2];
        // receiver on one of the other five faces: z=0, x=0, x=nx, y=0, y=ny
        const int f = s.below(5);
        for (int k = 0; k < 3; ++k) p1[k] = s.uniform() * N[k];
        switch (f) {
            case 0: p1[2] = 0.0; break;
            case 1: p1[0] = 0.0; break;
            case 2: p1[0] = N[0]; break;
            case 3: p1[1] = 0.0; break;
            default: p1[1] = N[1]; break;
        }
    }
}

// Calls emit(col, length) for every voxel the segment p0->p1 crosses with a
// positive length, in traversal order.
template <class Emit>
void traverse(const Grid &g, const double p0[3], const double p1[3], Emit &&emit) {
    double d[3];
    int iv[3], step[3];
    double tmax[3];
    double L2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        d[k] = p1[k] - p0[k];
        L2 += d[k] * d[k];
    }
    const double L = std::sqrt(L2);
    if (!(L > 0.0)) return;
    for (int k = 0; k < 3; ++k) {
        int i = (int)std::floor(p0[k]);
        if (i < 0) i = 0;
        if (i > g.n[k] - 1) i = g.n[k] - 1;
        iv[k] = i;
        if (d[k] > 0.0) {
            step[k] = 1;
            tmax[k] = ((double)(i + 1) - p0[k]) / d[k];
        } else if (d[k] < 0.0) {
            step[k] = -1;
            tmax[k] = (p0[k] - (double)i) / (-d[k]);
        } else {
            step[k] = 0;
            tmax[k] = INFINITY;
        }
    }
    double t = 0.0;
    for (;;) {
        int a = 0;
        if (tmax[1] < tmax[a]) a = 1;
        if (tmax[2] < tmax[a]) a = 2;
        const double tn = tmax[a] < 1.0 ? tmax[a] : 1.0;
        const double seg = (tn - t) * L;
        if (seg > 0.0) emit(iv[2] * g.n[1] * g.n[0] + iv[1] * g.n[0] + iv[0], seg);
        if (tn >= 1.0) break;
        t = tn;
        iv[a] += step[a];
        if (iv[a] < 0 || iv[a] >= g.n[a]) break;
        const double boundary = step[a] > 0 ? (double)(iv[a] + 1) : (double)iv[a];
        tmax[a] = step[a] > 0 ? (boundary - p0[a]) / d[a] : (p0[a] - boundary) / (-d[a]);
    }
}

// x_true: sum of Gaussian blobs sampled at voxel centres.
void make_blobs(const Grid &g, uint64_t seed, int n_blobs, double width, double *x) {
    const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
    std::vector<double> c(3 * (size_t)std::max(n_blobs, 0)), amp((size_t)std::max(n_blobs, 0));
    for (int k = 0; k < n_blobs; ++k) {
        Stream s(seed, STREAM_BLOB, (uint64_t)k);
        for (int a = 0; a < 3; ++a) c[3 * k + a] = s.uniform() * g.n[a];
        amp[k] = 2.0 * s.uniform() - 1.0;
    }
    const double inv2w2 = 1.0 / (2.0 * width * width);
    parallel_for(n, [&](long long b, long long e, unsigned) {
        for (long long j = b; j < e; ++j) {
            const int ix = (int)(j % g.n[0]);
            const int iy = (int)((j / g.n[0]) % g.n[1]);
            const int iz = (int)(j / ((long long)g.n[0] * g.n[1]));
            const double p[3] = {ix + 0.5, iy + 0.5, iz + 0.5};
            double v = 0.0;
            for (int k = 0; k < n_blobs; ++k) {
                double r2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double dd = p[a] - c[3 * k + a];
                    r2 += dd * dd;
                }
                v += amp[k] * std::exp(-r2 * inv2w2);
            }
            x[j] = v;
        }
    });
}

// b[r] = sum over row r (ascending column) of val * x + noise
void kernel_rhs(const int m_k, const long long *row_ptr, const int *col, const double *val,
                const double *x, uint64_t seed, double sigma, double *b) {
    parallel_for(m_k, [&](long long lo, long long hi, unsigned) {
        for (long long r = lo; r < hi; ++r) {
            double s = 0.0;
            for (long long e = row_ptr[r]; e < row_ptr[r + 1]; ++e) s += val[e] * x[col[e]];
            Stream st(seed, STREAM_NOISE, (uint64_t)r);
            b[r] = s + sigma * st.normal();
        }
    });
}

// Damping matrix in the reference tool's row order
// (damping_binary.py:84-479): parameter blocks one after the other (block p
// holds columns p*ncell.., :93-95), inside a block the cells x fastest
// (:89-92), and per cell the identity row, then the second differences
// along x, y and z (1,-2,1) and, with cross rows, the mixed differences in
// the xy, xz and yz planes (0.5,-0.5,-0.5,0.5 at the corners of a 3 x 3
// square).  w = {identity, xx, yy, zz, xy, xz, yz}.  Columns inside a row
// ascend.  The column layout holds the same FP64 values (the tool narrows
// its column copy to FP32, :70, which only differs for weights FP32 cannot
// represent).
int build_damping(int nx, int ny, int nz, int n_param, const double w[7], bool cross, plsqr_system *s) {
    const long long ncell = (long long)nx * ny * nz;
    const long long ex = std::max(nx - 2, 0), ey = std::max(ny - 2, 0), ez = std::max(nz - 2, 0);
    const long long rxx = ex * ny * nz, ryy = nx * ey * nz, rzz = (long long)nx * ny * ez;
    const long long rxy = cross ? ex * ey * nz : 0, rxz = cross ? ex * ny * ez : 0, ryz = cross ? nx * ey * ez : 0;
    const long long m_d = (ncell + rxx + ryy + rzz + rxy + rxz + ryz) * n_param;  // :50
    const long long nnz_d = (ncell + 3 * (rxx + ryy + rzz) + 4 * (rxy + rxz + ryz)) * n_param;  // :61
    const long long n = ncell * n_param;
    if (m_d >= (1LL << 31) || nnz_d >= (1LL << 31) || n >= (1LL << 31))
        return fail(PLSQR_ERR_ARG, "damping matrix too large for int32 indices");
    s->m_d = (int)m_d;
    s->nnz_d = nnz_d;
    s->d_row_ptr = (int *)malloc(sizeof(int) * (m_d + 1));
    s->d_col_idx = (int *)malloc(sizeof(int) * std::max(nnz_d, 1LL));
    s->d_val = (double *)malloc(sizeof(double) * std::max(nnz_d, 1LL));
    s->dc_col_ptr = (int *)calloc((size_t)n + 1, sizeof(int));
    s->dc_row_idx = (int *)malloc(sizeof(int) * std::max(nnz_d, 1LL));
    s->dc_val = (double *)malloc(sizeof(double) * std::max(nnz_d, 1LL));
    if (!s->d_row_ptr || !s->d_col_idx || !s->d_val || !s->dc_col_ptr || !s->dc_row_idx || !s->dc_val)
        return fail(PLSQR_ERR_NOMEM, "host allocation of the damping matrix failed");
    long long row = 0, e = 0;
    s->d_row_ptr[0] = 0;
    for (int p = 0; p < n_param; ++p) {
        const long long off = (long long)p * ncell;
        auto col = [&](int ix, int iy, int iz) { return (int)(off + (long long)iz * ny * nx + (long long)iy * nx + ix); };
        auto put = [&](int c, double v) {
            s->d_col_idx[e] = c;
            s->d_val[e++] = v;
        };
        auto end_row = [&] { s->d_row_ptr[++row] = (int)e; };
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    put(col(ix, iy, iz), 1.0 * w[0]);
                    end_row();
                    if (ix < nx - 2) {
                        put(col(ix, iy, iz), 1.0 * w[1]);
                        put(col(ix + 1, iy, iz), -2.0 * w[1]);
                        put(col(ix + 2, iy, iz), 1.0 * w[1]);
                        end_row();
                    }
                    if (iy < ny - 2) {
                        put(col(ix, iy, iz), 1.0 * w[2]);
                        put(col(ix, iy + 1, iz), -2.0 * w[2]);
                        put(col(ix, iy + 2, iz), 1.0 * w[2]);
                        end_row();
                    }
                    if (iz < nz - 2) {
                        put(col(ix, iy, iz), 1.0 * w[3]);
                        put(col(ix, iy, iz + 1), -2.0 * w[3]);
                        put(col(ix, iy, iz + 2), 1.0 * w[3]);
                        end_row();
                    }
                    if (!cross) continue;
                    if (ix < nx - 2 && iy < ny - 2) {
                        put(col(ix, iy, iz), 0.5 * w[4]);
                        put(col(ix + 2, iy, iz), -0.5 * w[4]);
                        put(col(ix, iy + 2, iz), -0.5 * w[4]);
                        put(col(ix + 2, iy + 2, iz), 0.5 * w[4]);
                        end_row();
                    }
                    if (ix < nx - 2 && iz < nz - 2) {
                        put(col(ix, iy, iz), 0.5 * w[5]);
                        put(col(ix + 2, iy, iz), -0.5 * w[5]);
                        put(col(ix, iy, iz + 2), -0.5 * w[5]);
                        put(col(ix + 2, iy, iz + 2), 0.5 * w[5]);
                        end_row();
                    }
                    if (iy < ny - 2 && iz < nz - 2) {
                        put(col(ix, iy, iz), 0.5 * w[6]);
                        put(col(ix, iy + 2, iz), -0.5 * w[6]);
                        put(col(ix, iy, iz + 2), -0.5 * w[6]);
                        put(col(ix, iy + 2, iz + 2), 0.5 * w[6]);
                        end_row();
                    }
                }
    }
    // column layout: stable by column so rows ascend inside a column
    // (damping_binary.py:523 sorts the row-ordered triplets by column)
    for (long long k = 0; k < nnz_d; ++k) s->dc_col_ptr[s->d_col_idx[k] + 1]++;
    for (long long c = 0; c < n; ++c) s->dc_col_ptr[c + 1] += s->dc_col_ptr[c];
    std::vector<int> fill(s->dc_col_ptr, s->dc_col_ptr + n);
    for (long long r = 0; r < m_d; ++r)
        for (int k = s->d_row_ptr[r]; k < s->d_row_ptr[r + 1]; ++k) {
            const int c = s->d_col_idx[k];
            const int pos = fill[(size_t)c]++;
            s->dc_row_idx[pos] = (int)r;
            s->dc_val[pos] = s->d_val[k];
        }
    return PLSQR_OK;
}

// CSR (row_ptr as long long) -> CSC with rows ascending inside each column.
// Each worker owns a contiguous row range, so a per-worker column histogram
// gives every worker its own write offsets and the result equals a serial
// row-ordered scatter.
void csr_to_csc(int m, int n, const long long *row_ptr, const int *col, const double *val,
                int *col_ptr, int *row_idx, double *cval) {
    const unsigned nt = n_workers(m);
    std::vector<std::vector<int>> cnt(nt, std::vector<int>((size_t)n, 0));
    std::vector<long long> rb(nt + 1);
    const long long chunk = ((long long)m + nt - 1) / nt;
    for (unsigned t = 0; t <= nt; ++t) rb[t] = std::min<long long>(m, (long long)t * chunk);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                auto &c = cnt[t];
                for (long long r = rb[t]; r < rb[t + 1]; ++r)
                    for (long long e = row_ptr[r]; e < row_ptr[r + 1]; ++e) c[col[e]]++;
            });
        for (auto &x : th) x.join();
    }
    col_ptr[0] = 0;
    for (int c = 0; c < n; ++c) {
        long long s = 0;
        for (unsigned t = 0; t < nt; ++t) s += cnt[t][c];
        col_ptr[c + 1] = (int)(col_ptr[c] + s);
    }
    // turn counts into per-worker start offsets
    for (int c = 0; c < n; ++c) {
        int off = col_ptr[c];
        for (unsigned t = 0; t < nt; ++t) {
            const int k = cnt[t][c];
            cnt[t][c] = off;
            off += k;
        }
    }
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                auto &o = cnt[t];
                for (long long r = rb[t]; r < rb[t + 1]; ++r)
                    for (long long e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
                        const int p = o[col[e]]++;
                        row_idx[p] = (int)r;
                        cval[p] = val[e];
                    }
            });
        for (auto &x : th) x.join();
    }
}

// CSC -> row_ptr/col/val (columns ascending inside each row)
void csc_to_csr(int m, int n, const int *col_ptr, const int *row_idx, const double *val,
                std::vector<long long> &row_ptr, std::vector<int> &col, std::vector<double> &rval) {
    const long long nnz = col_ptr[n];
    row_ptr.assign((size_t)m + 1, 0);
    for (long long e = 0; e < nnz; ++e) row_ptr[row_idx[e] + 1]++;
    for (int r = 0; r < m; ++r) row_ptr[r + 1] += row_ptr[r];
    col.resize((size_t)nnz);
    rval.resize((size_t)nnz);
    std::vector<long long> fill(row_ptr.begin(), row_ptr.end() - 1);
    for (int c = 0; c < n; ++c)
        for (int e = col_ptr[c]; e < col_ptr[c + 1]; ++e) {
            const long long p = fill[row_idx[e]]++;
            col[p] = c;
            rval[p] = val[e];
        }
}

struct FileCloser {
    FILE *f;
    ~FileCloser() {
        if (f) fclose(f);
    }
};

std::string join(const char *dir, const char *name) {
    std::string s = dir ? dir : ".";
    if (!s.empty() && s.back() != '/') s += '/';
    return s + name;
}

#pragma pack(push, 1)
struct Record {  // NONZERO in main.h:112-116: {int r, int c, double val}
    int32_t r, c;
    double v;
};
#pragma pack(pop)
static_assert(sizeof(Record) == 16, "16-byte records");

int write_records(const std::string &path, long long count,
                  const std::function<void(long long, Record &)> &get) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return fail(PLSQR_ERR_IO, "cannot open " + path + ": " + strerror(errno));
    FileCloser fc{f};
    const long long B = 1 << 16;
    std::vector<Record> buf((size_t)std::min(B, std::max(count, 1LL)));
    for (long long base = 0; base < count; base += B) {
        const long long k = std::min(B, count - base);
        for (long long i = 0; i < k; ++i) get(base + i, buf[(size_t)i]);
        if (fwrite(buf.data(), sizeof(Record), (size_t)k, f) != (size_t)k)
            return fail(PLSQR_ERR_IO, "short write on " + path);
    }
    return PLSQR_OK;
}

int read_records(const std::string &path, long long offset, long long count, std::vector<Record> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return fail(PLSQR_ERR_IO, "cannot open " + path + ": " + strerror(errno));
    FileCloser fc{f};
    if (fseeko(f, (off_t)offset * (off_t)sizeof(Record), SEEK_SET) != 0)
        return fail(PLSQR_ERR_IO, "seek failed on " + path);
    out.resize((size_t)count);
    if (count > 0 && fread(out.data(), sizeof(Record), (size_t)count, f) != (size_t)count)
        return fail(PLSQR_ERR_FORMAT, "file " + path + " holds fewer records than the info file declares");
    return PLSQR_OK;
}

template <class T>
T *xmalloc(size_t n) {
    return (T *)malloc(sizeof(T) * std::max<size_t>(n, 1));
}

}  // namespace

extern "C" {

const char *plsqr_last_global_error(void) { return g_err.c_str(); }

void plsqr_system_free(plsqr_system *s) {
    if (!s) return;
    if (s->owned) {
        free(s->k_col_ptr); free(s->k_row_idx); free(s->k_val);
        free(s->d_row_ptr); free(s->d_col_idx); free(s->d_val);
        free(s->dc_col_ptr); free(s->dc_row_idx); free(s->dc_val);
        free(s->b); free(s->x_true);
    }
    memset(s, 0, sizeof(*s));
}

int plsqr_gen_system(const plsqr_gen_params *p, plsqr_system *out) {
    if (!p || !out) return fail(PLSQR_ERR_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(PLSQR_ERR_ARG, "grid edges must be >= 3");
    const long long ncell = (long long)p->nx * p->ny * p->nz;
    if (ncell >= (1LL << 31)) return fail(PLSQR_ERR_ARG, "grid too large for int32 columns");
    if (p->n_rays <= 0 || p->n_rays >= (1LL << 31)) return fail(PLSQR_ERR_ARG, "n_rays out of range");
    if (p->geometry != PLSQR_GEOM_OPPOSITE_FACES && p->geometry != PLSQR_GEOM_TOP_SOURCES)
        return fail(PLSQR_ERR_ARG, "unknown geometry");
    Grid g{{p->nx, p->ny, p->nz}};
    const int m_k = (int)p->n_rays, n = (int)ncell;

    // pass 1: nonzeros per ray
    std::vector<long long> row_ptr((size_t)m_k + 1, 0);
    parallel_for(m_k, [&](long long b, long long e, unsigned) {
        for (long long r = b; r < e; ++r) {
            Stream s(p->seed, STREAM_RAY, (uint64_t)r);
            double p0[3], p1[3];
            ray_endpoints(g, p->geometry, s, p0, p1);
            long long k = 0;
            traverse(g, p0, p1, [&](int, double) { ++k; });
            row_ptr[(size_t)r + 1] = k;
        }
    });
    for (int r = 0; r < m_k; ++r) row_ptr[r + 1] += row_ptr[r];
    const long long nnz = row_ptr[m_k];
    if (nnz >= (1LL << 31)) return fail(PLSQR_ERR_ARG, "kernel nnz exceeds int32 column pointers");

    // pass 2: fill rows, then sort each row by column
    std::vector<int> col((size_t)nnz);
    std::vector<double> val((size_t)nnz);
    parallel_for(m_k, [&](long long b, long long e, unsigned) {
        std::vector<std::pair<int, double>> tmp;
        for (long long r = b; r < e; ++r) {
            Stream s(p->seed, STREAM_RAY, (uint64_t)r);
            double p0[3], p1[3];
            ray_endpoints(g, p->geometry, s, p0, p1);
            tmp.clear();
            traverse(g, p0, p1, [&](int c, double v) { tmp.emplace_back(c, v); });
            std::sort(tmp.begin(), tmp.end(),
                      [](const std::pair<int, double> &a, const std::pair<int, double> &b2) { return a.first < b2.first; });
            long long o = row_ptr[(size_t)r];
            for (auto &cv : tmp) {
                col[(size_t)o] = cv.first;
                val[(size_t)o++] = cv.second;
            }
        }
    });

    out->owned = 1;
    out->m_k = m_k;
    out->n = n;
    out->nnz_k = nnz;
    out->k_col_ptr = xmalloc<int>((size_t)n + 1);
    out->k_row_idx = xmalloc<int>((size_t)nnz);
    out->k_val = xmalloc<double>((size_t)nnz);
    out->b = xmalloc<double>((size_t)m_k);
    out->x_true = xmalloc<double>((size_t)n);
    if (!out->k_col_ptr || !out->k_row_idx || !out->k_val || !out->b || !out->x_true) {
        plsqr_system_free(out);
        return fail(PLSQR_ERR_NOMEM, "host allocation failed");
    }
    csr_to_csc(m_k, n, row_ptr.data(), col.data(), val.data(), out->k_col_ptr, out->k_row_idx, out->k_val);
    make_blobs(g, p->seed, p->n_blobs, p->blob_width, out->x_true);
    kernel_rhs(m_k, row_ptr.data(), col.data(), val.data(), out->x_true, p->seed, p->sigma, out->b);
    const double w[7] = {p->lambda, p->w_xx, p->w_yy, p->w_zz, 0.0, 0.0, 0.0};
    const int rc = build_damping(p->nx, p->ny, p->nz, 1, w, false, out);
    if (rc) plsqr_system_free(out);
    return rc;
}

int plsqr_gen_field(const plsqr_gen_params *p, double *x_true) {
    if (!p || !x_true) return fail(PLSQR_ERR_ARG, "null argument");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(PLSQR_ERR_ARG, "grid must be positive");
    Grid g{{p->nx, p->ny, p->nz}};
    make_blobs(g, p->seed, p->n_blobs, p->blob_width, x_true);
    return PLSQR_OK;
}

int plsqr_gen_rhs(const plsqr_system *sys, const plsqr_gen_params *p, double *b, double *x_true) {
    if (!sys || !p || !b || !x_true) return fail(PLSQR_ERR_ARG, "null argument");
    if ((long long)p->nx * p->ny * p->nz != sys->n) return fail(PLSQR_ERR_ARG, "grid does not match the system's columns");
    Grid g{{p->nx, p->ny, p->nz}};
    make_blobs(g, p->seed, p->n_blobs, p->blob_width, x_true);
    // b[r] = sum over ascending columns of val * x + noise, as kernel_rhs
    // forms it from the CSR; each host thread owns a range of rows and walks
    // the columns in order, so no CSR copy is needed
    const int *cp = sys->k_col_ptr, *ri = sys->k_row_idx;
    const double *kv = sys->k_val;
    parallel_for(sys->m_k, [&](long long lo, long long hi, unsigned) {
        std::vector<double> acc((size_t)(hi - lo), 0.0);
        for (int c = 0; c < sys->n; ++c) {
            const int *rb = ri + cp[c], *re = ri + cp[c + 1];
            const double xc = x_true[c];
            for (const int *q = std::lower_bound(rb, re, (int)lo); q != re && *q < hi; ++q)
                acc[(size_t)(*q - lo)] += kv[q - ri] * xc;
        }
        for (long long r = lo; r < hi; ++r) {
            Stream st(p->seed, STREAM_NOISE, (uint64_t)r);
            b[r] = acc[(size_t)(r - lo)] + p->sigma * st.normal();
        }
    });
    return PLSQR_OK;
}

int plsqr_write_reference_files(const plsqr_system *s, const char *dir, const char *ker_f,
                                const char *ker_i, const char *damp_f, const char *damp_i,
                                const char *damp_f_bycol, const char *damp_i_bycol, const char *b_k) {
    if (!s) return fail(PLSQR_ERR_ARG, "null system");
    int rc;
    if (ker_f) {
        // column-major records; the column of entry e is found by walking col_ptr
        std::vector<int> colof((size_t)s->nnz_k);
        for (int c = 0; c < s->n; ++c)
            for (int e = s->k_col_ptr[c]; e < s->k_col_ptr[c + 1]; ++e) colof[(size_t)e] = c;
        rc = write_records(join(dir, ker_f), s->nnz_k, [&](long long e, Record &R) {
            R.r = s->k_row_idx[e] + 1;
            R.c = colof[(size_t)e] + 1;
            R.v = s->k_val[e];
        });
        if (rc) return rc;
    }
    if (ker_i) {
        const std::string path = join(dir, ker_i);
        FILE *f = fopen(path.c_str(), "w");
        if (!f) return fail(PLSQR_ERR_IO, "cannot open " + path);
        FileCloser fc{f};
        long long start1 = 1;
        for (int c = 0; c < s->n; ++c) {
            const int k = s->k_col_ptr[c + 1] - s->k_col_ptr[c];
            fprintf(f, "%d %d %lld\n", c + 1, k, k == 0 ? 0LL : start1);
            start1 += k;
        }
    }
    if (damp_f) {
        std::vector<int> rowof((size_t)s->nnz_d);
        for (int r = 0; r < s->m_d; ++r)
            for (int e = s->d_row_ptr[r]; e < s->d_row_ptr[r + 1]; ++e) rowof[(size_t)e] = r;
        rc = write_records(join(dir, damp_f), s->nnz_d, [&](long long e, Record &R) {
            R.r = rowof[(size_t)e] + 1;
            R.c = s->d_col_idx[e] + 1;
            R.v = s->d_val[e];
        });
        if (rc) return rc;
    }
    if (damp_i) {
        const std::string path = join(dir, damp_i);
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) return fail(PLSQR_ERR_IO, "cannot open " + path);
        FileCloser fc{f};
        for (int r = 0; r < s->m_d; ++r) {
            const int32_t k = s->d_row_ptr[r + 1] - s->d_row_ptr[r];
            fwrite(&k, 4, 1, f);
        }
    }
    if (damp_f_bycol) {
        std::vector<int> colof((size_t)s->nnz_d);
        for (int c = 0; c < s->n; ++c)
            for (int e = s->dc_col_ptr[c]; e < s->dc_col_ptr[c + 1]; ++e) colof[(size_t)e] = c;
        rc = write_records(join(dir, damp_f_bycol), s->nnz_d, [&](long long e, Record &R) {
            R.r = s->dc_row_idx[e] + 1;
            R.c = colof[(size_t)e] + 1;
            R.v = s->dc_val[e];
        });
        if (rc) return rc;
    }
    if (damp_i_bycol) {
        const std::string path = join(dir, damp_i_bycol);
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) return fail(PLSQR_ERR_IO, "cannot open " + path);
        FileCloser fc{f};
        for (int c = 0; c < s->n; ++c) {
            const int32_t k = s->dc_col_ptr[c + 1] - s->dc_col_ptr[c];
            fwrite(&k, 4, 1, f);
        }
    }
    if (b_k) {
        const std::string path = join(dir, b_k);
        FILE *f = fopen(path.c_str(), "w");
        if (!f) return fail(PLSQR_ERR_IO, "cannot open " + path);
        FileCloser fc{f};
        for (int r = 0; r < s->m_k; ++r) fprintf(f, "%.17e\n", s->b[r]);
    }
    return PLSQR_OK;
}

int plsqr_read_reference_files(const char *dir, const char *ker_f, const char *ker_i,
                               const char *damp_f, const char *damp_i, const char *damp_f_bycol,
                               const char *damp_i_bycol, const char *b_k, int row_k, int row_d,
                               int col, plsqr_system *out) {
    if (!out || !ker_f || !ker_i || !damp_f || !damp_i || !damp_f_bycol || !damp_i_bycol || !b_k)
        return fail(PLSQR_ERR_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    if (row_k <= 0 || row_d < 0 || col <= 0) return fail(PLSQR_ERR_ARG, "sizes must be positive");
    out->owned = 1;
    out->m_k = row_k;
    out->m_d = row_d;
    out->n = col;
    int rc = PLSQR_OK;
    auto bail = [&](int code) {
        plsqr_system_free(out);
        return code;
    };

    // kernel .info (plsqr3.c:1379-1409): "col nnz start", 1-based
    std::vector<long long> begin((size_t)col);
    std::vector<int> cnt((size_t)col);
    {
        const std::string path = join(dir, ker_i);
        FILE *f = fopen(path.c_str(), "r");
        if (!f) return bail(fail(PLSQR_ERR_IO, "cannot open " + path));
        FileCloser fc{f};
        for (int i = 0; i < col; ++i) {
            int ci, k;
            long long st;
            if (fscanf(f, "%d %d %lld", &ci, &k, &st) != 3)
                return bail(fail(PLSQR_ERR_FORMAT, "kernel info file " + path + " ends before column " + std::to_string(i + 1)));
            if (k < 0) return bail(fail(PLSQR_ERR_FORMAT, "negative nnz in kernel info file"));
            cnt[(size_t)i] = k;
            begin[(size_t)i] = st - 1;
        }
    }
    long long nnz_k = 0, first = 0;
    bool found = false;
    for (int i = 0; i < col; ++i) {
        if (!found && cnt[(size_t)i] > 0) {
            first = begin[(size_t)i];
            found = true;
        }
        nnz_k += cnt[(size_t)i];
    }
    if (nnz_k >= (1LL << 31)) return bail(fail(PLSQR_ERR_FORMAT, "kernel nnz exceeds int32"));
    out->nnz_k = nnz_k;
    out->k_col_ptr = xmalloc<int>((size_t)col + 1);
    out->k_row_idx = xmalloc<int>((size_t)nnz_k);
    out->k_val = xmalloc<double>((size_t)nnz_k);
    if (!out->k_col_ptr || !out->k_row_idx || !out->k_val) return bail(fail(PLSQR_ERR_NOMEM, "host allocation failed"));
    {
        // one contiguous read starting at the first non-empty column, as
        // mpiio_load_kernelColBin2CSC_once does (plsqr3.c:1558-1590)
        std::vector<Record> rec;
        if ((rc = read_records(join(dir, ker_f), first, nnz_k, rec))) return bail(rc);
        long long t = 0;
        out->k_col_ptr[0] = 0;
        for (int i = 0; i < col; ++i) {
            out->k_col_ptr[i + 1] = out->k_col_ptr[i] + cnt[(size_t)i];
            for (int j = 0; j < cnt[(size_t)i]; ++j, ++t) {
                const int r0 = rec[(size_t)t].r - 1;
                if (r0 < 0 || r0 >= row_k)
                    return bail(fail(PLSQR_ERR_BOUNDS, "kernel row " + std::to_string(r0 + 1) + " out of range in column " + std::to_string(i + 1)));
                out->k_row_idx[t] = r0;
                out->k_val[t] = rec[(size_t)t].v;
            }
        }
    }

    // damping row info: int32 nnz per row (plsqr3.c:1008-1025)
    long long nnz_d = 0;
    {
        const std::string path = join(dir, damp_i);
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) return bail(fail(PLSQR_ERR_IO, "cannot open " + path));
        FileCloser fc{f};
        std::vector<int32_t> k((size_t)row_d);
        if (row_d > 0 && fread(k.data(), 4, (size_t)row_d, f) != (size_t)row_d)
            return bail(fail(PLSQR_ERR_FORMAT, "damping row info file is short"));
        for (auto v : k) nnz_d += v;
    }
    out->nnz_d = nnz_d;
    out->d_row_ptr = xmalloc<int>((size_t)row_d + 1);
    out->d_col_idx = xmalloc<int>((size_t)nnz_d);
    out->d_val = xmalloc<double>((size_t)nnz_d);
    out->dc_col_ptr = xmalloc<int>((size_t)col + 1);
    out->dc_row_idx = xmalloc<int>((size_t)nnz_d);
    out->dc_val = xmalloc<double>((size_t)nnz_d);
    if (!out->d_row_ptr || !out->d_col_idx || !out->d_val || !out->dc_col_ptr || !out->dc_row_idx || !out->dc_val)
        return bail(fail(PLSQR_ERR_NOMEM, "host allocation failed"));
    {
        // CSR: row pointers from the rows stored in the records (plsqr3.c:1133-1162)
        std::vector<Record> rec;
        if ((rc = read_records(join(dir, damp_f), 0, nnz_d, rec))) return bail(rc);
        std::vector<int> per((size_t)row_d, 0);
        for (long long i = 0; i < nnz_d; ++i) {
            const int r = rec[(size_t)i].r - 1, c = rec[(size_t)i].c - 1;
            if (r < 0 || r >= row_d || c < 0 || c >= col)
                return bail(fail(PLSQR_ERR_BOUNDS, "damping record " + std::to_string(i) + " out of range"));
            if (i > 0 && r < rec[(size_t)i - 1].r - 1)
                return bail(fail(PLSQR_ERR_FORMAT, "damping row file is not sorted by row"));
            per[(size_t)r]++;
            out->d_col_idx[i] = c;
            out->d_val[i] = rec[(size_t)i].v;
        }
        out->d_row_ptr[0] = 0;
        for (int r = 0; r < row_d; ++r) out->d_row_ptr[r + 1] = out->d_row_ptr[r] + per[(size_t)r];
    }
    {
        // column info then CSC (plsqr3.c:494-528, 1163-1189)
        const std::string path = join(dir, damp_i_bycol);
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) return bail(fail(PLSQR_ERR_IO, "cannot open " + path));
        FileCloser fc{f};
        std::vector<int32_t> k((size_t)col);
        if (fread(k.data(), 4, (size_t)col, f) != (size_t)col)
            return bail(fail(PLSQR_ERR_FORMAT, "damping column info file is short"));
        long long tot = 0;
        for (auto v : k) tot += v;
        if (tot != nnz_d) return bail(fail(PLSQR_ERR_FORMAT, "damping row and column files disagree on nnz"));
        std::vector<Record> rec;
        if ((rc = read_records(join(dir, damp_f_bycol), 0, nnz_d, rec))) return bail(rc);
        std::vector<int> per((size_t)col, 0);
        for (long long i = 0; i < nnz_d; ++i) {
            const int r = rec[(size_t)i].r - 1, c = rec[(size_t)i].c - 1;
            if (c < 0 || c >= col || r < 0 || r >= row_d)
                return bail(fail(PLSQR_ERR_BOUNDS, "damping-by-column record " + std::to_string(i) + " out of range"));
            if (i > 0 && c < rec[(size_t)i - 1].c - 1)
                return bail(fail(PLSQR_ERR_FORMAT, "damping column file is not sorted by column"));
            per[(size_t)c]++;
            out->dc_row_idx[i] = r;
            out->dc_val[i] = rec[(size_t)i].v;
        }
        out->dc_col_ptr[0] = 0;
        for (int c = 0; c < col; ++c) out->dc_col_ptr[c + 1] = out->dc_col_ptr[c] + per[(size_t)c];
    }
    {
        const std::string path = join(dir, b_k);
        FILE *f = fopen(path.c_str(), "r");
        if (!f) return bail(fail(PLSQR_ERR_IO, "cannot open " + path));
        FileCloser fc{f};
        out->b = xmalloc<double>((size_t)row_k);
        if (!out->b) return bail(fail(PLSQR_ERR_NOMEM, "host allocation failed"));
        for (int i = 0; i < row_k; ++i)
            if (fscanf(f, "%le", &out->b[i]) != 1) return bail(fail(PLSQR_ERR_FORMAT, "b file holds fewer than row_k values"));
    }
    return PLSQR_OK;
}

int plsqr_gen_damping(const plsqr_damping_spec *spec, plsqr_system *out) {
    if (!spec || !out) return fail(PLSQR_ERR_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    if (spec->nx < 1 || spec->ny < 1 || spec->nz < 1 || spec->n_param < 1)
        return fail(PLSQR_ERR_ARG, "grid edges and n_param must be >= 1");
    for (int k = 0; k < 7; ++k)
        if (!std::isfinite(spec->w[k])) return fail(PLSQR_ERR_NONFINITE, "damping weight is NaN or Inf");
    const long long n = (long long)spec->nx * spec->ny * spec->nz * spec->n_param;
    if (n >= (1LL << 31)) return fail(PLSQR_ERR_ARG, "too many columns for int32 indices");
    out->owned = 1;
    out->n = (int)n;
    const int rc = build_damping(spec->nx, spec->ny, spec->nz, spec->n_param, spec->w, spec->cross != 0, out);
    if (rc) plsqr_system_free(out);
    return rc;
}

int plsqr_write_damping_files(const plsqr_damping_spec *spec, const char *dir, const char *damp_f,
                              const char *damp_i, const char *damp_f_bycol, const char *damp_i_bycol) {
    plsqr_system d;
    int rc = plsqr_gen_damping(spec, &d);
    if (rc) return rc;
    rc = plsqr_write_reference_files(&d, dir, nullptr, nullptr, damp_f, damp_i, damp_f_bycol, damp_i_bycol, nullptr);
    plsqr_system_free(&d);
    return rc;
}

int plsqr_ker_ascii2bin(const char *list_path, const char *in_dir, const char *out_dir, const char *out_list) {
    if (!list_path || !in_dir || !out_dir) return fail(PLSQR_ERR_ARG, "list_path, in_dir and out_dir are required");
    FILE *lf = fopen(list_path, "r");
    if (!lf) return fail(PLSQR_ERR_IO, std::string("cannot open ") + list_path + ": " + strerror(errno));
    std::vector<std::string> names;
    {
        FileCloser fc{lf};
        int count = 0;  // the list: the number of rows, then one file name per row (ker_ascii2bin.c:41,58)
        if (fscanf(lf, "%d", &count) != 1 || count < 0)
            return fail(PLSQR_ERR_FORMAT, std::string(list_path) + ": the first token must be the row count");
        char name[4096];
        for (int i = 0; i < count; ++i) {
            if (fscanf(lf, "%4095s", name) != 1)
                return fail(PLSQR_ERR_FORMAT, std::string(list_path) + ": fewer file names than the row count");
            names.emplace_back(name);
        }
    }
    std::vector<long long> nnz(names.size(), 0);
    for (size_t i = 0; i < names.size(); ++i) {
        const std::string in = join(in_dir, names[i].c_str()), out = join(out_dir, (names[i] + ".bin").c_str());
        FILE *fi = fopen(in.c_str(), "r");
        if (!fi) return fail(PLSQR_ERR_IO, "cannot open " + in + ": " + strerror(errno));
        FileCloser ci{fi};
        FILE *fo = fopen(out.c_str(), "wb");
        if (!fo) return fail(PLSQR_ERR_IO, "cannot open " + out + ": " + strerror(errno));
        FileCloser co{fo};
        // "col ix iy iz value" per entry (ker_ascii2bin.c:76-77) -> {ix, col, value} (:81-83)
        int col, ix, iy, iz, k;
        double v;
        std::vector<Record> buf;
        buf.reserve(1 << 14);
        auto flush = [&]() -> bool {
            const bool ok = buf.empty() || fwrite(buf.data(), sizeof(Record), buf.size(), fo) == buf.size();
            buf.clear();
            return ok;
        };
        while ((k = fscanf(fi, "%d %d %d %d %lf", &col, &ix, &iy, &iz, &v)) == 5) {
            buf.push_back(Record{ix, col, v});
            ++nnz[i];
            if (buf.size() == (1 << 14) && !flush()) return fail(PLSQR_ERR_IO, "short write on " + out);
        }
        if (k != EOF)
            return fail(PLSQR_ERR_FORMAT, in + ": entry " + std::to_string(nnz[i] + 1) +
                                              " is not \"col ix iy iz value\"");
        if (!flush()) return fail(PLSQR_ERR_IO, "short write on " + out);
    }
    if (out_list) {
        const std::string path = join(out_dir, out_list);
        FILE *f = fopen(path.c_str(), "w");
        if (!f) return fail(PLSQR_ERR_IO, "cannot open " + path + ": " + strerror(errno));
        FileCloser fc{f};
        for (size_t i = 0; i < names.size(); ++i) fprintf(f, "%s.bin %lld\n", names[i].c_str(), nnz[i]);
    }
    return PLSQR_OK;
}

int plsqr_write_x(const char *path, const double *x, int n) {
    if (!path || (!x && n > 0) || n < 0) return fail(PLSQR_ERR_ARG, "bad argument");
    FILE *f = fopen(path, "wb");
    if (!f) return fail(PLSQR_ERR_IO, std::string("cannot open ") + path);
    FileCloser fc{f};
    if (fwrite(x, sizeof(double), (size_t)n, f) != (size_t)n) return fail(PLSQR_ERR_IO, "short write");
    return PLSQR_OK;
}

int plsqr_read_x(const char *path, double *x, int n) {
    if (!path || (!x && n > 0) || n < 0) return fail(PLSQR_ERR_ARG, "bad argument");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(PLSQR_ERR_IO, std::string("cannot open ") + path);
    FileCloser fc{f};
    if (fread(x, sizeof(double), (size_t)n, f) != (size_t)n) return fail(PLSQR_ERR_FORMAT, "short read");
    return PLSQR_OK;
}

}  // extern "C"

namespace plsqr {
int global_fail(int code, const char *msg) { return fail(code, msg); }
}  // namespace plsqr
