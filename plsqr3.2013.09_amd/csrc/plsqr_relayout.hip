// Kernel re-layout on the GPU: the kernel as rows (one record list per ray,
// the way ray-tracing tools write it) to the column-sorted form PLSQR3
// loads.  Counterpart of PLSQR3_tools/kernel_format/Ker2PLSQR3.c:358-415
// (rows read in order, entries distributed to their columns) and :595-606
// (column-major 16-byte records, rows ascending within a column), done as a
// stable radix sort of the entries by column on the device.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "global_error.h"
#include "plsqr_hip.h"

namespace {

// device allocation released with the scope
struct DevMem {
    void *p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

template <class T>
int dev_alloc(DevMem &m, size_t count, T **out) {
    if (count == 0) count = 1;
    const hipError_t e = hipMalloc(&m.p, sizeof(T) * count);
    if (e != hipSuccess) {
        m.p = nullptr;
        return plsqr::global_fail(PLSQR_ERR_NOMEM, (std::string("hipMalloc: ") + hipGetErrorString(e)).c_str());
    }
    *out = static_cast<T *>(m.p);
    return PLSQR_OK;
}

#define RL_TRY(expr)                                                                               \
    do {                                                                                           \
        const hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                      \
            return plsqr::global_fail(PLSQR_ERR_HIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

// row of every entry from the row pointers
__global__ void kr_row_of(int m, const int *__restrict__ rp, int *__restrict__ row_of) {
    for (int r = blockIdx.x * 256 + threadIdx.x; r < m; r += gridDim.x * 256)
        for (int e = rp[r]; e < rp[r + 1]; ++e) row_of[e] = r;
}

// key = 0-based column, value = entry index
__global__ void kr_keys(int nnz, const int *__restrict__ col1, int *__restrict__ key, int *__restrict__ idx) {
    for (int e = blockIdx.x * 256 + threadIdx.x; e < nnz; e += gridDim.x * 256) {
        key[e] = col1[e] - 1;
        idx[e] = e;
    }
}

__global__ void kr_gather(int nnz, const int *__restrict__ perm, const int *__restrict__ row_of,
                          const double *__restrict__ val, int *__restrict__ ri, double *__restrict__ vo) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nnz; i += gridDim.x * 256) {
        const int e = perm[i];
        ri[i] = row_of[e];
        vo[i] = val[e];
    }
}

// col_ptr[c] = the first sorted entry whose column is >= c, c = 0..n
__global__ void kr_col_ptr(int n, int nnz, const int *__restrict__ skey, int *__restrict__ col_ptr) {
    for (int c = blockIdx.x * 256 + threadIdx.x; c <= n; c += gridDim.x * 256) {
        int lo = 0, hi = nnz;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (skey[mid] < c)
                lo = mid + 1;
            else
                hi = mid;
        }
        col_ptr[c] = lo;
    }
}

int grid_of(long long work) {
    long long g = (work + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

int bits_of(int v) {
    int b = 1;
    while ((1LL << b) <= v) ++b;
    return b;
}

}  // namespace

extern "C" {

int plsqr_kernel_rows_to_csc(int m_k, int n, const int *row_ptr, const int *col1, const double *val, int *col_ptr,
                             int *row_idx, double *val_out) {
    if (m_k <= 0 || n <= 0 || !row_ptr || !col_ptr) return plsqr::global_fail(PLSQR_ERR_ARG, "m_k, n and the pointer arrays are required");
    if (row_ptr[0] != 0) return plsqr::global_fail(PLSQR_ERR_FORMAT, "row_ptr[0] != 0");
    for (int r = 0; r < m_k; ++r)
        if (row_ptr[r + 1] < row_ptr[r])
            return plsqr::global_fail(PLSQR_ERR_FORMAT, ("row_ptr decreases at row " + std::to_string(r)).c_str());
    const int nnz = row_ptr[m_k];
    if (nnz > 0 && (!col1 || !val || !row_idx || !val_out))
        return plsqr::global_fail(PLSQR_ERR_ARG, "entry arrays are required");
    for (int e = 0; e < nnz; ++e) {
        if (col1[e] < 1 || col1[e] > n)
            return plsqr::global_fail(PLSQR_ERR_BOUNDS, ("column out of range at entry " + std::to_string(e)).c_str());
        if (!std::isfinite(val[e]))
            return plsqr::global_fail(PLSQR_ERR_NONFINITE, ("value not finite at entry " + std::to_string(e)).c_str());
    }
    DevMem m_rp, m_c, m_v, m_row, m_k_in, m_k_out, m_i_in, m_i_out, m_ri, m_vo, m_cp, m_tmp;
    int *d_rp, *d_c, *d_row, *d_kin, *d_kout, *d_iin, *d_iout, *d_ri, *d_cp;
    double *d_v, *d_vo;
    int rc;
    if ((rc = dev_alloc(m_rp, (size_t)m_k + 1, &d_rp)) || (rc = dev_alloc(m_c, (size_t)nnz, &d_c)) ||
        (rc = dev_alloc(m_v, (size_t)nnz, &d_v)) || (rc = dev_alloc(m_row, (size_t)nnz, &d_row)) ||
        (rc = dev_alloc(m_k_in, (size_t)nnz, &d_kin)) || (rc = dev_alloc(m_k_out, (size_t)nnz, &d_kout)) ||
        (rc = dev_alloc(m_i_in, (size_t)nnz, &d_iin)) || (rc = dev_alloc(m_i_out, (size_t)nnz, &d_iout)) ||
        (rc = dev_alloc(m_ri, (size_t)nnz, &d_ri)) || (rc = dev_alloc(m_vo, (size_t)nnz, &d_vo)) ||
        (rc = dev_alloc(m_cp, (size_t)n + 1, &d_cp)))
        return rc;
    RL_TRY(hipMemcpy(d_rp, row_ptr, sizeof(int) * ((size_t)m_k + 1), hipMemcpyHostToDevice));
    if (nnz > 0) {
        RL_TRY(hipMemcpy(d_c, col1, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
        RL_TRY(hipMemcpy(d_v, val, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(kr_row_of, dim3(grid_of(m_k)), dim3(256), 0, 0, m_k, d_rp, d_row);
        hipLaunchKernelGGL(kr_keys, dim3(grid_of(nnz)), dim3(256), 0, 0, nnz, d_c, d_kin, d_iin);
        RL_TRY(hipGetLastError());
        // stable: entries of one column keep the row order they arrived in
        size_t tmp_bytes = 0;
        RL_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_kin, d_kout, d_iin, d_iout, nnz, 0,
                                                  bits_of(n)));
        char *d_tmp;
        if ((rc = dev_alloc(m_tmp, tmp_bytes, &d_tmp))) return rc;
        RL_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_bytes, d_kin, d_kout, d_iin, d_iout, nnz, 0,
                                                  bits_of(n)));
        hipLaunchKernelGGL(kr_gather, dim3(grid_of(nnz)), dim3(256), 0, 0, nnz, d_iout, d_row, d_v, d_ri, d_vo);
        RL_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(kr_col_ptr, dim3(grid_of((long long)n + 1)), dim3(256), 0, 0, n, nnz, d_kout, d_cp);
    RL_TRY(hipGetLastError());
    RL_TRY(hipMemcpy(col_ptr, d_cp, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost));
    if (nnz > 0) {
        RL_TRY(hipMemcpy(row_idx, d_ri, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost));
        RL_TRY(hipMemcpy(val_out, d_vo, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToHost));
    }
    return PLSQR_OK;
}

int plsqr_ker2plsqr3(const char *dir, const char *list_file, int n_col, const char *out_bin, const char *out_info) {
    if (!dir || !list_file || !out_bin || !out_info || n_col <= 0)
        return plsqr::global_fail(PLSQR_ERR_ARG, "dir, list_file, out_bin, out_info and n_col are required");
    const std::string base(dir);
    auto join = [&](const std::string &f) { return base.empty() ? f : base + "/" + f; };
    // the list: "file nnz" per line, one file per kernel row (FileIO.c:353-402)
    FILE *lf = fopen(join(list_file).c_str(), "r");
    if (!lf) return plsqr::global_fail(PLSQR_ERR_IO, ("cannot open " + join(list_file)).c_str());
    std::vector<std::string> names;
    std::vector<int> counts;
    char name[4096];
    int cnt = 0;
    while (fscanf(lf, "%4095s %d", name, &cnt) == 2) {
        if (cnt < 0) {
            fclose(lf);
            return plsqr::global_fail(PLSQR_ERR_FORMAT, ("negative count for " + std::string(name)).c_str());
        }
        names.emplace_back(name);
        counts.push_back(cnt);
    }
    fclose(lf);
    const int m_k = (int)names.size();
    if (m_k == 0) return plsqr::global_fail(PLSQR_ERR_FORMAT, "the list names no row file");
    std::vector<int> rp((size_t)m_k + 1, 0);
    for (int r = 0; r < m_k; ++r) {
        if ((long long)rp[(size_t)r] + counts[(size_t)r] > 0x7fffffffLL)
            return plsqr::global_fail(PLSQR_ERR_FORMAT, "more than 2^31-1 kernel nonzeros");
        rp[(size_t)r + 1] = rp[(size_t)r] + counts[(size_t)r];
    }
    const int nnz = rp[(size_t)m_k];
    std::vector<int> col1((size_t)std::max(nnz, 1));
    std::vector<double> val((size_t)std::max(nnz, 1));
    struct Rec {  // {int32 aux, int32 col (1-based), float64 val}: 16 bytes
        int aux, col;
        double v;
    };
    static_assert(sizeof(Rec) == 16, "16-byte records");
    std::vector<Rec> buf;
    for (int r = 0; r < m_k; ++r) {
        const int k = counts[(size_t)r];
        FILE *f = fopen(join(names[(size_t)r]).c_str(), "rb");
        if (!f) return plsqr::global_fail(PLSQR_ERR_IO, ("cannot open " + join(names[(size_t)r])).c_str());
        buf.resize((size_t)std::max(k, 1));
        const size_t got = k > 0 ? fread(buf.data(), sizeof(Rec), (size_t)k, f) : 0;
        fclose(f);
        if (got != (size_t)k)
            return plsqr::global_fail(PLSQR_ERR_IO, ("short read in " + join(names[(size_t)r])).c_str());
        for (int j = 0; j < k; ++j) {
            col1[(size_t)rp[(size_t)r] + j] = buf[(size_t)j].col;
            val[(size_t)rp[(size_t)r] + j] = buf[(size_t)j].v;
        }
    }
    std::vector<int> cp((size_t)n_col + 1), ri((size_t)std::max(nnz, 1));
    std::vector<double> cv((size_t)std::max(nnz, 1));
    int rc = plsqr_kernel_rows_to_csc(m_k, n_col, rp.data(), col1.data(), val.data(), cp.data(), ri.data(), cv.data());
    if (rc) return rc;
    plsqr_system s;
    memset(&s, 0, sizeof(s));
    s.m_k = m_k;
    s.n = n_col;
    s.nnz_k = nnz;
    s.k_col_ptr = cp.data();
    s.k_row_idx = ri.data();
    s.k_val = cv.data();
    return plsqr_write_reference_files(&s, dir, out_bin, out_info, nullptr, nullptr, nullptr, nullptr, nullptr);
}

}  // extern "C"
