// Version, error reporting and the standalone SpMV entry points.

#include "pdlp_common.hpp"

#include <algorithm>
#include <cstdio>

namespace pdlp {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    set_error(buf);
    return (int)e;
}

namespace {

// y = A x by groups of TPR lanes per row.
template <typename V, typename X, typename A, int TPR>
__global__ void __launch_bounds__(1024) spmv_kernel(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                    const V* __restrict__ val, const X* __restrict__ x,
                                                    A* __restrict__ y, int nrows) {
    const int gpb = blockDim.x / TPR, lane = threadIdx.x % TPR;
    const int r = blockIdx.x * gpb + threadIdx.x / TPR;
    if (r >= nrows) return;
    A dot, asum;
    row_dot<TPR, false, A>(rp, ci, val, r, lane, [&](int k) { return (A)x[k]; }, dot, asum);
    if (lane == 0) y[r] = dot;
}

template <typename V, typename X, typename A>
int launch_spmv(const int32_t* rp, const int32_t* ci, const V* val, const X* x, A* y, int nrows, int tpr,
                hipStream_t s) {
    constexpr int B = 256;
    if (nrows == 0) return 0;
    const unsigned g = (unsigned)(((long)nrows * tpr + B - 1) / B);
    switch (tpr) {
        case 1: spmv_kernel<V, X, A, 1><<<g, B, 0, s>>>(rp, ci, val, x, y, nrows); break;
        case 2: spmv_kernel<V, X, A, 2><<<g, B, 0, s>>>(rp, ci, val, x, y, nrows); break;
        case 4: spmv_kernel<V, X, A, 4><<<g, B, 0, s>>>(rp, ci, val, x, y, nrows); break;
        case 8: spmv_kernel<V, X, A, 8><<<g, B, 0, s>>>(rp, ci, val, x, y, nrows); break;
        case 16: spmv_kernel<V, X, A, 16><<<g, B, 0, s>>>(rp, ci, val, x, y, nrows); break;
        case 32: spmv_kernel<V, X, A, 32><<<g, B, 0, s>>>(rp, ci, val, x, y, nrows); break;
        case 64: spmv_kernel<V, X, A, 64><<<g, B, 0, s>>>(rp, ci, val, x, y, nrows); break;
        default: return fail(PDLP_E_TUNE, "tpr must be one of 1,2,4,8,16,32,64");
    }
    return hip_fail(hipGetLastError(), "spmv_kernel launch");
}

// y = A x with A as SELL-64 slices of its rows in natural order (pdlp_csr_to_sell64):
// one lane per row, no row pointer to read, every entry load a whole line; each
// row summed in CSR order (padding adds zeros), as the one-lane CSR kernel does.
template <typename T>
__global__ void __launch_bounds__(256) spmv_sell_kernel(const int32_t* __restrict__ sp, const int32_t* __restrict__ si,
                                                        const T* __restrict__ sv, const T* __restrict__ x,
                                                        T* __restrict__ y, int nrows) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = __builtin_amdgcn_readfirstlane(r >> 6);
    if (s >= (nrows + 63) >> 6) return;                // whole wavefronts leave together
    const int beg = __builtin_amdgcn_readfirstlane(sp[s]);
    const int width = __builtin_amdgcn_readfirstlane((sp[s + 1] - beg) >> 6);
    const int32_t* __restrict__ idx = si + beg + (r & 63);
    const T* __restrict__ val = sv + beg + (r & 63);
    T d = T(0);
    int k = 0;
    for (; k + kRowBatch <= width; k += kRowBatch) {
        T a[kRowBatch];
        int c[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            a[u] = val[64 * (k + u)];
            c[u] = idx[64 * (k + u)];
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) d += a[u] * x[c[u]];
    }
    for (; k < width; ++k) d += val[64 * k] * x[idx[64 * k]];
    if (r < nrows) y[r] = d;
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" {

int pdlp_abi_version(void) { return PDLP_ABI_VERSION; }

const char* pdlp_last_error(void) { return g_last_error.c_str(); }

int pdlp_spmv_csr(int32_t prec, const int32_t* row_ptr, const int32_t* col_idx, const void* val, const void* x,
                  void* y, int32_t nrows, int32_t tpr, void* stream) {
    if (nrows < 0) return fail(PDLP_E_ARG, "nrows must be >= 0");
    if (nrows > 0 && (!row_ptr || !col_idx || !val || !x || !y)) return fail(PDLP_E_ARG, "NULL pointer");
    if (prec == PDLP_F32)
        return launch_spmv<float, float, float>(row_ptr, col_idx, (const float*)val, (const float*)x, (float*)y,
                                                nrows, tpr, as_stream(stream));
    if (prec == PDLP_F64)
        return launch_spmv<double, double, double>(row_ptr, col_idx, (const double*)val, (const double*)x,
                                                   (double*)y, nrows, tpr, as_stream(stream));
    return fail(PDLP_E_ARG, "prec must be PDLP_F32 or PDLP_F64");
}

int pdlp_spmv_sell64(int32_t prec, const int32_t* sell_ptr, const int32_t* sell_idx, const void* sell_val,
                     const void* x, void* y, int32_t nrows, void* stream) {
    if (nrows < 0) return fail(PDLP_E_ARG, "nrows must be >= 0");
    if (nrows == 0) return 0;
    if (!sell_ptr || !sell_idx || !sell_val || !x || !y) return fail(PDLP_E_ARG, "NULL pointer");
    const unsigned g = (unsigned)((nrows + 255) / 256);
    hipStream_t s = as_stream(stream);
    if (prec == PDLP_F32)
        spmv_sell_kernel<float><<<g, 256, 0, s>>>(sell_ptr, sell_idx, (const float*)sell_val, (const float*)x,
                                                  (float*)y, nrows);
    else if (prec == PDLP_F64)
        spmv_sell_kernel<double><<<g, 256, 0, s>>>(sell_ptr, sell_idx, (const double*)sell_val, (const double*)x,
                                                   (double*)y, nrows);
    else
        return fail(PDLP_E_ARG, "prec must be PDLP_F32 or PDLP_F64");
    return hip_fail(hipGetLastError(), "spmv_sell_kernel launch");
}

int pdlp_spmv_csr_f64acc(const int32_t* row_ptr, const int32_t* col_idx, const float* val, const double* x,
                         double* y, int32_t nrows, int32_t tpr, void* stream) {
    if (nrows < 0) return fail(PDLP_E_ARG, "nrows must be >= 0");
    if (nrows > 0 && (!row_ptr || !col_idx || !val || !x || !y)) return fail(PDLP_E_ARG, "NULL pointer");
    return launch_spmv<float, double, double>(row_ptr, col_idx, val, x, y, nrows, tpr, as_stream(stream));
}

}  // extern "C"
