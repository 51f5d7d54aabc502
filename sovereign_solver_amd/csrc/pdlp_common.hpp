// Shared device helpers of libpdlp_hip.so (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <string>

#include "../../include/pdlp_hip.h"

namespace pdlp {

constexpr int kWave = 64;

// Entries per lane whose loads a row dot product issues together.
constexpr int kRowBatch = 4;

// The adaptive primal weight stays within [w0 / kWeightDown, w0 * kWeightUp]
// of its starting value w0.  With FP32 iterates an extreme weight (a tiny tau
// and a huge sigma, or the reverse) amplifies the rounding of the residual in
// the dual (primal) step until progress stalls.  Small random LPs with many
// empty columns (m = 50..120, n = 2.5m, 1% dense) drift downward: their FP32
// distance ratios keep shrinking the weight, which runs to any lower bound
// four decades away and never reaches 1e-4 KKT in 200,000 iterations, while
// with the bound at two decades they finish in 700..2,100 iterations.  The
// workloads that move the weight far all move it up (set-cover relaxations
// settle near 900x their start, transportation LPs near 140x), so the upper
// bound stays at four decades.
constexpr double kWeightDown = 1.0e2;
constexpr double kWeightUp = 1.0e4;

__host__ __device__ inline double clamp_weight(double w, double w0) {
    return fmin(fmax(w, w0 / kWeightDown), w0 * kWeightUp);
}

// The single-instance engine's argument checks and iteration blocks (pdlp_single.hip),
// shared with the library's other entry points.
int single_check_all(const PdlpDev* d, const PdlpTune* t);
int single_run(const PdlpDev& d, const PdlpTune& t, int iters, int check, hipStream_t s);

// Thread-local message behind pdlp_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Sum over aligned groups of W lanes inside one wavefront (W divides 64).
template <int W, typename A>
__device__ __forceinline__ A group_sum(A v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// Exact (non-contracted, round-to-nearest) arithmetic for the bitwise mode.
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }

// Workgroup index after the optional XCD remap.  Consecutive workgroups are
// dispatched round-robin over the 8 XCDs; with the remap, XCD k receives the
// contiguous index range [k*per, (k+1)*per), so rows that share gathered
// cache lines also share an L2.  The grid must be a multiple of 8 then.
__device__ __forceinline__ int block_index(int remap) {
    int b = blockIdx.x;
    if (remap) {
        int per = gridDim.x >> 3;
        b = (b & 7) * per + (b >> 3);
    }
    return b;
}

// Dot product of CSR row `row` with a gathered vector, plus the row's
// absolute sum, by a group of TPR lanes (lane = index inside the group).
// Exact mode: one lane, ascending column order, no FMA — the summation of
// SciPy's csr_matvec.
template <int TPR, bool EXACT, typename A, typename V, typename G>
__device__ __forceinline__ void row_dot(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                        const V* __restrict__ val, int row, int lane, G gather,
                                        A& dot, A& asum) {
    const int beg = rp[row], end = rp[row + 1];
    A d = A(0), s = A(0);
    if constexpr (EXACT) {
        for (int k = beg; k < end; ++k) {
            const A a = (A)val[k];
            d = add_rn(d, mul_rn(a, gather(ci[k])));
            s = add_rn(s, fabs(a));
        }
    } else {
        // kRowBatch entries per lane loaded before any is used (the same
        // summation order as one entry at a time)
        for (int k0 = beg + lane; k0 < end; k0 += TPR * kRowBatch) {
            A a[kRowBatch];
            int c[kRowBatch];
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) {
                const int k = k0 + u * TPR;
                a[u] = k < end ? (A)val[k] : A(0);
                c[u] = k < end ? ci[k] : 0;
            }
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) {
                if (k0 + u * TPR < end) {
                    d += a[u] * gather(c[u]);
                    s += fabs(a[u]);
                }
            }
        }
        if constexpr (TPR > 1) {
            d = group_sum<TPR>(d);
            s = group_sum<TPR>(s);
        }
    }
    dot = d;
    asum = s;
}

// Row dot product by a whole workgroup (TPR = blockDim.x > 64, one row per
// workgroup): strided partial sums, wavefront reduction, then the wavefront
// results through LDS.  Result valid in thread 0.  `lds` holds >= 2*16 A.
template <int TPR, typename A, typename V, typename G>
__device__ __forceinline__ void row_dot_block(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                              const V* __restrict__ val, int row, G gather, A& dot, A& asum,
                                              A* lds) {
    static_assert(TPR > 64 && TPR <= 1024 && TPR % 64 == 0, "block row width");
    const int beg = rp[row], end = rp[row + 1];
    A d = A(0), s = A(0);
    for (int k0 = beg + (int)threadIdx.x; k0 < end; k0 += TPR * kRowBatch) {
        A a[kRowBatch];
        int c[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            const int k = k0 + u * TPR;
            a[u] = k < end ? (A)val[k] : A(0);
            c[u] = k < end ? ci[k] : 0;
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            if (k0 + u * TPR < end) {
                d += a[u] * gather(c[u]);
                s += fabs(a[u]);
            }
        }
    }
    d = group_sum<kWave>(d);
    s = group_sum<kWave>(s);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        lds[2 * w] = d;
        lds[2 * w + 1] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d = A(0);
        s = A(0);
#pragma unroll
        for (int i = 0; i < TPR / kWave; ++i) {
            d += lds[2 * i];
            s += lds[2 * i + 1];
        }
    }
    dot = d;
    asum = s;
}

// Block-wide sum of NV doubles (blockDim.x a multiple of 64, <= 1024).
// Result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* >= 16*NV */) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = group_sum<kWave>(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[w * NV + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double t = 0.0;
            for (int i = 0; i < nw; ++i) t += lds[i * NV + k];
            v[k] = t;
        }
    }
    __syncthreads();
}

}  // namespace pdlp
