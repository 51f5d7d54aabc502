// MLP kernels of the learned warm start: BF16 MFMA GEMM with fused epilogues,
// loss gradient, bias gradients, Adam, casts.
//
// GEMM tiling: a 256-thread workgroup (4 wavefronts) computes a 64x64 tile of
// C; each wavefront owns a 32x32 sub-tile held in 16 FP32 accumulators per
// lane and issues v_mfma_f32_32x32x16_bf16 over K in steps of 16.  Operand
// tiles (64 x 32 of A, 32 x 64 of B) are staged in LDS with the K dimension
// contiguous, so each lane fetches its 8 consecutive K values of a row (A) or
// column (B) with one 16-byte LDS read.  Global loads map consecutive threads
// to the contiguous dimension of each operand (row- or column-major).
//
// MFMA operand/result layout (32x32x16): lane l supplies A[i][k] for
// i = l % 32, k = 8*(l/32) .. 8*(l/32)+7 and B[k][j] for j = l % 32 and the
// same k range; it receives D[i][j] for j = l % 32 and
// i = 8*(r/4) + 4*(l/32) + r%4 in accumulator r.

#include "pdlp_common.hpp"

#include <algorithm>

namespace pdlp {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                                                // nearest even
    return (uint16_t)(u >> 16);
}

constexpr int BM = 64, BN = 64, BK = 32, PAD = 8;

__global__ void __launch_bounds__(256) gemm_kernel(PdlpGemm g) {
    __shared__ __align__(16) uint16_t As[BM][BK + PAD];
    __shared__ __align__(16) uint16_t Bs[BN][BK + PAD];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;

    for (int k0 = 0; k0 < g.K; k0 += BK) {
        // A tile -> As[m][k]
#pragma unroll
        for (int e0 = 0; e0 < BM * BK; e0 += 256) {
            const int e = e0 + tid;
            int mm, kk;
            if (g.trans_a) { mm = e % BM; kk = e / BM; }     // A[k*lda + m]: m contiguous
            else { kk = e % BK; mm = e / BK; }               // A[m*lda + k]: k contiguous
            const int gm = m0 + mm, gk = k0 + kk;
            uint16_t v = 0;
            if (gm < g.M && gk < g.K) v = g.trans_a ? g.A[(size_t)gk * g.lda + gm] : g.A[(size_t)gm * g.lda + gk];
            As[mm][kk] = v;
        }
        // B tile -> Bs[n][k]
#pragma unroll
        for (int e0 = 0; e0 < BN * BK; e0 += 256) {
            const int e = e0 + tid;
            int nn, kk;
            if (g.trans_b) { kk = e % BK; nn = e / BK; }     // B[n*ldb + k]: k contiguous
            else { nn = e % BN; kk = e / BN; }               // B[k*ldb + n]: n contiguous
            const int gn = n0 + nn, gk = k0 + kk;
            uint16_t v = 0;
            if (gn < g.N && gk < g.K) v = g.trans_b ? g.B[(size_t)gn * g.ldb + gk] : g.B[(size_t)gk * g.ldb + gn];
            Bs[nn][kk] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const u16x8 a = *reinterpret_cast<const u16x8*>(&As[wm + (lane & 31)][kk + 8 * (lane >> 5)]);
            const u16x8 b = *reinterpret_cast<const u16x8*>(&Bs[wn + (lane & 31)][kk + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                          acc, 0, 0, 0);
        }
        __syncthreads();
    }

    const int col = n0 + wn + (lane & 31);
    if (col >= g.N) return;
    const float bias = g.bias != nullptr ? g.bias[col] : 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row >= g.M) continue;
        const size_t o = (size_t)row * g.ldc + col;
        float v = g.alpha * acc[r];
        if (g.accumulate && g.C != nullptr) v += g.C[o];
        v += bias;
        if (g.relu) v = v > 0.f ? v : 0.f;
        if (g.mask != nullptr && !(bf2f(g.mask[o]) > 0.f)) v = 0.f;
        if (g.C != nullptr) g.C[o] = v;
        if (g.Cb != nullptr) g.Cb[o] = f2bf(v);
    }
}

__global__ void zero1_kernel(float* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0.f;
}

// dout = 2 w (out - t) / (B N); loss += w (out - t)^2 / (B N)
__global__ void __launch_bounds__(256) loss_grad_kernel(const float* out, const float* t, const float* w, float* dout,
                                                        uint16_t* dout_b, float* loss, int64_t total, int N,
                                                        float inv) {
    __shared__ double lds[16 * 4];
    double part = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const float wj = w != nullptr ? w[e % N] : 1.f;
        const float d = out[e] - t[e];
        const float gr = 2.f * wj * d * inv;
        dout[e] = gr;
        dout_b[e] = f2bf(gr);
        part += (double)(wj * d * d) * inv;
    }
    double acc[4] = {part, 0.0, 0.0, 0.0};
    block_sum<4>(acc, lds);
    if (threadIdx.x == 0) atomicAdd(loss, (float)acc[0]);
}

// db[j] = sum_b dz[b][j]; dz FP32 (dzf) or BF16 (dzb)
__global__ void __launch_bounds__(256) colsum_kernel(const float* dzf, const uint16_t* dzb, float* db, int B, int N) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += dzf != nullptr ? dzf[(size_t)b * N + j] : bf2f(dzb[(size_t)b * N + j]);
    db[j] = s;
}

__global__ void __launch_bounds__(256) adam_kernel(float* w, const float* g, float* m, float* v, uint16_t* wb,
                                                   int64_t n, float lr, float b1, float b2, float eps, float c1,
                                                   float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        if (wb != nullptr) wb[i] = f2bf(wi);
    }
}

__global__ void __launch_bounds__(256) cast_kernel(const float* x, uint16_t* y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

int gemm(const PdlpGemm& g, hipStream_t s) {
    if (g.M < 0 || g.N < 0 || g.K < 0) return fail(PDLP_E_SHAPE, "gemm sizes must be >= 0");
    if (g.M == 0 || g.N == 0) return 0;
    if (!g.A || !g.B || (!g.C && !g.Cb)) return fail(PDLP_E_ARG, "gemm: A, B and an output are required");
    if (g.accumulate && !g.C) return fail(PDLP_E_ARG, "gemm: accumulate needs C");
    dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM);
    gemm_kernel<<<grid, 256, 0, s>>>(g);
    return hip_fail(hipGetLastError(), "gemm_kernel launch");
}

PdlpGemm make(int M, int N, int K, const uint16_t* A, int lda, int ta, const uint16_t* B, int ldb, int tb,
              const float* bias, int relu, const uint16_t* mask, float* C, uint16_t* Cb, int ldc) {
    PdlpGemm g{};
    g.M = M; g.N = N; g.K = K;
    g.lda = lda; g.ldb = ldb; g.ldc = ldc;
    g.trans_a = ta; g.trans_b = tb;
    g.relu = relu; g.accumulate = 0; g.alpha = 1.f;
    g.A = A; g.B = B; g.bias = bias; g.mask = mask; g.C = C; g.Cb = Cb;
    return g;
}

unsigned grid1(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" {

int pdlp_gemm_bf16(const PdlpGemm* g, void* stream) {
    if (g == nullptr) return fail(PDLP_E_ARG, "PdlpGemm is NULL");
    return gemm(*g, as_stream(stream));
}

int pdlp_mlp_fwd_bf16(const PdlpMlp* p, void* stream) {
    if (p == nullptr || !p->x || !p->w1 || !p->w2 || !p->w3 || !p->h1 || !p->h2 || !p->out)
        return fail(PDLP_E_ARG, "pdlp_mlp_fwd_bf16: NULL pointer");
    hipStream_t s = as_stream(stream);
    const int B = p->batch, di = p->d_in, dh = p->d_h, dout = p->d_out;
    int rc = gemm(make(B, dh, di, p->x, di, 0, p->w1, dh, 0, p->b1, 1, nullptr, nullptr, p->h1, dh), s);
    if (!rc) rc = gemm(make(B, dh, dh, p->h1, dh, 0, p->w2, dh, 0, p->b2, 1, nullptr, nullptr, p->h2, dh), s);
    if (!rc) rc = gemm(make(B, dout, dh, p->h2, dh, 0, p->w3, dout, 0, p->b3, 0, nullptr, p->out, nullptr, dout), s);
    return rc;
}

int pdlp_mlp_bwd_bf16(const PdlpMlp* p, void* stream) {
    if (p == nullptr || !p->target || !p->dout || !p->dout_b || !p->dz2 || !p->dz1 || !p->gw1 || !p->gw2 ||
        !p->gw3 || !p->gb1 || !p->gb2 || !p->gb3 || !p->loss)
        return fail(PDLP_E_ARG, "pdlp_mlp_bwd_bf16: NULL pointer");
    hipStream_t s = as_stream(stream);
    const int B = p->batch, di = p->d_in, dh = p->d_h, dout = p->d_out;
    const int64_t total = (int64_t)B * dout;
    zero1_kernel<<<1, 64, 0, s>>>(p->loss);
    loss_grad_kernel<<<grid1(total), 256, 0, s>>>(p->out, p->target, p->out_weight, p->dout, p->dout_b, p->loss,
                                                 total, dout, 1.f / (float)total);
    colsum_kernel<<<(dout + 255) / 256, 256, 0, s>>>(p->dout, nullptr, p->gb3, B, dout);
    int rc = hip_fail(hipGetLastError(), "mlp loss/bias kernels");
    // gw3 = h2^T dout ; dz2 = (dout w3^T) * [h2 > 0]
    if (!rc) rc = gemm(make(dh, dout, B, p->h2, dh, 1, p->dout_b, dout, 0, nullptr, 0, nullptr, p->gw3, nullptr, dout), s);
    if (!rc) rc = gemm(make(B, dh, dout, p->dout_b, dout, 0, p->w3, dout, 1, nullptr, 0, p->h2, nullptr, p->dz2, dh), s);
    if (!rc) {
        colsum_kernel<<<(dh + 255) / 256, 256, 0, s>>>(nullptr, p->dz2, p->gb2, B, dh);
        rc = hip_fail(hipGetLastError(), "colsum");
    }
    // gw2 = h1^T dz2 ; dz1 = (dz2 w2^T) * [h1 > 0]
    if (!rc) rc = gemm(make(dh, dh, B, p->h1, dh, 1, p->dz2, dh, 0, nullptr, 0, nullptr, p->gw2, nullptr, dh), s);
    if (!rc) rc = gemm(make(B, dh, dh, p->dz2, dh, 0, p->w2, dh, 1, nullptr, 0, p->h1, nullptr, p->dz1, dh), s);
    if (!rc) {
        colsum_kernel<<<(dh + 255) / 256, 256, 0, s>>>(nullptr, p->dz1, p->gb1, B, dh);
        rc = hip_fail(hipGetLastError(), "colsum");
    }
    // gw1 = x^T dz1
    if (!rc) rc = gemm(make(di, dh, B, p->x, di, 1, p->dz1, dh, 0, nullptr, 0, nullptr, p->gw1, nullptr, dh), s);
    return rc;
}

int pdlp_adam_step(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr, float beta1,
                   float beta2, float eps, int32_t step, void* stream) {
    if (n < 0 || step < 1) return fail(PDLP_E_ARG, "adam: n >= 0 and step >= 1");
    if (n == 0) return 0;
    if (!w || !g || !m || !v) return fail(PDLP_E_ARG, "adam: NULL pointer");
    const float c1 = 1.f / (1.f - powf(beta1, (float)step));
    const float c2 = 1.f / (1.f - powf(beta2, (float)step));
    adam_kernel<<<grid1(n), 256, 0, as_stream(stream)>>>(w, g, m, v, wb, n, lr, beta1, beta2, eps, c1, c2);
    return hip_fail(hipGetLastError(), "adam_kernel launch");
}

int pdlp_cast_bf16(const float* x, uint16_t* y, int64_t n, void* stream) {
    if (n < 0) return fail(PDLP_E_ARG, "cast: n >= 0");
    if (n == 0) return 0;
    if (!x || !y) return fail(PDLP_E_ARG, "cast: NULL pointer");
    cast_kernel<<<grid1(n), 256, 0, as_stream(stream)>>>(x, y, n);
    return hip_fail(hipGetLastError(), "cast_kernel launch");
}

}  // extern "C"
