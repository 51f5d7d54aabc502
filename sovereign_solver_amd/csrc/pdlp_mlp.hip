// MLP kernels of the learned warm start: BF16 MFMA GEMM with fused epilogues,
// loss gradient, bias gradients, Adam, casts.
//
// GEMM tiling: a 256-thread workgroup (4 wavefronts, 2 x 2) computes a
// 128 x 128 tile of C; each wavefront owns a 64 x 64 sub-tile as 2 x 2 blocks
// of 32 x 32, each held in 16 FP32 accumulators per lane and updated by
// v_mfma_f32_32x32x16_bf16 over K in steps of 16.  K advances in stages of 64:
// the next stage's operand tiles are loaded from global memory into registers
// while the MFMAs of the current stage run out of LDS (a second register set,
// two stages ahead, measured no faster on the MLP's shapes).  In LDS both operand
// tiles keep K contiguous (As[m][k], Bs[n][k]), so each lane fetches its 8
// consecutive K values of a row (A) or column (B) with one 16-byte LDS read.
// Global loads move 8 BF16 values (16 bytes) per thread along the operand's
// contiguous dimension when the leading dimensions are multiples of 8 and the
// bases 16-byte aligned; an operand whose contiguous dimension is M or N is
// transposed on its way into LDS.  Otherwise consecutive threads load
// consecutive 2-byte values.
//
// A long K with few output tiles (the MLP's dz2 = dout w3^T: M = batch,
// N = d_h, K = d_out) is split: workgroup z of split_k computes the partial
// product of its K range into work[z], and a second kernel sums the parts in
// order of z and applies the epilogue, so the result does not depend on
// scheduling.
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

constexpr int BM = 128, BN = 128, BK = 64, PAD = 8, LDT = BK + PAD;
constexpr int GEMM_THREADS = 256;

// One operand tile of R rows (M or N) by BK values of K, held in registers
// between its global load and its LDS store.  TR: the operand's contiguous
// dimension is R (op(A) = A^T / op(B) = B stored by K rows) instead of K.
// Transposed 16-byte path: a unit is two consecutive K rows (k, k+1) of 8
// consecutive R values; in LDS the pair goes out as 8 dword stores
// (S[r+u][k], S[r+u][k+1]).  A wavefront's 64 units are 32 K pairs of two
// R vectors 8 rows apart, whose dwords fall in 64 different banks (row
// stride 36 dwords).
template <bool VEC, bool TR, int R>
struct TileLoader {
    static constexpr int VPT = R * BK / 8 / GEMM_THREADS;       // 8-value vectors per thread
    static constexpr int EPT = R * BK / GEMM_THREADS;           // 2-byte values per thread
    static constexpr int NREG = VEC ? VPT : EPT;
    typedef unsigned short u16x8v __attribute__((ext_vector_type(8)));
    u16x8v v[VEC ? VPT : 1];
    uint16_t e[VEC ? 1 : EPT];

    // element (r, k) of the tile at (r0, k0): X[(r0+r)*ld + k0+k] (K contiguous) or X[(k0+k)*ld + r0+r]
    __device__ __forceinline__ void load(const uint16_t* __restrict__ X, int ld, int r0, int rmax, int k0, int kend) {
        const int t = threadIdx.x;
        if constexpr (VEC) {
#pragma unroll
            for (int q = 0; q < VPT; ++q) {
                int r, k;
                if constexpr (TR) {
                    const int w = t + (q >> 1) * GEMM_THREADS;       // unit: K pair w % 32, R vector w / 32
                    k = 2 * (w % (BK / 2)) + (q & 1);
                    r = (w / (BK / 2)) * 8;
                } else {
                    const int w = t + q * GEMM_THREADS;
                    r = w / (BK / 8);
                    k = (w % (BK / 8)) * 8;
                }
                const int gr = r0 + r, gk = k0 + k;
                u16x8v x;
                if constexpr (TR) {
                    if (gk < kend && gr + 7 < rmax) {
                        x = *reinterpret_cast<const u16x8v*>(X + (size_t)gk * ld + gr);
                    } else {
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            x[u] = (gk < kend && gr + u < rmax) ? X[(size_t)gk * ld + gr + u] : (uint16_t)0;
                    }
                } else {
                    if (gr < rmax && gk + 7 < kend) {
                        x = *reinterpret_cast<const u16x8v*>(X + (size_t)gr * ld + gk);
                    } else {
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            x[u] = (gr < rmax && gk + u < kend) ? X[(size_t)gr * ld + gk + u] : (uint16_t)0;
                    }
                }
                v[q] = x;
            }
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int w = t + q * GEMM_THREADS;
                int r, k;
                if constexpr (TR) { k = w / R; r = w % R; }
                else { r = w / BK; k = w % BK; }
                const int gr = r0 + r, gk = k0 + k;
                e[q] = (gr < rmax && gk < kend) ? (TR ? X[(size_t)gk * ld + gr] : X[(size_t)gr * ld + gk]) : (uint16_t)0;
            }
        }
    }

    __device__ __forceinline__ void store(uint16_t (*S)[LDT]) const {
        const int t = threadIdx.x;
        if constexpr (VEC) {
#pragma unroll
            for (int q = 0; q < VPT; ++q) {
                if constexpr (TR) {
                    if (q & 1) continue;
                    const int w = t + (q >> 1) * GEMM_THREADS;
                    const int k = 2 * (w % (BK / 2)), r = (w / (BK / 2)) * 8;
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        *reinterpret_cast<uint32_t*>(&S[r + u][k]) = (uint32_t)v[q][u] | ((uint32_t)v[q + 1][u] << 16);
                } else {
                    const int w = t + q * GEMM_THREADS;
                    const int r = w / (BK / 8), k = (w % (BK / 8)) * 8;
                    *reinterpret_cast<u16x8v*>(&S[r][k]) = v[q];
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int w = t + q * GEMM_THREADS;
                int r, k;
                if constexpr (TR) { k = w / R; r = w % R; }
                else { r = w / BK; k = w % BK; }
                S[r][k] = e[q];
            }
        }
    }
};

// alpha * acc (+ C) + bias, ReLU, mask; stores C and/or Cb
__device__ __forceinline__ void epilogue(const PdlpGemm& g, int row, int col, float acc) {
    const size_t o = (size_t)row * g.ldc + col;
    float v = g.alpha * acc;
    if (g.accumulate && g.C != nullptr) v += g.C[o];
    if (g.bias != nullptr) v += g.bias[col];
    if (g.relu) v = v > 0.f ? v : 0.f;
    if (g.mask != nullptr && !(bf2f(g.mask[o]) > 0.f)) v = 0.f;
    if (g.C != nullptr) g.C[o] = v;
    if (g.Cb != nullptr) g.Cb[o] = f2bf(v);
}

// blockIdx.x: N tile, blockIdx.y: M tile, blockIdx.z: K part (kc values of K each, a multiple of BK)
template <bool VEC, bool TA, bool TB>
__global__ void __launch_bounds__(GEMM_THREADS) gemm_kernel(PdlpGemm g, int kc) {
    __shared__ __align__(16) uint16_t As[BM][LDT];
    __shared__ __align__(16) uint16_t Bs[BN][LDT];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kbeg = blockIdx.z * kc, kend = min(g.K, kbeg + kc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    // the MFMAs of one stage out of LDS
    auto compute = [&]() {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            u16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const u16x8*>(&As[wm + 32 * i + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const u16x8*>(&Bs[wn + 32 * j + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        __builtin_bit_cast(bf16x8, a[i]), __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
        }
    };
    typedef TileLoader<VEC, TA, BM> LA;
    typedef TileLoader<VEC, !TB, BN> LB;  // op(B)(k, n): B[n*ldb + k] has K contiguous
    // one stage: the registers of stage k0 into LDS, stage knext into the same registers, MFMAs
    auto stage = [&](LA& la, LB& lb, int knext) {
        __syncthreads();                  // the previous stage's MFMAs are done with LDS
        la.store(As);
        lb.store(Bs);
        __syncthreads();
        if (knext < kend) {               // in flight during this stage's MFMAs
            la.load(g.A, g.lda, m0, g.M, knext, kend);
            lb.load(g.B, g.ldb, n0, g.N, knext, kend);
        }
        compute();
    };
    LA la;
    LB lb;
    if (kbeg < kend) {
        la.load(g.A, g.lda, m0, g.M, kbeg, kend);
        lb.load(g.B, g.ldb, n0, g.N, kbeg, kend);
    }
    for (int k0 = kbeg; k0 < kend; k0 += BK) stage(la, lb, k0 + BK);

    const bool split = gridDim.z > 1;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int col = n0 + wn + 32 * j + (lane & 31);
        if (col >= g.N) continue;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (row >= g.M) continue;
                if (split) g.work[((size_t)blockIdx.z * g.M + row) * g.N + col] = acc[i][j][r];
                else epilogue(g, row, col, acc[i][j][r]);
            }
        }
    }
}

// Sum of `parts` partial arrays part[z][0 .. total) in a fixed order (the same
// bits every run), then the GEMM epilogue (EPI) or a plain store.  A workgroup
// takes 64 consecutive elements; its 4 wavefronts sum every 4th part (four
// loads in flight per lane), then the 4 sums are added in LDS in wavefront order.
template <bool EPI>
__global__ void __launch_bounds__(256) sum_parts_kernel(const float* __restrict__ part, int64_t total, int parts,
                                                        float* out, PdlpGemm g) {
    __shared__ float red[4][64];
    const int c = threadIdx.x & 63, grp = threadIdx.x >> 6;
    for (int64_t base = blockIdx.x * 64LL; base < total; base += gridDim.x * 64LL) {
        const int64_t e = base + c;
        float s = 0.f;
        if (e < total) {
            int z = grp;
            for (; z + 12 < parts; z += 16) {
                const float v0 = part[(size_t)z * total + e], v1 = part[(size_t)(z + 4) * total + e];
                const float v2 = part[(size_t)(z + 8) * total + e], v3 = part[(size_t)(z + 12) * total + e];
                s += v0;
                s += v1;
                s += v2;
                s += v3;
            }
            for (; z < parts; z += 4) s += part[(size_t)z * total + e];
        }
        red[grp][c] = s;
        __syncthreads();
        if (grp == 0 && e < total) {
            const float t = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
            if constexpr (EPI) epilogue(g, (int)(e / g.N), (int)(e % g.N), t);
            else out[e] = t;
        }
        __syncthreads();
    }
}

unsigned parts_grid(int64_t total) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 63) / 64, 8192)); }

__global__ void zero1_kernel(float* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0.f;
}

// dout = 2 w (out - t) / (B N); loss += w (out - t)^2 / (B N), over the
// flattened B x N elements (any N, from one column to thousands); one atomic
// per workgroup for the loss.
__global__ void __launch_bounds__(256) loss_grad_kernel(const float* out, const float* t, const float* w, float* dout,
                                                        uint16_t* dout_b, float* loss, int64_t total, int N,
                                                        float inv) {
    __shared__ double lds[16 * 4];
    double part = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const float wj = w != nullptr ? w[total < INT32_MAX ? (int)e % N : (int)(e % N)] : 1.f;
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

// db[j] = sum_b dz[b][j]; dz FP32 (dzf) or BF16 (dzb).  A workgroup takes 64
// columns of one chunk of rows (blockIdx.y; rows_per rows each); its 16
// wavefronts sum every 16th row (8 rows' loads in flight per lane), then add
// their sums in LDS in a fixed order.  With more than one chunk the sums go
// to part[chunk][j] and sum_parts_kernel adds them in chunk order.
constexpr int CS_COLS = 64, CS_GROUPS = 16, CS_UNROLL = 8;

__global__ void __launch_bounds__(CS_COLS * CS_GROUPS) colsum_kernel(const float* dzf, const uint16_t* dzb, float* db,
                                                                   float* part_out, int B, int N, int rows_per) {
    __shared__ float part[CS_GROUPS][CS_COLS];
    const int c = threadIdx.x % CS_COLS, grp = threadIdx.x / CS_COLS;
    const int j = blockIdx.x * CS_COLS + c;
    const int rbeg = blockIdx.y * rows_per, rend = min(B, rbeg + rows_per);
    float s = 0.f;
    if (j < N) {
        for (int b0 = rbeg + grp; b0 < rend; b0 += CS_GROUPS * CS_UNROLL) {
            float v[CS_UNROLL];
#pragma unroll
            for (int u = 0; u < CS_UNROLL; ++u) {
                const int b = b0 + u * CS_GROUPS;
                v[u] = b < rend ? (dzf != nullptr ? dzf[(size_t)b * N + j] : bf2f(dzb[(size_t)b * N + j])) : 0.f;
            }
#pragma unroll
            for (int u = 0; u < CS_UNROLL; ++u) s += v[u];
        }
    }
    part[grp][c] = s;
    __syncthreads();
    if (grp == 0 && j < N) {
        float t = part[0][c];
#pragma unroll
        for (int g = 1; g < CS_GROUPS; ++g) t += part[g][c];
        if (gridDim.y > 1) part_out[(size_t)blockIdx.y * N + j] = t;
        else db[j] = t;
    }
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

// --------------------------------------------------------- fused inference
// The whole 3-layer forward pass of FI_ROWS input rows in one workgroup, for
// d_h = 512: out = relu(relu(x W1 + b1) W2 + b2) W3 + b3 with the hidden
// activations rounded to BF16 as the layer-by-layer path rounds them, but
// never written to memory.  Each of the 8 wavefronts owns 64 hidden units
// and all FI_ROWS rows; every product is computed transposed (units x rows),
// D = W^T X^T, so that:
//   layer 1: A = W1^T (global), B = x rows (global); the result, after bias
//            and ReLU, goes to LDS as H1[row][unit] in BF16;
//   layer 2: A = W2^T (global), B = H1 rows from LDS; a 64-k group is four
//            MFMA steps whose k slots map to k = kg + 32 h + 8 s + t (half h
//            of the wavefront, step s, element t), the same for both operands;
//   layer 3: B = the layer-2 accumulators themselves after bias, ReLU and
//            BF16 rounding: accumulator r of a 32x32 block holds unit
//            8 (r/4) + 4 h + r % 4 of the lane's row, so accumulators 8q..8q+7
//            are the B operand of one MFMA whose k slot (h, t) is unit
//            16 q + 8 (t/4) + 4 h + t % 4; A = W3 gathered into that order
//            (w3p).
// The weights come packed by pdlp_mlp_infer_pack in MFMA fragment order (one
// fragment = the 64 lanes' 16-byte operands, 1 KB of consecutive memory), so
// each A load instruction reads 8 whole cache lines instead of touching 32.
// The wavefronts' layer-3 partial sums (over their 64 units) meet in LDS (in
// the space of H1) and are added in wavefront order, so the result does not
// depend on scheduling.
constexpr int FI_ROWS = 128, FI_DH = 512, FI_WAVES = 8, FI_THREADS = FI_WAVES * 64;
constexpr int FI_LDH = FI_DH + 8;               // H1 row stride (BF16): a 16-byte pad per row
constexpr int FI_MAX_DIN = 128, FI_MAX_DOUT = 32;
constexpr size_t FI_H1 = (size_t)FI_ROWS * FI_LDH * 2;    // H1; later >= FI_WAVES * 32 * FI_ROWS floats
constexpr size_t FI_LDS = FI_H1 + FI_DH * 4;                // + b2 (FP32)

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ u16x8 ld16(const uint16_t* p) { return *reinterpret_cast<const u16x8*>(p); }

// FP32 -> BF16, round to nearest even (the hardware conversion of gfx950; finite
// values as f2bf)
__device__ __forceinline__ uint16_t bf16_rne(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }

__device__ __forceinline__ f32x16 mfma(const u16x8& a, const u16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// acc[ub][rb] += the 512-k product of this wavefront's A fragments (w2w: its first
// fragment of the packed 512 x 512 weight, fragment order (k group, unit block,
// step)) and the FI_ROWS x 512 BF16 rows in LDS (hrow: row li, column 32 h);
// `an` holds k group 0's fragments on entry.  Two register sets of A in turn: the
// loads of a group are issued a whole group (32 MFMAs) before its MFMAs; the loop
// is kept rolled and scheduling barriers keep the loads there.
__device__ __forceinline__ void hidden_product(const uint16_t* w2w, const uint16_t* hrow, u16x8 (&an)[2][4],
                                               f32x16 (&acc)[2][4]) {
    auto group_load = [&](int gk, u16x8 (&dst)[2][4]) {
        const uint16_t* g = w2w + (size_t)gk * 2 * 4 * 512;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            dst[0][s] = ld16(g + (size_t)(0 * 4 + s) * 512);
            dst[1][s] = ld16(g + (size_t)(1 * 4 + s) * 512);
        }
    };
    auto group_mfma = [&](int kg, const u16x8 (&a)[2][4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            u16x8 b[4];
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
                b[rb] = *reinterpret_cast<const u16x8*>(hrow + (size_t)32 * rb * FI_LDH + kg + 8 * s);
#pragma unroll
            for (int ub = 0; ub < 2; ++ub)
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) acc[ub][rb] = mfma(a[ub][s], b[rb], acc[ub][rb]);
        }
    };
    u16x8 ab[2][4];
#pragma unroll 1
    for (int kg = 0; kg < FI_DH; kg += 128) {
        group_load(kg / 64 + 1, ab);
        __builtin_amdgcn_sched_barrier(0);          // the loads stay ahead of the group's MFMAs
        group_mfma(kg, an);
        if (kg + 128 < FI_DH) group_load(kg / 64 + 2, an);
        __builtin_amdgcn_sched_barrier(0);
        group_mfma(kg + 64, ab);
    }
}

// NK: layer-1 MFMA steps (k1 = 16 NK).  Layer 1 runs one block of 32 rows at a
// time with every W1^T operand loaded once up front and the next block's x
// rows in flight during the current block's MFMAs; each block goes to LDS
// (bias, ReLU, BF16) as soon as it is done, so only its two accumulator blocks
// are live.  The first layer-2 operand group is issued before layer 1.
// ST: also store the hidden activations h1, h2 (BF16, [rows][512]) for a backward pass.
template <int NK, bool ST>
__global__ void __launch_bounds__(FI_THREADS) mlp_infer_kernel(PdlpMlpInfer p) {
    constexpr int k1 = 16 * NK;
    extern __shared__ __align__(16) uint16_t fi_lds[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 31, lh = lane >> 5;
    const int r0 = blockIdx.x * FI_ROWS;
    const int u0 = w * 64;
    // b2 to LDS for the layer-3 epilogue (the barrier after layer 1 orders it); this
    // lane's b1 values (units 8 g + 4 h .. + 3 of each block) in registers
    float* bl2 = reinterpret_cast<float*>(reinterpret_cast<char*>(fi_lds) + FI_H1);
    bl2[threadIdx.x] = p.b2[threadIdx.x];
    f32x4 bias1[2][4];
#pragma unroll
    for (int ub = 0; ub < 2; ++ub)
#pragma unroll
        for (int g = 0; g < 4; ++g) bias1[ub][g] = *reinterpret_cast<const f32x4*>(p.b1 + u0 + 32 * ub + 8 * g + 4 * lh);

    // The packed weights are in fragment order: the 64 lanes' operands of one MFMA
    // are 1 KB of consecutive memory, so every load instruction reads whole lines.
    // W2: fragment (wave, k group, unit block, step) at ((w * 8 + kg/64) * 2 + ub) * 4 + s.
    const uint16_t* w2w = p.w2p + (size_t)w * (FI_DH / 64) * 2 * 4 * 512 + lane * 8;
    // layer 2's first A group (no dependence on layer 1): in flight during layer 1
    u16x8 an[2][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        an[0][s] = ld16(w2w + (size_t)(0 * 4 + s) * 512);
        an[1][s] = ld16(w2w + (size_t)(1 * 4 + s) * 512);
    }

    // ---- layer 1: D[unit][row] = sum_k W1^T[unit][k] x[row][k], one row block at a time
    // W1: fragment (wave, unit block, step) at (w * 2 + ub) * NK + s
    u16x8 a1[2][NK], bx[NK], bn[NK];
#pragma unroll
    for (int ub = 0; ub < 2; ++ub)
#pragma unroll
        for (int s = 0; s < NK; ++s) a1[ub][s] = ld16(p.w1p + ((size_t)((w * 2 + ub) * NK + s) * 64 + lane) * 8);
    // branch-free: every lane loads from a valid address (row and k clamped) and
    // zeroes what lies beyond the rows or d_in, so the loads stay in flight together
    auto load_x = [&](int rb, u16x8 (&dst)[NK]) {
        const int row = r0 + 32 * rb + li;
        const uint16_t* xr = p.x + (size_t)min(row, p.rows - 1) * p.d_in;
#pragma unroll
        for (int s = 0; s < NK; ++s) {
            const int kk = 16 * s + 8 * lh;
            const u16x8 v = ld16(xr + min(kk, p.d_in - 8));
            const bool ok = row < p.rows && kk < p.d_in;
#pragma unroll
            for (int t = 0; t < 8; ++t) dst[s][t] = ok ? v[t] : (unsigned short)0;
        }
    };
    load_x(0, bx);
    __builtin_amdgcn_sched_barrier(0);              // every load above is issued before any wait
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        if (rb < 3) load_x(rb + 1, bn);
        f32x16 c[2];
#pragma unroll
        for (int ub = 0; ub < 2; ++ub)
#pragma unroll
            for (int r = 0; r < 16; ++r) c[ub][r] = 0.f;
#pragma unroll
        for (int s = 0; s < NK; ++s)
#pragma unroll
            for (int ub = 0; ub < 2; ++ub) c[ub] = mfma(a1[ub][s], bx[s], c[ub]);
        // bias, ReLU, BF16 -> H1[row][unit]: accumulators 4g..4g+3 are 4 consecutive units
#pragma unroll
        for (int ub = 0; ub < 2; ++ub)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int unit = u0 + 32 * ub + 8 * g + 4 * lh;
                const f32x4 bias = bias1[ub][g];
                const float v0 = fmaxf(c[ub][4 * g + 0] + bias[0], 0.f);
                const float v1 = fmaxf(c[ub][4 * g + 1] + bias[1], 0.f);
                const float v2 = fmaxf(c[ub][4 * g + 2] + bias[2], 0.f);
                const float v3 = fmaxf(c[ub][4 * g + 3] + bias[3], 0.f);
                uint2 pk;
                pk.x = (uint32_t)bf16_rne(v0) | ((uint32_t)bf16_rne(v1) << 16);
                pk.y = (uint32_t)bf16_rne(v2) | ((uint32_t)bf16_rne(v3) << 16);
                *reinterpret_cast<uint2*>(fi_lds + (size_t)(32 * rb + li) * FI_LDH + unit) = pk;
            }
        if (rb < 3) {
#pragma unroll
            for (int s = 0; s < NK; ++s) bx[s] = bn[s];
        }
    }
    __syncthreads();
    if constexpr (ST) {                             // H1 rows to h1, 16 bytes a thread at a time
        for (int e = threadIdx.x; e < FI_ROWS * (FI_DH / 8); e += FI_THREADS) {
            const int row = e / (FI_DH / 8), c8 = (e % (FI_DH / 8)) * 8;
            if (r0 + row < p.rows)
                *reinterpret_cast<u16x8*>(p.h1 + (size_t)(r0 + row) * FI_DH + c8) =
                    *reinterpret_cast<const u16x8*>(fi_lds + (size_t)row * FI_LDH + c8);
        }
    }

    f32x16 acc[2][4];
    // ---- layer 2: D[unit][row] = sum_k W2^T[unit][k] H1[row][k], k in groups of 64
#pragma unroll
    for (int ub = 0; ub < 2; ++ub)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[ub][rb][r] = 0.f;
    hidden_product(w2w, fi_lds + (size_t)li * FI_LDH + 32 * lh, an, acc);

    // ---- layer 3 from the accumulators: D3[channel][row] over this wavefront's units
    f32x16 acc3[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc3[rb][r] = 0.f;
#pragma unroll
    for (int ub = 0; ub < 2; ++ub)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int blk = (u0 >> 4) + 2 * ub + q;
            const u16x8 a3 = ld16(p.w3p + ((size_t)(blk * 32 + li) * 2 + lh) * 8);
            const int ubase = u0 + 32 * ub + 16 * q + 4 * lh;
            const f32x4 bl = *reinterpret_cast<const f32x4*>(bl2 + ubase);
            const f32x4 bh = *reinterpret_cast<const f32x4*>(bl2 + ubase + 8);
            const float bias[8] = {bl[0], bl[1], bl[2], bl[3], bh[0], bh[1], bh[2], bh[3]};
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                u16x8 hv;
#pragma unroll
                for (int t = 0; t < 8; ++t) hv[t] = bf16_rne(fmaxf(acc[ub][rb][8 * q + t] + bias[t], 0.f));
                if constexpr (ST) {                 // units ubase + 0..3 and ubase + 8..11 of this row
                    const int row = r0 + 32 * rb + li;
                    if (row < p.rows) {
                        uint16_t* h = p.h2 + (size_t)row * FI_DH + ubase;
                        *reinterpret_cast<uint2*>(h) = make_uint2((uint32_t)hv[0] | ((uint32_t)hv[1] << 16),
                                                                  (uint32_t)hv[2] | ((uint32_t)hv[3] << 16));
                        *reinterpret_cast<uint2*>(h + 8) = make_uint2((uint32_t)hv[4] | ((uint32_t)hv[5] << 16),
                                                                      (uint32_t)hv[6] | ((uint32_t)hv[7] << 16));
                    }
                }
                acc3[rb] = mfma(a3, hv, acc3[rb]);
            }
        }
    __syncthreads();                                // every wavefront is done with H1
    float* red = reinterpret_cast<float*>(fi_lds);  // [wave][channel 32][row]
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int ch = 8 * (r >> 2) + 4 * lh + (r & 3);
            if (ch < p.d_out) red[((size_t)w * 32 + ch) * FI_ROWS + 32 * rb + li] = acc3[rb][r];
        }
    __syncthreads();
    for (int e = threadIdx.x; e < FI_ROWS * p.d_out; e += FI_THREADS) {
        const int row = e / p.d_out, ch = e - row * p.d_out;
        float s = 0.f;
#pragma unroll
        for (int v = 0; v < FI_WAVES; ++v) s += red[((size_t)v * 32 + ch) * FI_ROWS + row];
        if (r0 + row < p.rows) p.out[(size_t)(r0 + row) * p.d_out + ch] = s + p.b3[ch];
    }
}

// ------------------------------------------------------- fused backward dz
// dz2 = BF16((dout W3^T) * [h2 > 0]) and dz1 = BF16((dz2 W2^T) * [h1 > 0]) of FI_ROWS
// rows per workgroup in one kernel (the two masked products of pdlp_mlp_bwd_bf16,
// same rounding points), transposed as in mlp_infer_kernel: each of the 8
// wavefronts owns 64 hidden units.  h2 is staged into LDS; the dz2 tile replaces
// it in place (each element read and written by the same lane), goes to global
// memory for the weight and bias gradients, and is the B operand of the second
// product (A = W2 rows in fragment order, w2b).  NK3: MFMA steps over d_out.
template <int NK3>
__global__ void __launch_bounds__(FI_THREADS) mlp_dz_kernel(PdlpMlp p) {
    extern __shared__ __align__(16) uint16_t fi_lds[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 31, lh = lane >> 5;
    const int r0 = blockIdx.x * FI_ROWS, R = p.batch;
    const int u0 = w * 64;
    // the second product's first A group, in flight during the first product
    const uint16_t* w2w = p.w2b + (size_t)w * (FI_DH / 64) * 2 * 4 * 512 + lane * 8;
    u16x8 an[2][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        an[0][s] = ld16(w2w + (size_t)(0 * 4 + s) * 512);
        an[1][s] = ld16(w2w + (size_t)(1 * 4 + s) * 512);
    }
    // first product's operands: W3 fragments ((w*2 + ub)*NK3 + s) and dout rows (zero beyond d_out)
    u16x8 a3[2][NK3], bd[4][NK3];
#pragma unroll
    for (int ub = 0; ub < 2; ++ub)
#pragma unroll
        for (int s = 0; s < NK3; ++s) a3[ub][s] = ld16(p.w3b + ((size_t)((w * 2 + ub) * NK3 + s) * 64 + lane) * 8);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const int row = r0 + 32 * rb + li;
        const uint16_t* dr = p.dout_b + (size_t)min(row, R - 1) * p.d_out;
#pragma unroll
        for (int s = 0; s < NK3; ++s) {
            const int kk = 16 * s + 8 * lh;
            const u16x8 v = ld16(dr + min(kk, p.d_out - 8));
            const bool ok = row < R && kk < p.d_out;
#pragma unroll
            for (int t = 0; t < 8; ++t) bd[rb][s][t] = ok ? v[t] : (unsigned short)0;
        }
    }
    // h2 tile -> LDS [row][unit] (zeros beyond the batch: their dz2 is 0)
    for (int e = threadIdx.x; e < FI_ROWS * (FI_DH / 8); e += FI_THREADS) {
        const int row = e / (FI_DH / 8), c8 = (e % (FI_DH / 8)) * 8;
        u16x8 v = {};
        if (r0 + row < R) v = ld16(p.h2 + (size_t)(r0 + row) * FI_DH + c8);
        *reinterpret_cast<u16x8*>(fi_lds + (size_t)row * FI_LDH + c8) = v;
    }
    __syncthreads();
    // ---- dz2[unit][row] = [h2 > 0] sum_c W3[unit][c] dout[row][c], in place of h2
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        f32x16 c[2];
#pragma unroll
        for (int ub = 0; ub < 2; ++ub) {
#pragma unroll
            for (int r = 0; r < 16; ++r) c[ub][r] = 0.f;
#pragma unroll
            for (int s = 0; s < NK3; ++s) c[ub] = mfma(a3[ub][s], bd[rb][s], c[ub]);
        }
#pragma unroll
        for (int ub = 0; ub < 2; ++ub)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                uint2* cell = reinterpret_cast<uint2*>(fi_lds + (size_t)(32 * rb + li) * FI_LDH + u0 + 32 * ub + 8 * g +
                                                      4 * lh);
                const uint2 m = *cell;
                const float h[4] = {bf2f((uint16_t)(m.x & 0xffff)), bf2f((uint16_t)(m.x >> 16)),
                                    bf2f((uint16_t)(m.y & 0xffff)), bf2f((uint16_t)(m.y >> 16))};
                uint16_t v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = bf16_rne(h[e] > 0.f ? c[ub][4 * g + e] : 0.f);
                *cell = make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
            }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < FI_ROWS * (FI_DH / 8); e += FI_THREADS) {     // dz2 to global
        const int row = e / (FI_DH / 8), c8 = (e % (FI_DH / 8)) * 8;
        if (r0 + row < R)
            *reinterpret_cast<u16x8*>(p.dz2 + (size_t)(r0 + row) * FI_DH + c8) =
                *reinterpret_cast<const u16x8*>(fi_lds + (size_t)row * FI_LDH + c8);
    }
    // ---- dz1[unit][row] = [h1 > 0] sum_j W2[unit][j] dz2[row][j]
    f32x16 acc[2][4];
#pragma unroll
    for (int ub = 0; ub < 2; ++ub)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[ub][rb][r] = 0.f;
    hidden_product(w2w, fi_lds + (size_t)li * FI_LDH + 32 * lh, an, acc);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const int row = r0 + 32 * rb + li;
        const uint16_t* hr = p.h1 + (size_t)min(row, R - 1) * FI_DH;
        uint2 m[2][4];
#pragma unroll
        for (int ub = 0; ub < 2; ++ub)
#pragma unroll
            for (int g = 0; g < 4; ++g) m[ub][g] = *reinterpret_cast<const uint2*>(hr + u0 + 32 * ub + 8 * g + 4 * lh);
        if (row < R) {
#pragma unroll
            for (int ub = 0; ub < 2; ++ub)
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const float h[4] = {bf2f((uint16_t)(m[ub][g].x & 0xffff)), bf2f((uint16_t)(m[ub][g].x >> 16)),
                                        bf2f((uint16_t)(m[ub][g].y & 0xffff)), bf2f((uint16_t)(m[ub][g].y >> 16))};
                    uint16_t v[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = bf16_rne(h[e] > 0.f ? acc[ub][rb][4 * g + e] : 0.f);
                    *reinterpret_cast<uint2*>(p.dz1 + (size_t)row * FI_DH + u0 + 32 * ub + 8 * g + 4 * lh) =
                        make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
                }
        }
    }
}

// Backward operands in fragment order (element t of lane l, unit 64 w + 32 ub + l % 32,
// half h = l / 32):  w2b: W2 rows (k = 64 g + 32 h + 8 s + t) at
// (((((w*8 + g)*2 + ub)*4 + s)*64 + l)*8 + t;  w3b: W3 rows (k = 16 s + 8 h + t, zero
// for k >= d_out) at (((w*2 + ub)*NK3 + s)*64 + l)*8 + t.
__global__ void __launch_bounds__(256) mlp_dz_pack_kernel(const uint16_t* w2, const uint16_t* w3, int d_out, int nk3,
                                                          uint16_t* w2b, uint16_t* w3b) {
    const int64_t n2 = (int64_t)FI_DH * FI_DH, n3 = (int64_t)FI_DH * 16 * nk3;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n2 + n3;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < n2) {
            const int t = (int)(e & 7), l = (int)((e >> 3) & 63), s = (int)((e >> 9) & 3);
            const int ub = (int)((e >> 11) & 1), g = (int)((e >> 12) & 7), w = (int)(e >> 15);
            const int unit = 64 * w + 32 * ub + (l & 31), k = 64 * g + 32 * (l >> 5) + 8 * s + t;
            w2b[e] = w2[(size_t)unit * FI_DH + k];
        } else {
            const int64_t f = e - n2;
            const int t = (int)(f & 7), l = (int)((f >> 3) & 63);
            const int rest = (int)(f >> 9), s = rest % nk3, ub = (rest / nk3) & 1, w = (rest / nk3) >> 1;
            const int unit = 64 * w + 32 * ub + (l & 31), k = 16 * s + 8 * (l >> 5) + t;
            w3b[f] = k < d_out ? w3[(size_t)unit * d_out + k] : (uint16_t)0;
        }
    }
}

template <int NK3>
int launch_dz(const PdlpMlp& p, hipStream_t s) {
    const hipError_t e = hipFuncSetAttribute((const void*)mlp_dz_kernel<NK3>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)FI_H1);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
    const unsigned grid = (unsigned)((p.batch + FI_ROWS - 1) / FI_ROWS);
    mlp_dz_kernel<NK3><<<grid, FI_THREADS, FI_H1, s>>>(p);
    return hip_fail(hipGetLastError(), "mlp_dz_kernel launch");
}

template <int NK, bool ST>
int launch_infer_st(const PdlpMlpInfer& p, unsigned grid, hipStream_t s) {
    const hipError_t e = hipFuncSetAttribute((const void*)mlp_infer_kernel<NK, ST>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)FI_LDS);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
    mlp_infer_kernel<NK, ST><<<grid, FI_THREADS, FI_LDS, s>>>(p);
    return hip_fail(hipGetLastError(), "mlp_infer_kernel launch");
}

template <int NK>
int launch_infer(const PdlpMlpInfer& p, unsigned grid, hipStream_t s) {
    return p.h1 != nullptr ? launch_infer_st<NK, true>(p, grid, s) : launch_infer_st<NK, false>(p, grid, s);
}

// The weights in the fragment order of mlp_infer_kernel: element t of lane l's
// A operand (unit 64 w + 32 ub + l % 32, half h = l / 32) of
//   w1p: step s (k = 16 s + 8 h + t; zero for k >= d_in) at (((w*2 + ub)*NK + s)*64 + l)*8 + t,
//   w2p: k group g, step s (k = 64 g + 32 h + 8 s + t) at (((((w*8 + g)*2 + ub)*4 + s)*64 + l)*8 + t,
//   w3p: (channel ch, unit block blk) at ((blk*32 + ch)*2 + h)*8 + t, zero for ch >= d_out.
__global__ void __launch_bounds__(256) mlp_infer_pack_kernel(const uint16_t* w1, const uint16_t* w2,
                                                             const uint16_t* w3, int d_in, int k1, int d_out,
                                                             uint16_t* w1p, uint16_t* w2p, uint16_t* w3p) {
    const int64_t n1 = (int64_t)FI_DH * k1, n2 = (int64_t)FI_DH * FI_DH, n3 = (int64_t)(FI_DH / 16) * 32 * 16;
    const int nk = k1 / 16;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n1 + n2 + n3;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < n1) {
            const int t = (int)(e & 7), l = (int)((e >> 3) & 63);
            const int rest = (int)(e >> 9), s = rest % nk, ub = (rest / nk) & 1, w = (rest / nk) >> 1;
            const int unit = 64 * w + 32 * ub + (l & 31), k = 16 * s + 8 * (l >> 5) + t;
            w1p[e] = k < d_in ? w1[(size_t)k * FI_DH + unit] : (uint16_t)0;
        } else if (e < n1 + n2) {
            const int64_t f = e - n1;
            const int t = (int)(f & 7), l = (int)((f >> 3) & 63), s = (int)((f >> 9) & 3);
            const int ub = (int)((f >> 11) & 1), g = (int)((f >> 12) & 7), w = (int)(f >> 15);
            const int unit = 64 * w + 32 * ub + (l & 31), k = 64 * g + 32 * (l >> 5) + 8 * s + t;
            w2p[f] = w2[(size_t)k * FI_DH + unit];
        } else {
            const int64_t f = e - n1 - n2;                  // ((blk * 32 + ch) * 2 + h) * 8 + t
            const int t = (int)(f & 7), h = (int)((f >> 3) & 1), ch = (int)((f >> 4) & 31), blk = (int)(f >> 9);
            const int unit = 16 * blk + 8 * (t >> 2) + 4 * h + (t & 3);
            w3p[f] = ch < d_out ? w3[(size_t)unit * d_out + ch] : (uint16_t)0;
        }
    }
}

template <bool VEC, bool TA, bool TB>
void launch_gemm(const PdlpGemm& g, dim3 grid, int kc, hipStream_t s) {
    gemm_kernel<VEC, TA, TB><<<grid, GEMM_THREADS, 0, s>>>(g, kc);
}

unsigned grid1(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

int gemm(const PdlpGemm& g, hipStream_t s) {
    if (g.M < 0 || g.N < 0 || g.K < 0) return fail(PDLP_E_SHAPE, "gemm sizes must be >= 0");
    if (g.M == 0 || g.N == 0) return 0;
    if (!g.A || !g.B || (!g.C && !g.Cb)) return fail(PDLP_E_ARG, "gemm: A, B and an output are required");
    if (g.accumulate && !g.C) return fail(PDLP_E_ARG, "gemm: accumulate needs C");
    if (g.split_k > 1 && !g.work) return fail(PDLP_E_ARG, "gemm: split_k > 1 needs work");
    if (g.split_k > 1024) return fail(PDLP_E_LIMIT, "gemm: split_k at most 1024");
    const int stages = (g.K + BK - 1) / BK;
    const int parts = std::max(1, std::min(g.split_k, stages));
    const int kc = std::max(1, (stages + parts - 1) / parts) * BK;   // K values per part, a multiple of BK
    const int used = std::max(1, (g.K + kc - 1) / kc);               // parts that hold any K
    dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, used);
    const bool vec = g.lda % 8 == 0 && g.ldb % 8 == 0 && (reinterpret_cast<uintptr_t>(g.A) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(g.B) & 15) == 0;
    const int sel = (vec ? 4 : 0) | (g.trans_a ? 2 : 0) | (g.trans_b ? 1 : 0);
    switch (sel) {
        case 0: launch_gemm<false, false, false>(g, grid, kc, s); break;
        case 1: launch_gemm<false, false, true>(g, grid, kc, s); break;
        case 2: launch_gemm<false, true, false>(g, grid, kc, s); break;
        case 3: launch_gemm<false, true, true>(g, grid, kc, s); break;
        case 4: launch_gemm<true, false, false>(g, grid, kc, s); break;
        case 5: launch_gemm<true, false, true>(g, grid, kc, s); break;
        case 6: launch_gemm<true, true, false>(g, grid, kc, s); break;
        default: launch_gemm<true, true, true>(g, grid, kc, s); break;
    }
    int rc = hip_fail(hipGetLastError(), "gemm_kernel launch");
    if (!rc && used > 1) {
        const int64_t total = (int64_t)g.M * g.N;
        sum_parts_kernel<true><<<parts_grid(total), 256, 0, s>>>(g.work, total, used, nullptr, g);
        rc = hip_fail(hipGetLastError(), "sum_parts_kernel launch");
    }
    return rc;
}

// Column sums of a B x N gradient: with work, the rows are cut into chunks of at
// least 256 rows so the launch has about 1024 workgroups (a long batch with few columns).
int colsum(const float* dzf, const uint16_t* dzb, float* db, int B, int N, float* work, int64_t work_floats,
           hipStream_t s) {
    const int cblocks = (N + CS_COLS - 1) / CS_COLS;
    int chunks = 1;
    if (work != nullptr && N > 0) {
        const int64_t cap = std::min<int64_t>(work_floats / N, 1024);
        chunks = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)(1024 + cblocks - 1) / cblocks,
                                                                (int64_t)(B + 255) / 256, cap}));
    }
    const int rows_per = (B + chunks - 1) / chunks;
    colsum_kernel<<<dim3(cblocks, chunks), CS_COLS * CS_GROUPS, 0, s>>>(dzf, dzb, db, work, B, N, rows_per);
    if (chunks > 1) sum_parts_kernel<false><<<parts_grid(N), 256, 0, s>>>(work, N, chunks, db, PdlpGemm{});
    return hip_fail(hipGetLastError(), "colsum");
}

// K parts of a weight-gradient GEMM (M x N output, K = batch): about 512
// workgroups, at least 4 K stages per part, within the work space
int weight_split(int M, int N, int K, int64_t work_floats) {
    if (work_floats <= 0) return 1;
    const int64_t tiles = (int64_t)((M + BM - 1) / BM) * ((N + BN - 1) / BN);
    const int64_t stages = (K + BK - 1) / BK;
    const int64_t fit = work_floats / std::max<int64_t>(1, (int64_t)M * N);
    return (int)std::max<int64_t>(1, std::min<int64_t>({(512 + tiles - 1) / tiles, stages / 4, fit, 1024}));
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
    if (p->work_floats < 0 || (p->work_floats > 0 && p->work == nullptr))
        return fail(PDLP_E_ARG, "pdlp_mlp_bwd_bf16: work_floats > 0 needs work");
    hipStream_t s = as_stream(stream);
    const int B = p->batch, di = p->d_in, dh = p->d_h, dout = p->d_out;
    const int64_t total = (int64_t)B * dout;
    float* work = p->work_floats > 0 ? p->work : nullptr;
    const int64_t wf = work != nullptr ? p->work_floats : 0;
    zero1_kernel<<<1, 64, 0, s>>>(p->loss);
    const unsigned lg = (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 2048));
    loss_grad_kernel<<<lg, 256, 0, s>>>(p->out, p->target, p->out_weight, p->dout, p->dout_b, p->loss, total, dout,
                                        1.f / (float)total);
    int rc = hip_fail(hipGetLastError(), "mlp loss kernel");
    if (!rc) rc = colsum(p->dout, nullptr, p->gb3, B, dout, work, wf, s);
    // gw3 = h2^T dout ; dz2 = (dout w3^T) * [h2 > 0]
    auto wgrad = [&](PdlpGemm g) {
        g.split_k = weight_split(g.M, g.N, g.K, wf);
        g.work = work;
        return gemm(g, s);
    };
    if (!rc) rc = wgrad(make(dh, dout, B, p->h2, dh, 1, p->dout_b, dout, 0, nullptr, 0, nullptr, p->gw3, nullptr, dout));
    // the fused kernel for both masked products where it applies (d_h = 512, packed weights given)
    const bool fused = p->w2b != nullptr && p->w3b != nullptr && dh == FI_DH && dout % 8 == 0 && dout <= 32 &&
                       !((reinterpret_cast<uintptr_t>(p->dout_b) | reinterpret_cast<uintptr_t>(p->h1) |
                          reinterpret_cast<uintptr_t>(p->h2) | reinterpret_cast<uintptr_t>(p->dz1) |
                          reinterpret_cast<uintptr_t>(p->dz2) | reinterpret_cast<uintptr_t>(p->w2b) |
                          reinterpret_cast<uintptr_t>(p->w3b)) & 15);
    if (!rc && fused && B > 0) {
        rc = dout <= 16 ? launch_dz<1>(*p, s) : launch_dz<2>(*p, s);
        if (!rc) rc = colsum(nullptr, p->dz2, p->gb2, B, dh, work, wf, s);
        if (!rc) rc = wgrad(make(dh, dh, B, p->h1, dh, 1, p->dz2, dh, 0, nullptr, 0, nullptr, p->gw2, nullptr, dh));
    } else if (!rc) {
        // K = d_out is long and the output (batch x d_h) has few tiles: split K, with dout
        // (free once gb3 and gw3 are taken from it) as the work space of split_k * batch * d_h floats
        PdlpGemm g = make(B, dh, dout, p->dout_b, dout, 0, p->w3, dout, 1, nullptr, 0, p->h2, nullptr, p->dz2, dh);
        const int tiles = ((B + BM - 1) / BM) * ((dh + BN - 1) / BN);
        const int stages = (dout + BK - 1) / BK;
        g.split_k = std::max(1, std::min({dout / std::max(dh, 1), 256 / tiles, stages / 4}));
        g.work = p->dout;
        rc = gemm(g, s);
        if (!rc) rc = colsum(nullptr, p->dz2, p->gb2, B, dh, work, wf, s);
        // gw2 = h1^T dz2 ; dz1 = (dz2 w2^T) * [h1 > 0]
        if (!rc) rc = wgrad(make(dh, dh, B, p->h1, dh, 1, p->dz2, dh, 0, nullptr, 0, nullptr, p->gw2, nullptr, dh));
        if (!rc) rc = gemm(make(B, dh, dh, p->dz2, dh, 0, p->w2, dh, 1, nullptr, 0, p->h1, nullptr, p->dz1, dh), s);
    }
    if (!rc) rc = colsum(nullptr, p->dz1, p->gb1, B, dh, work, wf, s);
    // gw1 = x^T dz1
    if (!rc) rc = wgrad(make(di, dh, B, p->x, di, 1, p->dz1, dh, 0, nullptr, 0, nullptr, p->gw1, nullptr, dh));
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

int pdlp_mlp_infer_pack(const uint16_t* w1, const uint16_t* w2, const uint16_t* w3, int32_t d_in, int32_t d_h,
                        int32_t d_out, uint16_t* w1p, uint16_t* w2p, uint16_t* w3p, void* stream) {
    if (d_h != FI_DH) return fail(PDLP_E_SHAPE, "mlp infer: d_h must be 512");
    if (d_in <= 0 || d_in > FI_MAX_DIN || d_in % 8 != 0) return fail(PDLP_E_SHAPE, "mlp infer: d_in in 8..128, multiple of 8");
    if (d_out <= 0 || d_out > FI_MAX_DOUT) return fail(PDLP_E_SHAPE, "mlp infer: d_out in 1..32");
    if (!w1 || !w2 || !w3 || !w1p || !w2p || !w3p) return fail(PDLP_E_ARG, "mlp infer pack: NULL pointer");
    const int k1 = (d_in + 15) / 16 * 16;
    const int64_t total = (int64_t)FI_DH * k1 + (int64_t)FI_DH * FI_DH + (int64_t)(FI_DH / 16) * 32 * 16;
    mlp_infer_pack_kernel<<<grid1(total), 256, 0, as_stream(stream)>>>(w1, w2, w3, d_in, k1, d_out, w1p, w2p, w3p);
    return hip_fail(hipGetLastError(), "mlp_infer_pack_kernel launch");
}

int pdlp_mlp_dz_pack(const uint16_t* w2, const uint16_t* w3, int32_t d_h, int32_t d_out, uint16_t* w2b,
                     uint16_t* w3b, void* stream) {
    if (d_h != FI_DH) return fail(PDLP_E_SHAPE, "mlp dz pack: d_h must be 512");
    if (d_out <= 0 || d_out > 32 || d_out % 8 != 0) return fail(PDLP_E_SHAPE, "mlp dz pack: d_out in 8..32, multiple of 8");
    if (!w2 || !w3 || !w2b || !w3b) return fail(PDLP_E_ARG, "mlp dz pack: NULL pointer");
    const int nk3 = (d_out + 15) / 16;
    const int64_t total = (int64_t)FI_DH * FI_DH + (int64_t)FI_DH * 16 * nk3;
    mlp_dz_pack_kernel<<<grid1(total), 256, 0, as_stream(stream)>>>(w2, w3, d_out, nk3, w2b, w3b);
    return hip_fail(hipGetLastError(), "mlp_dz_pack_kernel launch");
}

int pdlp_mlp_infer_bf16(const PdlpMlpInfer* p, void* stream) {
    if (p == nullptr) return fail(PDLP_E_ARG, "PdlpMlpInfer is NULL");
    if (p->d_h != FI_DH) return fail(PDLP_E_SHAPE, "mlp infer: d_h must be 512");
    if (p->d_in <= 0 || p->d_in > FI_MAX_DIN || p->d_in % 8 != 0)
        return fail(PDLP_E_SHAPE, "mlp infer: d_in in 8..128, multiple of 8");
    if (p->d_out <= 0 || p->d_out > FI_MAX_DOUT) return fail(PDLP_E_SHAPE, "mlp infer: d_out in 1..32");
    if (p->rows < 0) return fail(PDLP_E_SHAPE, "mlp infer: rows >= 0");
    if (p->rows == 0) return 0;
    if (!p->x || !p->w1p || !p->w2p || !p->w3p || !p->b1 || !p->b2 || !p->b3 || !p->out)
        return fail(PDLP_E_ARG, "mlp infer: NULL pointer");
    if ((p->h1 == nullptr) != (p->h2 == nullptr)) return fail(PDLP_E_ARG, "mlp infer: h1 and h2 go together");
    if ((reinterpret_cast<uintptr_t>(p->h1) | reinterpret_cast<uintptr_t>(p->h2)) & 15)
        return fail(PDLP_E_ARG, "mlp infer: h1 and h2 must be 16-byte aligned");
    // 16-byte operand loads: x rows (d_in a multiple of 8) and the packed weights
    if ((reinterpret_cast<uintptr_t>(p->x) | reinterpret_cast<uintptr_t>(p->w1p) |
         reinterpret_cast<uintptr_t>(p->w2p) | reinterpret_cast<uintptr_t>(p->w3p) |
         reinterpret_cast<uintptr_t>(p->b1) | reinterpret_cast<uintptr_t>(p->b2)) & 15)
        return fail(PDLP_E_ARG, "mlp infer: x, the packed weights, b1 and b2 must be 16-byte aligned");
    const unsigned grid = (unsigned)((p->rows + FI_ROWS - 1) / FI_ROWS);
    switch ((p->d_in + 15) / 16) {
        case 1: return launch_infer<1>(*p, grid, as_stream(stream));
        case 2: return launch_infer<2>(*p, grid, as_stream(stream));
        case 3: return launch_infer<3>(*p, grid, as_stream(stream));
        case 4: return launch_infer<4>(*p, grid, as_stream(stream));
        case 5: return launch_infer<5>(*p, grid, as_stream(stream));
        case 6: return launch_infer<6>(*p, grid, as_stream(stream));
        case 7: return launch_infer<7>(*p, grid, as_stream(stream));
        default: return launch_infer<8>(*p, grid, as_stream(stream));
    }
}

int pdlp_cast_bf16(const float* x, uint16_t* y, int64_t n, void* stream) {
    if (n < 0) return fail(PDLP_E_ARG, "cast: n >= 0");
    if (n == 0) return 0;
    if (!x || !y) return fail(PDLP_E_ARG, "cast: NULL pointer");
    cast_kernel<<<grid1(n), 256, 0, as_stream(stream)>>>(x, y, n);
    return hip_fail(hipGetLastError(), "cast_kernel launch");
}

}  // extern "C"
