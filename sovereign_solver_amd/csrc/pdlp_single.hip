// Single-instance restarted PDHG on gfx950.
//
// Replaces the per-iteration Python dispatch of solver/lp/pdhg.py:209-244 and
// solver/lp/pdhg_gpu.py:196-282, 378-427.  Per iteration two kernels run:
//   primal: one K^T row (= one column of K) per group of TPR lanes; the
//           gathered K^T y, the step tau_j (from the same row's |values|), the
//           box projection, the extrapolation and the ergodic sum are fused;
//   dual:   one K row per group; gathers xbar, forms sigma_i, projects y.
// A group is 1..64 lanes of a wavefront, or a whole workgroup (128..1024
// threads) for long rows.  Every check_every iterations four small kernels
// evaluate the reference KKT of the window average in FP64 (two partial-sum
// passes, one finalize that decides termination/restart, one apply that
// restarts to the average and folds the FP32 partial sums).  The control
// block lives on the device; kernels return at once after the status leaves
// RUNNING, so blocks of iterations can be enqueued (and captured in a HIP
// graph) ahead of the host's status poll.

#include "pdlp_common.hpp"

#include <algorithm>
#include <type_traits>

namespace pdlp {

namespace {

template <typename T>
__device__ __forceinline__ T clamp_box(T v, const T* lo, const T* hi, int j) {
    if (lo != nullptr) v = v < lo[j] ? lo[j] : v;
    else v = v < T(0) ? T(0) : v;
    if (hi != nullptr) v = v > hi[j] ? hi[j] : v;
    return v;
}

// Row of this group and the lane inside the group.
template <int TPR>
__device__ __forceinline__ void group_of(int remap, int& row, int& lane) {
    if constexpr (TPR > 64) {
        row = block_index(remap);
        lane = threadIdx.x;
    } else {
        row = block_index(remap) * (blockDim.x / TPR) + threadIdx.x / TPR;
        lane = threadIdx.x % TPR;
    }
}

template <int TPR, bool EXACT, typename T, typename G>
__device__ __forceinline__ void any_row_dot(const int32_t* rp, const int32_t* ci, const T* val, int row, int lane,
                                            G gather, T& dot, T& asum) {
    if constexpr (TPR > 64) {
        __shared__ T lds[2 * 16];
        row_dot_block<TPR, T>(rp, ci, val, row, gather, dot, asum, lds);
    } else {
        row_dot<TPR, EXACT, T>(rp, ci, val, row, lane, gather, dot, asum);
    }
}

// ---------------------------------------------------------------- primal step
template <typename T, int TPR, bool EXACT>
__global__ void __launch_bounds__(1024) primal_kernel(PdlpDev d, int remap) {
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const T tau_scale = (T)ctrl->tau_scale;
    const T* __restrict__ y = (const T*)d.y;
    int j, lane;
    group_of<TPR>(remap, j, lane);
    if (j >= d.n) return;                               // whole groups leave together
    const T* c = (const T*)d.c;
    const T* lo = (const T*)d.lo;
    const T* hi = (const T*)d.hi;
    T* x = (T*)d.x;
    T* xbar = (T*)d.xbar;
    // the column's own data first (group lane 0), so its loads are in flight with the dot product's
    const bool own = lane == 0;
    const T xj = own ? x[j] : T(0), cj = own ? c[j] : T(0);
    const T lj = (own && lo != nullptr) ? lo[j] : T(0), hj = (own && hi != nullptr) ? hi[j] : T(0);
    const float psf = (own && d.x_psum != nullptr) ? d.x_psum[j] : 0.f;
    const double psd = (own && d.x_psum == nullptr) ? d.x_sum[j] : 0.0;
    T aty, asum;
    any_row_dot<TPR, EXACT, T>(d.t_row_ptr, d.t_col_idx, (const T*)d.t_val, j, lane,
                               [&](int k) { return y[k]; }, aty, asum);
    if (lane != 0) return;
    const T colsum = asum > T(1e-4) ? asum : T(1e-4);
    T tau = T(0.999) / colsum;
    if (tau_scale != T(1)) tau *= tau_scale;
    T xn;
    if constexpr (EXACT) xn = add_rn(xj, -mul_rn(tau, add_rn(cj, -aty)));
    else xn = xj - tau * (cj - aty);
    // the box projection of clamp_box with the preloaded bounds
    if (lo != nullptr) xn = xn < lj ? lj : xn;
    else xn = xn < T(0) ? T(0) : xn;
    if (hi != nullptr) xn = xn > hj ? hj : xn;
    if constexpr (EXACT) xbar[j] = add_rn(mul_rn(T(2), xn), -xj);
    else xbar[j] = T(2) * xn - xj;
    x[j] = xn;
    if (d.x_psum != nullptr) d.x_psum[j] = psf + (float)xn;
    else d.x_sum[j] = psd + (double)xn;
}

// Primal step over the SELL-64 copy of K^T: one lane per column, no column
// pointer to read; slice s = 64 columns, entries column-major, so every entry
// load is a whole line.  The column's entries are summed in CSR order (padding
// adds zeros), which is the one-lane CSR kernel's arithmetic.
template <typename T>
__global__ void __launch_bounds__(1024) primal_sell_kernel(PdlpDev d, int remap) {
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const T tau_scale = (T)ctrl->tau_scale;
    const int j = block_index(remap) * blockDim.x + threadIdx.x;
    const int s = __builtin_amdgcn_readfirstlane(j >> 6);
    if (s >= d.t_sell_slices) return;               // whole wavefronts leave together
    const bool live = j < d.n;
    const int jj = live ? j : d.n - 1;
    const T* __restrict__ y = (const T*)d.y;
    const T* c = (const T*)d.c;
    const T* lo = (const T*)d.lo;
    const T* hi = (const T*)d.hi;
    // the column's own data first, in flight with the entries
    const T xj = ((const T*)d.x)[jj], cj = c[jj];
    const T lj = lo != nullptr ? lo[jj] : T(0), hj = hi != nullptr ? hi[jj] : T(0);
    const float psf = d.x_psum != nullptr ? d.x_psum[jj] : 0.f;
    const double psd = d.x_psum == nullptr ? d.x_sum[jj] : 0.0;
    const int beg = __builtin_amdgcn_readfirstlane(d.t_sell_ptr[s]);
    const int width = __builtin_amdgcn_readfirstlane((d.t_sell_ptr[s + 1] - beg) >> 6);
    const int32_t* __restrict__ idx = d.t_sell_idx + beg + (j & 63);
    const T* __restrict__ val = (const T*)d.t_sell_val + beg + (j & 63);
    const T kc = (T)d.k_const;                      // nonzero: every entry holds kc (no values read)
    T aty = T(0), asum = T(0);
    int k = 0;
    for (; k + kRowBatch <= width; k += kRowBatch) {
        T a[kRowBatch];
        int r[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            a[u] = kc != T(0) ? kc : val[64 * (k + u)];
            r[u] = idx[64 * (k + u)];
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            aty += a[u] * y[r[u]];
            asum += fabs(a[u]);
        }
    }
    for (; k < width; ++k) {
        const T a = kc != T(0) ? kc : val[64 * k];
        aty += a * y[idx[64 * k]];
        asum += fabs(a);
    }
    if (!live) return;
    const T colsum = asum > T(1e-4) ? asum : T(1e-4);
    T tau = T(0.999) / colsum;
    if (tau_scale != T(1)) tau *= tau_scale;
    T xn = xj - tau * (cj - aty);
    if (lo != nullptr) xn = xn < lj ? lj : xn;
    else xn = xn < T(0) ? T(0) : xn;
    if (hi != nullptr) xn = xn > hj ? hj : xn;
    ((T*)d.xbar)[j] = T(2) * xn - xj;
    ((T*)d.x)[j] = xn;
    if (d.x_psum != nullptr) d.x_psum[j] = psf + (float)xn;
    else d.x_sum[j] = psd + (double)xn;
}

// The same step with two slices per wavefront (columns 128 w .. 128 w + 127 for
// wavefront w, lane l owning columns 128 w + l and 128 w + 64 + l): both columns'
// own data, row indices and gathers are issued together, so a lane keeps twice
// the loads in flight; each column is summed in the same order as above (same
// bits).
template <typename T>
__global__ void __launch_bounds__(1024) primal_sell2_kernel(PdlpDev d, int remap) {
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const T tau_scale = (T)ctrl->tau_scale;
    const int wv = block_index(remap) * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    const int s0 = __builtin_amdgcn_readfirstlane(2 * wv);
    if (s0 >= d.t_sell_slices) return;              // whole wavefronts leave together
    const int ns = s0 + 1 < d.t_sell_slices ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const T* __restrict__ y = (const T*)d.y;
    const T* c = (const T*)d.c;
    const T* lo = (const T*)d.lo;
    const T* hi = (const T*)d.hi;
    const T kc = (T)d.k_const;                      // nonzero: every entry holds kc (no values read)
    int j[2], beg[2], width[2];
    bool live[2];
    T xj[2], cj[2], lj[2], hj[2];
    float psf[2];
    double psd[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        j[u] = (s0 + u) * 64 + lane;
        live[u] = u < ns && j[u] < d.n;
        const int jj = j[u] < d.n ? j[u] : d.n - 1;
        xj[u] = ((const T*)d.x)[jj];
        cj[u] = c[jj];
        lj[u] = lo != nullptr ? lo[jj] : T(0);
        hj[u] = hi != nullptr ? hi[jj] : T(0);
        psf[u] = d.x_psum != nullptr ? d.x_psum[jj] : 0.f;
        psd[u] = d.x_psum == nullptr ? d.x_sum[jj] : 0.0;
        const int sl = u < ns ? s0 + u : s0;
        beg[u] = __builtin_amdgcn_readfirstlane(d.t_sell_ptr[sl]);
        width[u] = u < ns ? __builtin_amdgcn_readfirstlane((d.t_sell_ptr[sl + 1] - beg[u]) >> 6) : 0;
    }
    const int32_t* __restrict__ idx0 = d.t_sell_idx + beg[0] + lane;
    const int32_t* __restrict__ idx1 = d.t_sell_idx + beg[1] + lane;
    const T* __restrict__ val0 = (const T*)d.t_sell_val + beg[0] + lane;
    const T* __restrict__ val1 = (const T*)d.t_sell_val + beg[1] + lane;
    const int wmax = width[0] > width[1] ? width[0] : width[1];
    T aty[2] = {T(0), T(0)}, asum[2] = {T(0), T(0)};
    for (int k = 0; k < wmax; k += kRowBatch) {
        T a[2][kRowBatch];
        int r[2][kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            const bool in0 = k + u < width[0], in1 = k + u < width[1];
            a[0][u] = in0 ? (kc != T(0) ? kc : val0[64 * (k + u)]) : T(0);
            r[0][u] = in0 ? idx0[64 * (k + u)] : 0;
            a[1][u] = in1 ? (kc != T(0) ? kc : val1[64 * (k + u)]) : T(0);
            r[1][u] = in1 ? idx1[64 * (k + u)] : 0;
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            if (k + u < width[0]) {
                aty[0] += a[0][u] * y[r[0][u]];
                asum[0] += fabs(a[0][u]);
            }
            if (k + u < width[1]) {
                aty[1] += a[1][u] * y[r[1][u]];
                asum[1] += fabs(a[1][u]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        if (!live[u]) continue;
        const T colsum = asum[u] > T(1e-4) ? asum[u] : T(1e-4);
        T tau = T(0.999) / colsum;
        if (tau_scale != T(1)) tau *= tau_scale;
        T xn = xj[u] - tau * (cj[u] - aty[u]);
        if (lo != nullptr) xn = xn < lj[u] ? lj[u] : xn;
        else xn = xn < T(0) ? T(0) : xn;
        if (hi != nullptr) xn = xn > hj[u] ? hj[u] : xn;
        ((T*)d.xbar)[j[u]] = T(2) * xn - xj[u];
        ((T*)d.x)[j[u]] = xn;
        if (d.x_psum != nullptr) d.x_psum[j[u]] = psf[u] + (float)xn;
        else d.x_sum[j[u]] = psd[u] + (double)xn;
    }
}

// ------------------------------------------------------------------ dual step
// Dual update of row i from ax = K_i xbar and asum = sum_j |K_ij|.
template <typename T, bool EXACT>
__device__ __forceinline__ void dual_update(const PdlpDev& d, T sigma_scale, int i, T ax, T asum) {
    const T* b = (const T*)d.b;
    T* y = (T*)d.y;
    const T rowsum = asum > T(1e-4) ? asum : T(1e-4);
    T sigma = T(0.999) / rowsum;
    if (sigma_scale != T(1)) sigma *= sigma_scale;
    T yn;
    if constexpr (EXACT) yn = add_rn(y[i], mul_rn(sigma, add_rn(b[i], -ax)));
    else yn = y[i] + sigma * (b[i] - ax);
    if (d.row_kind != nullptr && d.row_kind[i] == PDLP_ROW_LE && yn > T(0)) yn = T(0);
    y[i] = yn;
    if (d.y_psum != nullptr) d.y_psum[i] += (float)yn;
    else d.y_sum[i] += (double)yn;
}

// A row's own data for its dual update, loaded ahead of the row's dot
// product so that the loads overlap (only the thread that applies the update
// loads: own).
template <typename T>
struct RowOwn {
    T y, b;
    bool le;
    float psf;
    double psd;
};

template <typename T>
__device__ __forceinline__ RowOwn<T> row_own(const PdlpDev& d, int i, bool own) {
    RowOwn<T> o;
    o.y = own ? ((const T*)d.y)[i] : T(0);
    o.b = own ? ((const T*)d.b)[i] : T(0);
    o.le = own && d.row_kind != nullptr && d.row_kind[i] == PDLP_ROW_LE;
    o.psf = (own && d.y_psum != nullptr) ? d.y_psum[i] : 0.f;
    o.psd = (own && d.y_psum == nullptr) ? d.y_sum[i] : 0.0;
    return o;
}

// dual_update with the row's own data already loaded (same arithmetic)
template <typename T, bool EXACT>
__device__ __forceinline__ void dual_update_own(const PdlpDev& d, T sigma_scale, int i, T ax, T asum,
                                                const RowOwn<T>& o) {
    const T rowsum = asum > T(1e-4) ? asum : T(1e-4);
    T sigma = T(0.999) / rowsum;
    if (sigma_scale != T(1)) sigma *= sigma_scale;
    T yn;
    if constexpr (EXACT) yn = add_rn(o.y, mul_rn(sigma, add_rn(o.b, -ax)));
    else yn = o.y + sigma * (o.b - ax);
    if (o.le && yn > T(0)) yn = T(0);
    ((T*)d.y)[i] = yn;
    if (d.y_psum != nullptr) d.y_psum[i] = o.psf + (float)yn;
    else d.y_sum[i] = o.psd + (double)yn;
}

// CSR dual step over all rows (use_list = 0) or over csr_rows (the rows the
// SELL slices leave).  Block 0 counts the iteration.
template <typename T, int TPR, bool EXACT>
__global__ void __launch_bounds__(1024) dual_kernel(PdlpDev d, int remap, int use_list) {
    PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // single writer; read only by later kernels
        ctrl->iter += 1;
        ctrl->window += 1;
    }
    const T sigma_scale = (T)ctrl->sigma_scale;
    const T* __restrict__ xbar = (const T*)d.xbar;
    int i, lane;
    group_of<TPR>(remap, i, lane);
    if (use_list) {
        if (i >= d.n_csr_rows) return;
        i = d.csr_rows[i];
    } else if (i >= d.m) {
        return;
    }
    // the row's own data first (group lane 0), so its loads are in flight with the dot product's
    const RowOwn<T> o = row_own<T>(d, i, lane == 0);
    T ax, asum;
    any_row_dot<TPR, EXACT, T>(d.row_ptr, d.col_idx, (const T*)d.val, i, lane,
                               [&](int k) { return xbar[k]; }, ax, asum);
    if (lane != 0) return;
    dual_update_own<T, EXACT>(d, sigma_scale, i, ax, asum, o);
}

// SELL-64 dual step: one wavefront per slice, one row per lane; entry k of
// every lane is read in one coalesced 256-byte (FP32) access, and the
// gathers of neighbouring columns share cache lines.  Entries are summed in
// CSR order, as the one-lane CSR kernel does.
template <typename T>
__global__ void __launch_bounds__(256) dual_sell_kernel(PdlpDev d) {
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const int s = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    if (s >= d.sell_slices) return;
    const int lane = threadIdx.x & 63;
    const T sigma_scale = (T)ctrl->sigma_scale;
    const T* __restrict__ xbar = (const T*)d.xbar;
    const T* __restrict__ val = (const T*)d.sell_val;
    const int32_t* __restrict__ col = d.sell_col;
    const int beg = d.sell_ptr[s], width = (d.sell_ptr[s + 1] - beg) >> 6;
    T ax = T(0), asum = T(0);
    int q = beg + lane;
#pragma unroll 8
    for (int k = 0; k < width; ++k, q += 64) {
        const T a = val[q];
        ax += a * xbar[col[q]];
        asum += fabs(a);
    }
    const int i = d.sell_rows[s * 64 + lane];
    if (i < 0) return;
    dual_update<T, false>(d, sigma_scale, i, ax, asum);
}

// SELL-64 slices and CSR rows in one launch, workgroups of NT threads.  The
// first slices * (64 / R) workgroups each take R consecutive rows of a slice:
// a wavefront's lane l works on row l % R of them at entry phase l / R, so one
// access reads entry k .. k + 64/R - 1 of the R rows as 64/R contiguous pieces
// of R values (and their gathers of xbar share cache lines); the slice's
// entries are split into NT / 64 contiguous ranges, one per wavefront.  The
// partial sums meet in LDS and are added in a fixed order (wavefront, then
// phase).  The remaining workgroups take NT / TC rows of csr_rows each, TC
// threads per row (TC / 64 wavefronts; their sums meet in LDS in wavefront
// order).  Block 0 counts the iteration.  SPMV: the same sums of K vx stored in
// vy (pdlp_spmv_dev), no dual update and no iteration state.
template <typename T, int NT, int R, int TC, bool SPMV = false>
__global__ void __launch_bounds__(NT) dual_sell_wg_kernel(PdlpDev d, const T* __restrict__ vx, T* __restrict__ vy) {
    constexpr int NW = NT / 64, P = 64 / R;
    static_assert(R >= 1 && R <= 64 && 64 % R == 0, "rows per workgroup divide the slice");
    static_assert(TC >= 64 && TC <= NT && NT % TC == 0 && TC % 64 == 0, "threads per CSR row");
    __shared__ T part[2][NW][64];
    T sigma_scale = T(1);
    if constexpr (!SPMV) {
        PdlpCtrl* ctrl = d.ctrl;
        if (ctrl->status != PDLP_RUNNING) return;
        if (blockIdx.x == 0 && threadIdx.x == 0) {   // single writer; read only by later kernels
            ctrl->iter += 1;
            ctrl->window += 1;
        }
        sigma_scale = (T)ctrl->sigma_scale;
    }
    const T* __restrict__ xbar = SPMV ? vx : (const T*)d.xbar;
    const int b = blockIdx.x;
    if (b < d.sell_slices * P) {
        const int sl = b / P, rb = (b % P) * R;
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int r = lane % R, ph = lane / R;
        const T* __restrict__ val = (const T*)d.sell_val;
        const int32_t* __restrict__ col = d.sell_col;
        const int beg = d.sell_ptr[sl], width = (d.sell_ptr[sl + 1] - beg) >> 6;
        // the rows' own data, by the threads that apply the updates
        const int ir = threadIdx.x < R ? d.sell_rows[sl * 64 + rb + threadIdx.x] : -1;
        const RowOwn<T> o = row_own<T>(d, ir, !SPMV && ir >= 0);
        const int kc = (width + NW - 1) / NW;
        const int k0 = min(width, w * kc), k1 = min(width, k0 + kc);
        const T kv = (T)d.k_const;                  // nonzero: every entry holds kv (no values read)
        // nonzero: the slice's rows step their columns by st (only each row's first index read)
        const int st = d.sell_stride != nullptr ? d.sell_stride[sl] : 0;
        T ax = T(0), asum = T(0);
        int q = beg + (k0 + ph) * 64 + rb + r;
        if (st != 0) {
            // a lane without a row (the slice's padding lanes) gathers column 0
            const bool row = d.sell_rows[sl * 64 + rb + r] >= 0;
            const int c0 = row ? col[beg + rb + r] : 0, dc = row ? st : 0;
            int cc = c0 + (k0 + ph) * dc;
            if (kv != T(0)) {
#pragma unroll 8
                for (int k = k0 + ph; k < k1; k += P, cc += P * dc) {
                    ax += kv * xbar[cc];
                    asum += fabs(kv);
                }
            } else {
#pragma unroll 8
                for (int k = k0 + ph; k < k1; k += P, q += 64 * P, cc += P * dc) {
                    const T a = val[q];
                    ax += a * xbar[cc];
                    asum += fabs(a);
                }
            }
        } else if (kv != T(0)) {
#pragma unroll 8
            for (int k = k0 + ph; k < k1; k += P, q += 64 * P) {
                ax += kv * xbar[col[q]];
                asum += fabs(kv);
            }
        } else {
#pragma unroll 8
            for (int k = k0 + ph; k < k1; k += P, q += 64 * P) {
                const T a = val[q];
                ax += a * xbar[col[q]];
                asum += fabs(a);
            }
        }
        part[0][w][lane] = ax;
        part[1][w][lane] = asum;
        __syncthreads();
        if (threadIdx.x < R) {
            T tx = T(0), ta = T(0);
#pragma unroll
            for (int v = 0; v < NW; ++v)
#pragma unroll
                for (int f = 0; f < P; ++f) {
                    tx += part[0][v][f * R + threadIdx.x];
                    ta += part[1][v][f * R + threadIdx.x];
                }
            if (ir >= 0) {
                if constexpr (SPMV) vy[ir] = tx;
                else dual_update_own<T, false>(d, sigma_scale, ir, tx, ta, o);
            }
        }
        return;
    }
    // CSR rows: every thread reaches the barrier below, rows past the end included
    const int rr = (b - d.sell_slices * P) * (NT / TC) + (int)threadIdx.x / TC;
    const int tc = threadIdx.x % TC;
    const int i = rr < d.n_csr_rows ? d.csr_rows[rr] : -1;
    const RowOwn<T> o = row_own<T>(d, i, !SPMV && i >= 0 && tc == 0);
    T ax = T(0), asum = T(0);
    if (i >= 0) {
        const int beg = d.row_ptr[i], end = d.row_ptr[i + 1];
        const int32_t* __restrict__ ci = d.col_idx;
        const T* __restrict__ val = (const T*)d.val;
        const T kv = (T)d.k_const;
        const bool uni = kv != T(0);
        // nonzero: the row's columns step by rs from its first (the only index read)
        const int rs = d.csr_stride != nullptr ? d.csr_stride[rr] : 0;
        const int c0 = rs != 0 && end > beg ? ci[beg] : 0;
        for (int k0 = beg + tc; k0 < end; k0 += TC * kRowBatch) {
            T a[kRowBatch];
            int c[kRowBatch];
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) {
                const int k = k0 + u * TC;
                a[u] = k < end ? (uni ? kv : val[k]) : T(0);
                c[u] = k < end ? (rs != 0 ? c0 + (k - beg) * rs : ci[k]) : 0;
            }
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) {
                if (k0 + u * TC < end) {
                    ax += a[u] * xbar[c[u]];
                    asum += fabs(a[u]);
                }
            }
        }
    }
    ax = group_sum<kWave>(ax);
    asum = group_sum<kWave>(asum);
    if constexpr (TC > 64) {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            part[0][w][0] = ax;
            part[1][w][0] = asum;
        }
        __syncthreads();
        if (tc == 0) {
            ax = T(0);
            asum = T(0);
#pragma unroll
            for (int v = 0; v < TC / 64; ++v) {
                ax += part[0][w + v][0];
                asum += part[1][w + v][0];
            }
        }
    }
    if (i >= 0 && tc == 0) {
        if constexpr (SPMV) vy[i] = ax;
        else dual_update_own<T, false>(d, sigma_scale, i, ax, asum, o);
    }
}

// ------------------------------------------------------------- KKT partials
// A check evaluates one or two candidates: src 0 = current iterate, src 1 =
// window average ((FP64 sum + FP32 partial sum) / window).  Each workgroup
// writes NP FP64 sums to partials[((src*2 + part)*BLOCKS + blk)*NP + k], part
// 0 = rows, 1 = columns; k = 0 rp^2, 1 rd^2, 2 primal objective, 3 dual
// objective, 4 |dx|^2 and 5 |dy|^2: the distance of the candidate from the
// last restart point in the step-size metric (sum dx_j^2 / tau_hat_j and
// sum dy_i^2 / sigma_hat_i), used by the adaptive primal weight; then the
// Farkas test of the candidate's y (box form): 6 b_shift'y, 7 the maximum of
// (K'y)'(x - lo_shift) over the shifted box, 8 sum |b_shift_i y_i| and 9 the
// sum of the |terms| of 7.  Then the unboundedness test of the ray estimate
// d = x - (window average), in the average's launches when `ray` is set
// (zero otherwise), all in the original LP's units: 10 the squared violations
// of K d against the rows' recession cone (= 0, <= 0), 11 the squared row sums
// of |K_ij d_j|, 12 c'd, 13 |y|^2 of the average, 14 the squared violations of
// d against the bounds' recession cone, 15 |c - K'y|^2 of the average.

constexpr int NP = 16;
constexpr int NP_BASE = 10;   // the sums of a check without the unboundedness test
constexpr double kRayEps = 1e-6;   // normalised violation a ray of decrease must reach

__device__ __forceinline__ double* part_ptr(const PdlpDev& d, int src, int part, int blk) {
    return d.partials + ((size_t)(src * 2 + part) * PDLP_PARTIAL_BLOCKS + blk) * NP;
}

__device__ __forceinline__ double avg_of(const double* s, const float* ps, int k, double w) {
    return ps != nullptr ? (s[k] + (double)ps[k]) / w : s[k] / w;
}

// Largest value of g * z over z in [lo - lo_shift, span]: z >= 0 up to span when the
// lower bound is finite, z <= span without one; +inf when unbounded (then y
// certifies nothing).
__device__ __forceinline__ double farkas_support(double g, double span, bool lo_inf) {
    if (g > 0.0) return isfinite(span) ? g * span : INFINITY;
    if (g < 0.0 && lo_inf) return INFINITY;
    return 0.0;
}

// (K d)_i and sum_j |K_ij d_j| of one row, by TPR lanes (result in every lane).
template <int TPR, typename V, typename G>
__device__ __forceinline__ void row_ray(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                        const V* __restrict__ val, int row, int lane, G dir, double& kd,
                                        double& kabs) {
    double s = 0.0, a = 0.0;
    for (int k = rp[row] + lane; k < rp[row + 1]; k += TPR) {
        const double t = (double)val[k] * dir(ci[k]);
        s += t;
        a += fabs(t);
    }
    if constexpr (TPR > 1) {
        s = group_sum<TPR>(s);
        a = group_sum<TPR>(a);
    }
    kd = s;
    kabs = a;
}

template <typename T, int TPR, bool RAY>
__global__ void __launch_bounds__(256) kkt_rows_kernel(PdlpDev d, int src) {
    constexpr int NS = RAY ? NP : NP_BASE;
    __shared__ double lds[16 * NS];
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const double w = src ? (double)ctrl->window : 1.0;
    const T* __restrict__ v = (const T*)d.val;
    const T* __restrict__ x = (const T*)d.x;
    const T* __restrict__ y = (const T*)d.y;
    const T* ylast = (const T*)d.y_last;
    const int gpb = blockDim.x / TPR, lane = threadIdx.x % TPR;
    double acc[NS] = {};
    for (int i = blockIdx.x * gpb + threadIdx.x / TPR; i < d.m; i += gridDim.x * gpb) {
        double ax, asum;
        if (src) row_dot<TPR, false, double>(d.row_ptr, d.col_idx, v, i, lane,
                                             [&](int k) { return avg_of(d.x_sum, d.x_psum, k, w); }, ax, asum);
        else row_dot<TPR, false, double>(d.row_ptr, d.col_idx, v, i, lane,
                                         [&](int k) { return (double)x[k]; }, ax, asum);
        double kd = 0.0, kabs = 0.0;
        if constexpr (RAY) row_ray<TPR>(d.row_ptr, d.col_idx, v, i, lane,
                              [&](int k) { return (double)x[k] - avg_of(d.x_sum, d.x_psum, k, w); }, kd, kabs);
        if (lane == 0) {
            const bool le = d.row_kind != nullptr && d.row_kind[i] == PDLP_ROW_LE;
            double res = ax - d.b64[i];
            if (le && res < 0.0) res = 0.0;
            const double yi = src ? avg_of(d.y_sum, d.y_psum, i, w) : (double)y[i];
            if (d.row_scale != nullptr) {              // residual and dual of the original row
                const double dr = d.row_scale[i];
                res /= dr;
                const double yo = yi * dr;
                if (le && yo > 0.0) acc[1] += yo * yo;
            } else if (le && yi > 0.0) {
                acc[1] += yi * yi;                      // slack column reduced cost -y_i
            }
            acc[0] += res * res;
            if constexpr (RAY) {
                const double dr = d.row_scale != nullptr ? d.row_scale[i] : 1.0;
                const double viol = (le && kd < 0.0 ? 0.0 : kd) / dr, ka = kabs / dr, yo = yi * dr;
                acc[10] += viol * viol;
                acc[11] += ka * ka;
                acc[13] += yo * yo;
            }
            const double by = d.b_shift[i] * yi;
            acc[3] += by;
            acc[6] += by;
            acc[8] += fabs(by);
            if (ylast != nullptr) {
                const double dy = yi - (double)ylast[i];
                acc[5] += dy * dy * (asum > 1e-4 ? asum : 1e-4) / 0.999;
            }
        }
    }
    block_sum<NS>(acc, lds);
    if (threadIdx.x == 0) {
        double* p = part_ptr(d, src, 0, blockIdx.x);
        for (int k = 0; k < NP; ++k) p[k] = k < NS ? acc[k] : 0.0;
    }
}

template <typename T, int TPR, bool RAY>
__global__ void __launch_bounds__(256) kkt_cols_kernel(PdlpDev d, int src) {
    constexpr int NS = RAY ? NP : NP_BASE;
    __shared__ double lds[16 * NS];
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const double w = src ? (double)ctrl->window : 1.0;
    const T* __restrict__ tv = (const T*)d.t_val;
    const T* __restrict__ x = (const T*)d.x;
    const T* __restrict__ y = (const T*)d.y;
    const T* xlast = (const T*)d.x_last;
    const int gpb = blockDim.x / TPR, lane = threadIdx.x % TPR;
    double acc[NS] = {};
    for (int j = blockIdx.x * gpb + threadIdx.x / TPR; j < d.n; j += gridDim.x * gpb) {
        double aty, asum;
        if (src) row_dot<TPR, false, double>(d.t_row_ptr, d.t_col_idx, tv, j, lane,
                                             [&](int k) { return avg_of(d.y_sum, d.y_psum, k, w); }, aty, asum);
        else row_dot<TPR, false, double>(d.t_row_ptr, d.t_col_idx, tv, j, lane,
                                         [&](int k) { return (double)y[k]; }, aty, asum);
        if (lane == 0) {
            const double r = d.c64[j] - aty;
            const double span = d.ub_span != nullptr ? d.ub_span[j] : INFINITY;
            const double rneg = r < 0.0 ? r : 0.0;
            if (isfinite(span)) {
                acc[3] += span * rneg;                   // upper-bound row dual w_j = min(r_j, 0)
            } else {
                const double ro = d.col_scale != nullptr ? rneg / d.col_scale[j] : rneg;
                acc[1] += ro * ro;                       // reduced cost of the original column
            }
            // Farkas test: max over x - lo_shift in [lo - lo_shift, span] of (K'y)_j (x - lo_shift)_j
            const bool lo_inf = d.lo != nullptr && isinf((double)((const T*)d.lo)[j]);
            const double u = farkas_support(aty, span, lo_inf);
            acc[7] += u;
            acc[9] += fabs(u);
            const double xj = src ? avg_of(d.x_sum, d.x_psum, j, w) : (double)x[j];
            const double sh = d.lo_shift != nullptr ? d.lo_shift[j] : 0.0;
            acc[2] += d.c64[j] * (xj - sh);
            if constexpr (RAY) {
                const double dj = (double)x[j] - xj, dc = d.col_scale != nullptr ? d.col_scale[j] : 1.0;
                const double dor = dj * dc, ro = r / dc;
                double bv = 0.0;                         // d must keep finite bounds' sides
                if (!lo_inf && dor < 0.0) bv = dor;
                if (isfinite(span) && dor > 0.0) bv = dor;
                acc[12] += d.c64[j] * dj;
                acc[14] += bv * bv;
                acc[15] += ro * ro;
            }
            if (xlast != nullptr) {
                const double dx = xj - (double)xlast[j];
                acc[4] += dx * dx * (asum > 1e-4 ? asum : 1e-4) / 0.999;
            }
        }
    }
    block_sum<NS>(acc, lds);
    if (threadIdx.x == 0) {
        double* p = part_ptr(d, src, 1, blockIdx.x);
        for (int k = 0; k < NP; ++k) p[k] = k < NS ? acc[k] : 0.0;
    }
}

struct KktVal {
    double rp, rd, gap, err, po, dobj, dx, dy;
    bool infeasible;     // the candidate's y is a Farkas certificate of primal infeasibility
    bool unbounded;      // the ray estimate d passes the dual infeasibility test
};

__device__ KktVal kkt_value(const double (&s)[NP], const PdlpCtrl* ctrl) {
    KktVal k;
    k.rp = sqrt(s[0]) / ctrl->norm_b;
    k.rd = sqrt(s[1]) / ctrl->norm_c;
    k.po = s[2];
    k.dobj = s[3];
    k.gap = fabs(k.po - k.dobj) / (1.0 + fabs(k.po) + fabs(k.dobj));
    k.err = fmax(k.rp, fmax(k.rd, k.gap));
    if (!isfinite(k.po) || !isfinite(k.dobj)) k.err = NAN;
    k.dx = sqrt(s[4]);
    k.dy = sqrt(s[5]);
    // F(y) = b_shift'y - max_box (K'y)'(x - lo_shift) > 0 proves that no x in the box satisfies
    // the rows (y <= 0 on <= rows holds by projection, for the average too).  The margin 1e-6
    // of the summed magnitudes exceeds the FP32 rounding of K's values (6e-8) and the FP64
    // rounding of the sums, so a feasible LP is never flagged.
    const double F = s[6] - s[7], scale = s[8] + s[9];
    k.infeasible = isfinite(F) && F > 0.0 && F > 1e-6 * scale;
    // d is a ray of decrease (dual infeasibility) when c'd < 0 and its violations are
    // negligible against any dual solution: for dual-feasible (y, r = c - K'y),
    // c'd >= -|y| |viol(K d)| - |r| |viol_bounds(d)|, so a bounded LP fails
    // |viol(Kd)| (1 + |y|) + |viol_bounds(d)| (1 + |r|) <= M^-1 (-c'd) while the
    // candidate's (y, r) is within a factor M of a dual solution in norm.  M = 1e3 when
    // the rows' violation is also small against their terms, |viol(K d)| <= 1e-3 |(|K||d|)|
    // (K d cancels, as along a ray and unlike the early iterations' moves), else 1e6: a
    // ray on columns that no row holds has no terms to cancel.
    // Second, and independent of the iterate's duals: the violations of d itself must be
    // below kRayEps of the decrease, (|viol(K d)| + |viol_bounds(d)|) <= kRayEps (-c'd).  By
    // the same bound a bounded LP passes this only when every dual solution has
    // |y*| + |r*| >= 1 / kRayEps, so LPs whose optimum is merely far away (small row
    // coefficients, large but finite duals) keep running while their iterates travel.
    const double cd = s[12], vr = sqrt(s[10]), vb = sqrt(s[14]);
    const double lhs = vr * (1.0 + sqrt(s[13])) + vb * (1.0 + sqrt(s[15]));
    k.unbounded = isfinite(cd) && cd < 0.0 && lhs <= (vr <= 1e-3 * sqrt(s[11]) ? 1e-3 : 1e-6) * -cd &&
                  vr + vb <= kRayEps * -cd;
    return k;
}

// mode 0: initial point (candidate 0 only; sets the reference KKT).
// mode 1: check; `cands` = bit mask of the evaluated candidates (bit c = src c).
template <bool RAY>
__global__ void __launch_bounds__(256) kkt_finalize_kernel(PdlpDev d, int nb_rows, int nb_cols, int mode,
                                                           int scheme, int cands) {
    constexpr int NS = RAY ? NP : NP_BASE;      // sums beyond NS are zero
    __shared__ double lds[16 * NS];
    PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) {
        if (threadIdx.x == 0) ctrl->restart_pending = 0;
        return;
    }
    double sums[2][NP];
    for (int c = 0; c < 2; ++c) {
        double acc[NS] = {};
        if (cands & (1 << c)) {
            for (int b = threadIdx.x; b < nb_rows; b += blockDim.x) {
                const double* p = part_ptr(d, c, 0, b);
                for (int k = 0; k < NS; ++k) acc[k] += p[k];
            }
            for (int b = threadIdx.x; b < nb_cols; b += blockDim.x) {
                const double* p = part_ptr(d, c, 1, b);
                for (int k = 0; k < NS; ++k) acc[k] += p[k];
            }
        }
        block_sum<NS>(acc, lds);
        for (int k = 0; k < NP; ++k) sums[c][k] = k < NS ? acc[k] : 0.0;
    }
    if (threadIdx.x != 0) return;
    const KktVal k0 = kkt_value(sums[0], ctrl), k1 = kkt_value(sums[1], ctrl);
    int c;
    if (cands == 1) c = 0;
    else if (cands == 2) c = 1;
    else c = (isfinite(k1.err) && !(k0.err < k1.err)) ? 1 : 0;   // the better candidate; NaN loses
    const KktVal& k = c ? k1 : k0;
    ctrl->rp = k.rp;
    ctrl->rd = k.rd;
    ctrl->gap = k.gap;
    ctrl->p_obj = k.po;
    ctrl->d_obj = k.dobj;
    ctrl->last_err = k.err;
    ctrl->restart_pending = 0;
    // a candidate whose y is a Farkas certificate ends the solve (not with the reference
    // scheme: the reference reports such LPs at the iteration limit)
    const bool inf0 = (cands & 1) && k0.infeasible, inf1 = (cands & 2) && k1.infeasible;
    if (scheme != PDLP_SCHEME_REFERENCE && (inf0 || inf1)) {
        ctrl->status = PDLP_INFEASIBLE;
        ctrl->done_iter = ctrl->iter;
        if (mode != 0) ctrl->checks += 1;
        ctrl->cand = inf0 ? 0 : 1;
        ctrl->restart_pending = inf0 ? 0 : 1;   // finish at the average when it holds the certificate
        ctrl->apply_window = ctrl->window;
        return;
    }
    // the same for a ray of decrease: finish at the current iterate with the window
    // sums kept (apply folds them), so the host forms d = x - x_sum / apply_window
    if (scheme != PDLP_SCHEME_REFERENCE && (cands & 2) && k1.unbounded) {
        ctrl->status = PDLP_UNBOUNDED;
        ctrl->done_iter = ctrl->iter;
        if (mode != 0) ctrl->checks += 1;
        ctrl->cand = 0;
        ctrl->restart_pending = 0;
        ctrl->apply_window = ctrl->window;
        return;
    }
    if (!isfinite(k.err)) {
        ctrl->status = PDLP_NUMERICAL;
        ctrl->done_iter = ctrl->iter;
        return;
    }
    if (mode == 0) {
        ctrl->ref_err = k.err;
        ctrl->err_prev = INFINITY;
        ctrl->last_restart = ctrl->iter;
        return;
    }
    ctrl->checks += 1;
    if (k.err < ctrl->tol) {                     // pdhg.py:231-235
        ctrl->status = PDLP_OPTIMAL;
        ctrl->done_iter = ctrl->iter;
        ctrl->restart_pending = 1;
        ctrl->cand = c;
        ctrl->apply_window = ctrl->window;
        return;
    }
    bool restart;
    if (scheme == PDLP_SCHEME_REFERENCE) {
        restart = k.err < 0.5 * ctrl->ref_err;   // pdhg.py:238-244
    } else {
        const int len = ctrl->iter - ctrl->last_restart;
        restart = k.err <= 0.2 * ctrl->ref_err ||
                  (k.err <= 0.8 * ctrl->ref_err && k.err > ctrl->err_prev) ||
                  (double)len >= 0.36 * (double)ctrl->iter;
    }
    if (!restart) {
        ctrl->err_prev = k.err;
        return;
    }
    if ((scheme == PDLP_SCHEME_ADAPTIVE || scheme == PDLP_SCHEME_ADAPTIVE_START) && k.dx > 1e-10 &&
        k.dy > 1e-10) {
        // primal weight w (tau = w tau_hat, sigma = sigma_hat / w): geometric mean of the
        // old weight and the distance ratio (from the start point, or since the last restart)
        const double lw = 0.5 * log(k.dx / k.dy) + 0.5 * log(ctrl->tau_scale);
        const double wt = clamp_weight(exp(lw), ctrl->weight0);
        ctrl->tau_scale = wt;
        ctrl->sigma_scale = 1.0 / wt;
    }
    // 2: the restart point becomes the reference of the next distances
    ctrl->restart_pending = scheme == PDLP_SCHEME_ADAPTIVE_START ? 1 : 2;
    ctrl->cand = c;
    ctrl->apply_window = ctrl->window;
    ctrl->window = 0;
    ctrl->ref_err = k.err;
    ctrl->err_prev = INFINITY;
    ctrl->last_restart = ctrl->iter;
    ctrl->restarts += 1;
}

// After each check: restart to (or, on convergence, finish at) the chosen
// candidate — the window average or the current iterate — and fold the FP32
// partial sums of the last check_every iterations into the FP64 window sums.
// A restart also records the new reference point (x_last, y_last) of the
// weight update, except with PDLP_SCHEME_ADAPTIVE_START (reference = start).  At
// convergence the FP64 sums are kept (folded) so the host can read the exact
// average.
template <typename T>
__global__ void __launch_bounds__(256) apply_kernel(PdlpDev d) {
    const PdlpCtrl* ctrl = d.ctrl;
    const bool pending = ctrl->restart_pending != 0;
    if (!pending && d.x_psum == nullptr) return;
    const bool keep_sums = ctrl->status == PDLP_OPTIMAL;
    const bool to_avg = ctrl->cand == 1;
    const double w = (double)ctrl->apply_window;
    T* x = (T*)d.x;
    T* y = (T*)d.y;
    T* xl = (T*)d.x_last;
    T* yl = (T*)d.y_last;
    const bool move_ref = ctrl->restart_pending == 2 && xl != nullptr && yl != nullptr;
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < d.n; j += stride) {
        double s = d.x_sum[j];
        if (d.x_psum != nullptr) {
            s += (double)d.x_psum[j];
            d.x_psum[j] = 0.f;
        }
        if (pending) {
            if (to_avg) x[j] = (T)(s / w);
            if (!keep_sums) s = 0.0;
            if (move_ref) xl[j] = x[j];
        }
        d.x_sum[j] = s;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.m; i += stride) {
        double s = d.y_sum[i];
        if (d.y_psum != nullptr) {
            s += (double)d.y_psum[i];
            d.y_psum[i] = 0.f;
        }
        if (pending) {
            if (to_avg) y[i] = (T)(s / w);
            if (!keep_sums) s = 0.0;
            if (move_ref) yl[i] = y[i];
        }
        d.y_sum[i] = s;
    }
}

template <typename T>
__global__ void init_kernel(PdlpDev d) {
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < d.n; j += stride) {
        d.x_sum[j] = 0.0;
        if (d.x_psum != nullptr) d.x_psum[j] = 0.f;
        if (d.x_last != nullptr) ((T*)d.x_last)[j] = ((const T*)d.x)[j];
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.m; i += stride) {
        d.y_sum[i] = 0.0;
        if (d.y_psum != nullptr) d.y_psum[i] = 0.f;
        if (d.y_last != nullptr) ((T*)d.y_last)[i] = ((const T*)d.y)[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        PdlpCtrl* c = d.ctrl;
        c->iter = 0;
        c->window = 0;
        c->status = PDLP_RUNNING;
        c->restarts = 0;
        c->done_iter = 0;
        c->restart_pending = 0;
        c->apply_window = 0;
        c->checks = 0;
        c->cand = 1;
        c->last_restart = 0;
        c->err_prev = INFINITY;
        c->weight0 = c->tau_scale;
    }
}

// ------------------------------------------------------------ launch helpers
template <int V>
using ic = std::integral_constant<int, V>;

// Calls f(ic<TPR>{}) for a supported lanes-per-row value.
template <typename F>
int with_tpr(int tpr, F&& f) {
    switch (tpr) {
        case 1: return f(ic<1>{});
        case 2: return f(ic<2>{});
        case 4: return f(ic<4>{});
        case 8: return f(ic<8>{});
        case 16: return f(ic<16>{});
        case 32: return f(ic<32>{});
        case 64: return f(ic<64>{});
        case 128: return f(ic<128>{});
        case 256: return f(ic<256>{});
        case 512: return f(ic<512>{});
        case 1024: return f(ic<1024>{});
        default: return fail(PDLP_E_TUNE, "tpr must be a power of two in [1, 1024]");
    }
}

bool valid_tpr(int t) { return t >= 1 && t <= 1024 && (t & (t - 1)) == 0; }

int check_dev(const PdlpDev* d) {
    if (d == nullptr) return fail(PDLP_E_ARG, "PdlpDev is NULL");
    if (d->prec != PDLP_F32 && d->prec != PDLP_F64) return fail(PDLP_E_ARG, "prec must be PDLP_F32 or PDLP_F64");
    if (d->m <= 0 || d->n <= 0 || d->nnz < 0) return fail(PDLP_E_SHAPE, "m and n must be positive");
    if (!d->row_ptr || !d->col_idx || !d->val || !d->t_row_ptr || !d->t_col_idx || !d->t_val || !d->c ||
        !d->b || !d->c64 || !d->b64 || !d->b_shift || !d->x || !d->xbar || !d->y || !d->x_sum ||
        !d->y_sum || !d->partials || !d->ctrl)
        return fail(PDLP_E_ARG, "a required PdlpDev pointer is NULL");
    if ((d->x_psum == nullptr) != (d->y_psum == nullptr))
        return fail(PDLP_E_ARG, "x_psum and y_psum must both be set or both be NULL");
    if ((d->row_scale == nullptr) != (d->col_scale == nullptr))
        return fail(PDLP_E_ARG, "row_scale and col_scale must both be set or both be NULL");
    if (d->x_psum != nullptr && d->prec != PDLP_F32)
        return fail(PDLP_E_ARG, "partial sums are for PDLP_F32 only");
    if (d->sell_slices < 0 || d->n_csr_rows < 0 || d->n_csr_rows > d->m)
        return fail(PDLP_E_SHAPE, "sell_slices >= 0 and n_csr_rows in [0, m]");
    if (d->sell_slices > 0 && (!d->sell_ptr || !d->sell_rows || !d->sell_col || !d->sell_val ||
                               (d->n_csr_rows > 0 && !d->csr_rows) || !d->sell_part || !d->sell_asum ||
                               d->sell_max_width <= 0))
        return fail(PDLP_E_ARG, "SELL slices need sell_ptr/rows/col/val/part/asum, csr_rows and sell_max_width");
    if (d->t_sell_slices != 0 && (d->t_sell_slices != (d->n + 63) / 64 || !d->t_sell_ptr || !d->t_sell_idx ||
                                  !d->t_sell_val))
        return fail(PDLP_E_ARG, "the SELL copy of K^T needs ceil(n / 64) slices and t_sell_ptr/idx/val");
    return 0;
}

int check_tune(const PdlpTune* t) {
    if (t == nullptr) return fail(PDLP_E_ARG, "PdlpTune is NULL");
    if (!valid_tpr(t->tpr_primal) || !valid_tpr(t->tpr_dual))
        return fail(PDLP_E_TUNE, "tpr must be a power of two in [1, 1024]");
    if (t->block < 64 || t->block > 1024 || (t->block & 63))
        return fail(PDLP_E_TUNE, "block must be a multiple of 64 in [64, 1024]");
    if (t->block_dual != 0 && (t->block_dual < 64 || t->block_dual > 1024 || (t->block_dual & 63)))
        return fail(PDLP_E_TUNE, "block_dual must be 0 or a multiple of 64 in [64, 1024]");
    if (t->check_every <= 0) return fail(PDLP_E_TUNE, "check_every must be positive");
    if (t->exact && (t->tpr_primal != 1 || t->tpr_dual != 1))
        return fail(PDLP_E_TUNE, "exact mode needs tpr_primal = tpr_dual = 1");
    if (t->scheme < PDLP_SCHEME_REFERENCE || t->scheme > PDLP_SCHEME_ADAPTIVE_START)
        return fail(PDLP_E_TUNE, "scheme must be one of the PDLP_SCHEME_* values");
    if (t->sell < 0 || (t->sell > 7 && (t->sell & 7)))
        return fail(PDLP_E_TUNE, "sell must be 0..7 or a multiple of 8");
    if (t->exact && t->scheme != PDLP_SCHEME_REFERENCE)
        return fail(PDLP_E_TUNE, "exact mode reproduces the reference scheme only");
    if (t->sell_primal < 0 || t->sell_primal > 2) return fail(PDLP_E_TUNE, "sell_primal must be 0, 1 or 2");
    if (t->sell_primal && t->exact) return fail(PDLP_E_TUNE, "exact mode keeps the CSR primal step");
    return 0;
}

int check_all(const PdlpDev* d, const PdlpTune* t) {
    int rc = check_dev(d);
    if (rc || (rc = check_tune(t))) return rc;
    if (t->scheme != PDLP_SCHEME_REFERENCE && (d->x_last == nullptr || d->y_last == nullptr))
        return fail(PDLP_E_ARG, "the adaptive scheme needs x_last and y_last");
    if (t->sell_primal && d->t_sell_slices == 0)
        return fail(PDLP_E_TUNE, "sell_primal needs the SELL copy of K^T (PdlpDev.t_sell_*)");
    return 0;
}

// Workgroups and threads per workgroup for `rows` rows at `tpr` lanes each.
inline void shape_for(int rows, int tpr, int block, int remap, unsigned& grid, unsigned& threads) {
    long g;
    if (tpr > 64) {
        threads = (unsigned)tpr;
        g = rows;
    } else {
        threads = (unsigned)block;
        const long groups = block / tpr;
        g = (rows + groups - 1) / groups;
    }
    if (g < 1) g = 1;
    if (remap) g = (g + 7) / 8 * 8;
    grid = (unsigned)g;
}

template <typename T, bool EXACT>
int launch_primal_t(const PdlpDev& d, const PdlpTune& t, hipStream_t s) {
    const int remap = t.xcd_remap & 1;
    unsigned g, nt;
    if constexpr (!EXACT) {
        if (t.sell_primal == 2) {                     // two slices per wavefront
            shape_for((d.t_sell_slices + 1) / 2 * 64, 1, t.block, remap, g, nt);   // a wavefront per slice pair
            primal_sell2_kernel<T><<<g, nt, 0, s>>>(d, remap);
            return hip_fail(hipGetLastError(), "primal_sell2_kernel launch");
        }
        if (t.sell_primal) {                          // one lane per column over the SELL-64 slices
            shape_for(d.n, 1, t.block, remap, g, nt);
            primal_sell_kernel<T><<<g, nt, 0, s>>>(d, remap);
            return hip_fail(hipGetLastError(), "primal_sell_kernel launch");
        }
    }
    shape_for(d.n, t.tpr_primal, t.block, remap, g, nt);
    if constexpr (EXACT) {
        primal_kernel<T, 1, true><<<g, nt, 0, s>>>(d, remap);
    } else {
        int rc = with_tpr(t.tpr_primal, [&](auto k) {
            primal_kernel<T, decltype(k)::value, false><<<g, nt, 0, s>>>(d, remap);
            return 0;
        });
        if (rc) return rc;
    }
    return hip_fail(hipGetLastError(), "primal_kernel launch");
}

// Split SELL-64 dual step, pass 1: wavefront (slice, chunk c) sums entries
// [c kc, (c+1) kc) of each lane's row into sell_part.
template <typename T>
__global__ void __launch_bounds__(256) dual_sell_part_kernel(PdlpDev d, int kc, int chunks) {
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const int s = blockIdx.y;
    const int c = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    const int beg = d.sell_ptr[s], width = (d.sell_ptr[s + 1] - beg) >> 6;
    const int k0 = c * kc;
    if (c >= chunks || k0 >= width) return;
    const int k1 = min(width, k0 + kc);
    const int lane = threadIdx.x & 63;
    const T* __restrict__ xbar = (const T*)d.xbar;
    const T* __restrict__ val = (const T*)d.sell_val;
    const int32_t* __restrict__ col = d.sell_col;
    T ax = T(0);
    int q = beg + k0 * 64 + lane;
#pragma unroll 8
    for (int k = k0; k < k1; ++k, q += 64) ax += val[q] * xbar[col[q]];
    ((T*)d.sell_part)[((size_t)s * chunks + c) * 64 + lane] = ax;
}

// Pass 2: one wavefront per slice adds the chunks in order and updates y.
template <typename T>
__global__ void __launch_bounds__(256) dual_sell_finish_kernel(PdlpDev d, int kc, int chunks) {
    const PdlpCtrl* ctrl = d.ctrl;
    if (ctrl->status != PDLP_RUNNING) return;
    const int s = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    if (s >= d.sell_slices) return;
    const int lane = threadIdx.x & 63;
    const int width = (d.sell_ptr[s + 1] - d.sell_ptr[s]) >> 6;
    const int used = (width + kc - 1) / kc;
    const T* part = (const T*)d.sell_part + (size_t)s * chunks * 64 + lane;
    T ax = T(0);
    for (int c = 0; c < used; ++c) ax += part[(size_t)c * 64];
    const int i = d.sell_rows[s * 64 + lane];
    if (i < 0) return;
    dual_update<T, false>(d, (T)ctrl->sigma_scale, i, ax, ((const T*)d.sell_asum)[s * 64 + lane]);
}

// The one-launch SELL kernel in the layout of t.sell (2..7): NT threads, R rows
// per slice piece, tpr_dual threads per CSR row (64 .. NT).
template <typename T, bool SPMV>
void launch_sell_wg(const PdlpDev& d, const PdlpTune& t, const T* vx, T* vy, hipStream_t s) {
    static constexpr int kRows[8] = {0, 0, 64, 64, 16, 8, 32, 16};
    static constexpr int kThreads[8] = {0, 0, 1024, 512, 1024, 1024, 1024, 512};
    const int nt = kThreads[t.sell];
    const int tc = std::min(std::max(t.tpr_dual, 64), nt);
    const int per = nt / tc;
    const unsigned items = (unsigned)(d.sell_slices * (64 / kRows[t.sell]) + (d.n_csr_rows + per - 1) / per);
    auto go = [&](auto nt_c, auto r_c) {
        constexpr int NT = decltype(nt_c)::value, R = decltype(r_c)::value;
        switch (tc) {
            case 64: dual_sell_wg_kernel<T, NT, R, 64, SPMV><<<items, NT, 0, s>>>(d, vx, vy); break;
            case 128: dual_sell_wg_kernel<T, NT, R, 128, SPMV><<<items, NT, 0, s>>>(d, vx, vy); break;
            case 256: dual_sell_wg_kernel<T, NT, R, 256, SPMV><<<items, NT, 0, s>>>(d, vx, vy); break;
            case 512: dual_sell_wg_kernel<T, NT, R, 512, SPMV><<<items, NT, 0, s>>>(d, vx, vy); break;
            default: dual_sell_wg_kernel<T, NT, R, NT, SPMV><<<items, NT, 0, s>>>(d, vx, vy); break;
        }
    };
    switch (t.sell) {
        case 2: go(ic<1024>{}, ic<64>{}); break;
        case 3: go(ic<512>{}, ic<64>{}); break;
        case 4: go(ic<1024>{}, ic<16>{}); break;
        case 5: go(ic<1024>{}, ic<8>{}); break;
        case 6: go(ic<1024>{}, ic<32>{}); break;
        default: go(ic<512>{}, ic<16>{}); break;
    }
}

template <typename T, bool EXACT>
int launch_dual_t(const PdlpDev& d, const PdlpTune& t, hipStream_t s) {
    const int remap = (t.xcd_remap >> 1) & 1;
    const bool sell = !EXACT && t.sell && d.sell_slices > 0;
    unsigned g, nt;
    // with SELL slices the CSR kernel covers csr_rows only (at least one workgroup: it counts the iteration)
    shape_for(sell ? std::max(d.n_csr_rows, 1) : d.m, t.tpr_dual, (t.block_dual > 0 ? t.block_dual : t.block),
              remap, g, nt);
    if constexpr (EXACT) {
        dual_kernel<T, 1, true><<<g, nt, 0, s>>>(d, remap, 0);
    } else {
        if (sell && t.sell >= 2 && t.sell <= 7) {
            // slices and CSR rows in one launch (block 0 counts the iteration)
            launch_sell_wg<T, false>(d, t, nullptr, nullptr, s);
            return hip_fail(hipGetLastError(), "dual_sell_wg_kernel launch");
        }
        if (sell && t.sell == 1) {
            dual_sell_kernel<T><<<(unsigned)((d.sell_slices + 3) / 4), 256, 0, s>>>(d);
        } else if (sell) {
            const int kc = t.sell, chunks = (d.sell_max_width + kc - 1) / kc;
            dual_sell_part_kernel<T><<<dim3((unsigned)((chunks + 3) / 4), (unsigned)d.sell_slices), 256, 0, s>>>(
                d, kc, chunks);
            dual_sell_finish_kernel<T><<<(unsigned)((d.sell_slices + 3) / 4), 256, 0, s>>>(d, kc, chunks);
        }
        int rc = with_tpr(t.tpr_dual, [&](auto k) {
            dual_kernel<T, decltype(k)::value, false><<<g, nt, 0, s>>>(d, remap, sell ? 1 : 0);
            return 0;
        });
        if (rc) return rc;
    }
    return hip_fail(hipGetLastError(), "dual_kernel launch");
}

template <typename T>
int launch_kkt_t(const PdlpDev& d, const PdlpTune& t, int mode, hipStream_t s) {
    constexpr int B = 256;
    auto nb = [](int rows, int tpr) {
        long g = (rows + (B / tpr) - 1) / (B / tpr);
        return (int)std::max(1L, std::min<long>(g, PDLP_PARTIAL_BLOCKS));
    };
    // the check's reductions use at most a wavefront per row
    const int td = std::min(t.tpr_dual, 64), tp = std::min(t.tpr_primal, 64);
    const int nr = nb(d.m, td), nc = nb(d.n, tp);
    // candidates: the initial point is "current"; a reference check uses the
    // average only; an adaptive check evaluates both
    const int cands = mode == 0 ? 1 : (t.scheme != PDLP_SCHEME_REFERENCE ? 3 : 2);
    const bool check_ray = mode == 1 && t.scheme != PDLP_SCHEME_REFERENCE && !t.no_ray;
    for (int src = 0; src < 2; ++src) {
        if (!(cands & (1 << src))) continue;
        // the unboundedness test rides on the average's launches of an adaptive check
        const bool ray = src == 1 && check_ray;
        int rc = with_tpr(td, [&](auto k) {
            constexpr int P = decltype(k)::value;
            if constexpr (P <= 64) {
                if (ray) kkt_rows_kernel<T, P, true><<<nr, B, 0, s>>>(d, src);
                else kkt_rows_kernel<T, P, false><<<nr, B, 0, s>>>(d, src);
            }
            return 0;
        });
        if (rc) return rc;
        rc = with_tpr(tp, [&](auto k) {
            constexpr int P = decltype(k)::value;
            if constexpr (P <= 64) {
                if (ray) kkt_cols_kernel<T, P, true><<<nc, B, 0, s>>>(d, src);
                else kkt_cols_kernel<T, P, false><<<nc, B, 0, s>>>(d, src);
            }
            return 0;
        });
        if (rc) return rc;
    }
    if (check_ray) kkt_finalize_kernel<true><<<1, B, 0, s>>>(d, nr, nc, mode, t.scheme, cands);
    else kkt_finalize_kernel<false><<<1, B, 0, s>>>(d, nr, nc, mode, t.scheme, cands);
    if (mode == 1) {
        const int work = std::max(d.n, d.m);
        const int g = std::min((work + B - 1) / B, 4096);
        apply_kernel<T><<<g, B, 0, s>>>(d);
    }
    return hip_fail(hipGetLastError(), "kkt kernels launch");
}

int primal(const PdlpDev& d, const PdlpTune& t, hipStream_t s) {
    if (d.prec == PDLP_F64)
        return t.exact ? launch_primal_t<double, true>(d, t, s) : launch_primal_t<double, false>(d, t, s);
    return t.exact ? launch_primal_t<float, true>(d, t, s) : launch_primal_t<float, false>(d, t, s);
}

int dual(const PdlpDev& d, const PdlpTune& t, hipStream_t s) {
    if (d.prec == PDLP_F64)
        return t.exact ? launch_dual_t<double, true>(d, t, s) : launch_dual_t<double, false>(d, t, s);
    return t.exact ? launch_dual_t<float, true>(d, t, s) : launch_dual_t<float, false>(d, t, s);
}

int kkt(const PdlpDev& d, const PdlpTune& t, int mode, hipStream_t s) {
    return d.prec == PDLP_F64 ? launch_kkt_t<double>(d, t, mode, s) : launch_kkt_t<float>(d, t, mode, s);
}

int run(const PdlpDev& d, const PdlpTune& t, int iters, int check, hipStream_t s) {
    for (int k = 0; k < iters; ++k) {
        int rc = primal(d, t, s);
        if (rc) return rc;
        rc = dual(d, t, s);
        if (rc) return rc;
    }
    return check ? kkt(d, t, 1, s) : 0;
}

}  // namespace

// Library-internal entry points for the other sources of libpdlp_hip.so.
int single_check_all(const PdlpDev* d, const PdlpTune* t) { return check_all(d, t); }
int single_run(const PdlpDev& d, const PdlpTune& t, int iters, int check, hipStream_t s) {
    return run(d, t, iters, check, s);
}

}  // namespace pdlp

using namespace pdlp;

extern "C" {

int pdlp_init(const PdlpDev* dev, const PdlpTune* tune, void* stream) {
    int rc = check_all(dev, tune);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const int work = std::max(dev->n, dev->m);
    if (dev->prec == PDLP_F64) init_kernel<double><<<std::min((work + 255) / 256, 4096), 256, 0, s>>>(*dev);
    else init_kernel<float><<<std::min((work + 255) / 256, 4096), 256, 0, s>>>(*dev);
    if ((rc = hip_fail(hipGetLastError(), "init_kernel launch"))) return rc;
    return kkt(*dev, *tune, 0, s);
}

int pdlp_primal_step(const PdlpDev* dev, const PdlpTune* tune, void* stream) {
    int rc = check_all(dev, tune);
    if (rc) return rc;
    return primal(*dev, *tune, as_stream(stream));
}

int pdlp_dual_step(const PdlpDev* dev, const PdlpTune* tune, void* stream) {
    int rc = check_all(dev, tune);
    if (rc) return rc;
    return dual(*dev, *tune, as_stream(stream));
}

int pdlp_kkt_check(const PdlpDev* dev, const PdlpTune* tune, void* stream) {
    int rc = check_all(dev, tune);
    if (rc) return rc;
    return kkt(*dev, *tune, 1, as_stream(stream));
}

int pdlp_run(const PdlpDev* dev, const PdlpTune* tune, int32_t iters, int32_t check, void* stream) {
    int rc = check_all(dev, tune);
    if (rc) return rc;
    if (iters < 0) return fail(PDLP_E_ARG, "iters must be >= 0");
    return run(*dev, *tune, iters, check, as_stream(stream));
}

int pdlp_graph_capture(const PdlpDev* dev, const PdlpTune* tune, int32_t blocks, void* stream, void** exec) {
    int rc = check_all(dev, tune);
    if (rc) return rc;
    if (blocks <= 0 || exec == nullptr) return fail(PDLP_E_ARG, "blocks must be positive and exec non-NULL");
    if (stream == nullptr) return fail(PDLP_E_ARG, "graph capture needs a non-default stream");
    hipStream_t s = as_stream(stream);
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    for (int k = 0; k < blocks && rc == 0; ++k) rc = run(*dev, *tune, tune->check_every, 1, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *exec = (void*)ge;
    return 0;
}

int pdlp_graph_capture_step(const PdlpDev* dev, const PdlpTune* tune, int32_t which, int32_t reps, void* stream,
                            void** exec) {
    int rc = check_all(dev, tune);
    if (rc) return rc;
    if (reps <= 0 || exec == nullptr || (which != 0 && which != 1))
        return fail(PDLP_E_ARG, "reps > 0, which in {0, 1} and exec non-NULL");
    if (stream == nullptr) return fail(PDLP_E_ARG, "graph capture needs a non-default stream");
    hipStream_t s = as_stream(stream);
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    for (int k = 0; k < reps && rc == 0; ++k) rc = which == 0 ? primal(*dev, *tune, s) : dual(*dev, *tune, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *exec = (void*)ge;
    return 0;
}

int pdlp_spmv_dev(const PdlpDev* dev, const PdlpTune* tune, const void* x, void* y, void* stream) {
    if (int rc = check_dev(dev)) return rc;
    if (int rc = check_tune(tune)) return rc;
    if (tune->sell < 2 || tune->sell > 7) return fail(PDLP_E_TUNE, "spmv_dev: tune->sell must be 2..7");
    if (dev->sell_slices <= 0) return fail(PDLP_E_ARG, "spmv_dev: dev has no SELL-64 rows (use pdlp_spmv_csr)");
    if (x == nullptr || y == nullptr) return fail(PDLP_E_ARG, "spmv_dev: x or y is NULL");
    hipStream_t s = as_stream(stream);
    if (dev->prec == PDLP_F32) launch_sell_wg<float, true>(*dev, *tune, (const float*)x, (float*)y, s);
    else launch_sell_wg<double, true>(*dev, *tune, (const double*)x, (double*)y, s);
    return hip_fail(hipGetLastError(), "spmv_dev launch");
}

int pdlp_graph_launch(void* exec, void* stream) {
    if (exec == nullptr) return fail(PDLP_E_ARG, "graph exec is NULL");
    return hip_fail(hipGraphLaunch((hipGraphExec_t)exec, as_stream(stream)), "hipGraphLaunch");
}

int pdlp_graph_destroy(void* exec) {
    if (exec == nullptr) return 0;
    return hip_fail(hipGraphExecDestroy((hipGraphExec_t)exec), "hipGraphExecDestroy");
}

}  // extern "C"
