// Batched restarted PDHG: one persistent launch, one workgroup per LP at a time.
//
// Per LP (box form, rows '=' or '<=', lo <= x <= hi), LDS holds
//   R   : 2n floats — xbar[0..n) and the K^T y accumulator aty[n..2n);
//         reused as n doubles during a KKT check
//   yl  : m floats — the dual iterate
//   ypl : m floats — its sum since the last check
// and each iteration is
//   column pass  x+ = clip(x - tau (c - aty)), xbar = 2x+ - x, aty = 0
//   row pass     d = K_i xbar; y_i+ = proj(y_i + sigma_i (b_i - d));
//                aty[j] += K_ij y_i+ for the row's entries (LDS atomics)
// so the LP's CSR is read once per iteration.  Every check_every iterations
// the partial sums are folded into the FP64 window sums and the reference KKT
// of the window average is evaluated in FP64 (pdhg.py:184-196 on the lifted
// form), followed by the termination/restart rule of pdhg.py:225-244.
// The float atomics make the summation order of K^T y run-dependent; results
// agree to rounding, not bit for bit, between runs.

#include "pdlp_common.hpp"

#include <algorithm>
#include <type_traits>

namespace pdlp {
namespace {

struct LPView {
    int m, n;
    const int32_t* rp;     // m+1, relative to nnz base
    const int32_t* ci;     // relative to column base
    const float* val;
    const float *c, *lo, *hi;
    const double *c64, *lo_shift, *ub_span;
    const float* b;
    const int8_t* kind;
    const double *b64, *b_shift;
    float *x, *x_psum, *tau;
    double* x_sum;
    float* y;
    double* y_sum;
    float *x_last, *y_last;   // last restart point (adaptive scheme)
    double norm_b, norm_c;
    float weight;             // starting primal weight w: tau = w tau_hat, sigma = sigma_hat / w
    bool adaptive;            // adaptive restarts (flags bit 0)
    bool fixed_weight;        // keep the primal weight (flags bit 1)
    bool move_ref;            // distances since the last restart (flags bit 2 clear) or from the start
};

__device__ __forceinline__ LPView view_of(const PdlpBatch& a, const PdlpBatchLP& lp) {
    LPView v;
    v.m = lp.m;
    v.n = lp.n;
    v.rp = a.row_ptr + lp.rp_off;
    v.ci = a.col_idx + lp.nnz_off;
    v.val = a.val + lp.nnz_off;
    v.c = a.c + lp.col_off;
    v.lo = a.lo + lp.col_off;
    v.hi = a.hi + lp.col_off;
    v.c64 = a.c64 + lp.col_off;
    v.lo_shift = a.lo_shift + lp.col_off;
    v.ub_span = a.ub_span + lp.col_off;
    v.b = a.b + lp.row_off;
    v.kind = a.row_kind + lp.row_off;
    v.b64 = a.b64 + lp.row_off;
    v.b_shift = a.b_shift + lp.row_off;
    v.x = a.x + lp.col_off;
    v.x_psum = a.x_psum + lp.col_off;
    v.x_sum = a.x_sum + lp.col_off;
    v.tau = a.tau + lp.col_off;
    v.y = a.y + lp.row_off;
    v.y_sum = a.y_sum + lp.row_off;
    v.x_last = a.x_last + lp.col_off;
    v.y_last = a.y_last + lp.row_off;
    v.norm_b = lp.norm_b;
    v.norm_c = lp.norm_c;
    v.weight = (float)lp.tau_scale;
    v.adaptive = (lp.flags & 1) != 0;
    v.fixed_weight = (lp.flags & 2) != 0;
    v.move_ref = (lp.flags & 4) == 0;
    return v;
}

// Scatter acc[col] += scale(row) * K_ij over all rows (LDS atomics), TPR lanes per row.
template <int NT, int TPR, typename A, typename F>
__device__ __forceinline__ void scatter_rows(const LPView& v, A* acc, F row_scale) {
    const int g = threadIdx.x / TPR, lane = threadIdx.x % TPR;
    for (int i = g; i < v.m; i += NT / TPR) {
        const A s = row_scale(i);
        if (s == A(0)) continue;                      // uniform across the group
        const int beg = v.rp[i], end = v.rp[i + 1];
        for (int k = beg + lane; k < end; k += TPR) atomicAdd(&acc[v.ci[k]], (A)v.val[k] * s);
    }
}

// Block-wide sum of NV doubles, result broadcast to all threads.
template <int NT, int NV>
__device__ __forceinline__ void block_sum_all(double (&acc)[NV], double* red) {
    constexpr int NW = NT / kWave;
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = group_sum<kWave>(acc[k]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[w * NV + k] = acc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = 0.0;
        for (int i = 0; i < NW; ++i) t += red[i * NV + k];
        acc[k] = t;
    }
    __syncthreads();
}

constexpr int NS = 6;   // rp^2, rd^2, primal obj, dual obj, |dx|^2, |dy|^2

// Reference KKT of (x, y): src 0 = current iterate, 1 = window average (w).
// Uses R as n doubles.  acc (all threads): rp^2, rd^2, primal and dual
// objective, and for the adaptive scheme the squared distances of the point
// from the last restart point in the step-size metric (dx^2/tau_hat, dy^2/sigma_hat).
template <int NT, int TPR>
__device__ void kkt_eval(const LPView& v, float* R, const float* yl, int src, double w, double (&acc)[NS],
                         double* red) {
    double* R64 = reinterpret_cast<double*>(R);
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;
    // C1: primal point into LDS (FP64), primal objective, primal distance
    for (int j = threadIdx.x; j < v.n; j += NT) {
        const double xj = src ? v.x_sum[j] / w : (double)v.x[j];
        R64[j] = xj;
        acc[2] += v.c64[j] * (xj - v.lo_shift[j]);
        if (v.adaptive) {
            const double dx = xj - (double)v.x_last[j];
            acc[4] += dx * dx / (double)v.tau[j];
        }
    }
    __syncthreads();
    // C2: rows — primal residual, dual objective, slack dual residual, dual distance
    {
        const int g = threadIdx.x / TPR, lane = threadIdx.x % TPR;
        for (int i = g; i < v.m; i += NT / TPR) {
            const int beg = v.rp[i], end = v.rp[i + 1];
            double d = 0.0, as = 0.0;
            for (int k = beg + lane; k < end; k += TPR) {
                const double kv = (double)v.val[k];
                d += kv * R64[v.ci[k]];
                as += fabs(kv);
            }
            if constexpr (TPR > 1) {
                d = group_sum<TPR>(d);
                as = group_sum<TPR>(as);
            }
            if (lane == 0) {
                const bool le = v.kind[i] == PDLP_ROW_LE;
                double res = d - v.b64[i];
                if (le && res < 0.0) res = 0.0;
                const double yi = src ? v.y_sum[i] / w : (double)yl[i];
                acc[0] += res * res;
                if (le && yi > 0.0) acc[1] += yi * yi;
                acc[3] += v.b_shift[i] * yi;
                if (v.adaptive) {
                    const double dy = yi - (double)v.y_last[i];
                    acc[5] += dy * dy * (as > 1e-4 ? as : 1e-4) / 0.999;
                }
            }
        }
    }
    __syncthreads();
    // C3: K^T y into LDS (FP64 atomics)
    for (int j = threadIdx.x; j < v.n; j += NT) R64[j] = 0.0;
    __syncthreads();
    scatter_rows<NT, TPR, double>(v, R64, [&](int i) { return src ? v.y_sum[i] / w : (double)yl[i]; });
    __syncthreads();
    // C4: columns — dual residual / upper-bound-row duals
    for (int j = threadIdx.x; j < v.n; j += NT) {
        const double r = v.c64[j] - R64[j];
        const double rneg = r < 0.0 ? r : 0.0;
        const double span = v.ub_span[j];
        if (isfinite(span)) acc[3] += span * rneg;
        else acc[1] += rneg * rneg;
    }
    block_sum_all<NT, NS>(acc, red);
}

// aty = K^T yl (float), R's upper half.
template <int NT, int TPR>
__device__ __forceinline__ void rebuild_aty(const LPView& v, float* R, const float* yl) {
    float* aty = R + v.n;
    for (int j = threadIdx.x; j < v.n; j += NT) aty[j] = 0.f;
    __syncthreads();
    scatter_rows<NT, TPR, float>(v, aty, [&](int i) { return yl[i]; });
    __syncthreads();
}

struct Kkt {
    double rp, rd, gap, err, po, dobj, dx, dy;
};

__device__ __forceinline__ Kkt kkt_of(const double (&s)[NS], const LPView& v) {
    Kkt k;
    k.rp = sqrt(s[0]) / v.norm_b;
    k.rd = sqrt(s[1]) / v.norm_c;
    k.po = s[2];
    k.dobj = s[3];
    k.gap = fabs(k.po - k.dobj) / (1.0 + fabs(k.po) + fabs(k.dobj));
    k.err = fmax(k.rp, fmax(k.rd, k.gap));
    if (!isfinite(k.po) || !isfinite(k.dobj)) k.err = NAN;
    k.dx = sqrt(s[4]);
    k.dy = sqrt(s[5]);
    return k;
}

template <int NT, int TPR>
__device__ void solve_one(const PdlpBatch& a, int k, float* lds, double* red) {
    const PdlpBatchLP lp = a.lps[k];
    const LPView v = view_of(a, lp);
    float* R = lds;
    float* xb = R;
    float* aty = R + v.n;
    float* yl = R + 2 * v.n;
    float* ypl = yl + v.m;

    // ---- set-up: y into LDS, sums cleared, restart point = start, tau_hat from column sums of |K|
    for (int i = threadIdx.x; i < v.m; i += NT) {
        yl[i] = v.y[i];
        ypl[i] = 0.f;
        v.y_sum[i] = 0.0;
        v.y_last[i] = v.y[i];
    }
    for (int j = threadIdx.x; j < v.n; j += NT) {
        aty[j] = 0.f;
        v.x_sum[j] = 0.0;
        v.x_psum[j] = 0.f;
        v.x_last[j] = v.x[j];
    }
    __syncthreads();
    {
        const int g = threadIdx.x / TPR, lane = threadIdx.x % TPR;
        for (int i = g; i < v.m; i += NT / TPR)
            for (int q = v.rp[i] + lane; q < v.rp[i + 1]; q += TPR) atomicAdd(&aty[v.ci[q]], fabsf(v.val[q]));
    }
    __syncthreads();
    for (int j = threadIdx.x; j < v.n; j += NT) {
        const float s = aty[j];
        v.tau[j] = 0.999f / (s > 1e-4f ? s : 1e-4f);
    }
    __syncthreads();

    // starting primal weight: given, or (NaN) ||sqrt(sigma_hat) b|| / ||sqrt(tau_hat) c||
    float w_start = v.weight;
    if (!(w_start > 0.f)) {
        double nw[2] = {0.0, 0.0};
        for (int j = threadIdx.x; j < v.n; j += NT) nw[1] += v.c64[j] * v.c64[j] * (double)v.tau[j];
        const int g = threadIdx.x / TPR, lane = threadIdx.x % TPR;
        for (int i = g; i < v.m; i += NT / TPR) {
            double s = 0.0;
            for (int q = v.rp[i] + lane; q < v.rp[i + 1]; q += TPR) s += fabs((double)v.val[q]);
            if constexpr (TPR > 1) s = group_sum<TPR>(s);
            if (lane == 0) nw[0] += v.b64[i] * v.b64[i] * 0.999 / (s > 1e-4 ? s : 1e-4);
        }
        block_sum_all<NT, 2>(nw, red);
        const double nb = sqrt(nw[0]), nc = sqrt(nw[1]);
        w_start = (nb > 1e-12 && nc > 1e-12) ? (float)fmin(fmax(nb / nc, 1e-6), 1e6) : 1.f;
    }

    double acc[NS], acc2[NS];
    kkt_eval<NT, TPR>(v, R, yl, 0, 1.0, acc, red);
    Kkt kk = kkt_of(acc, v);
    double ref_err = kk.err, err_prev = INFINITY;
    int status = isfinite(kk.err) ? PDLP_RUNNING : PDLP_NUMERICAL;
    int window = 0, restarts = 0, checks = 0, it = 0, done_it = 0, last_restart = 0;
    bool end_on_avg = false;
    double w_final = 1.0;
    float wgt = w_start;
    rebuild_aty<NT, TPR>(v, R, yl);

    const int ce = a.check_every;
    const int g = threadIdx.x / TPR, lane = threadIdx.x % TPR;
    while (status == PDLP_RUNNING && it < a.max_iters) {
        const float inv_wgt = 1.f / wgt;
        // ---- column pass
        for (int j = threadIdx.x; j < v.n; j += NT) {
            const float xo = v.x[j];
            float xn = xo - (wgt * v.tau[j]) * (v.c[j] - aty[j]);
            xn = xn < v.lo[j] ? v.lo[j] : xn;
            xn = xn > v.hi[j] ? v.hi[j] : xn;
            xb[j] = 2.f * xn - xo;
            aty[j] = 0.f;
            v.x[j] = xn;
            v.x_psum[j] += xn;
        }
        __syncthreads();
        // ---- row pass: dual update and K^T y+ scatter
        for (int i = g; i < v.m; i += NT / TPR) {
            const int beg = v.rp[i], end = v.rp[i + 1];
            float d = 0.f, s = 0.f;
            for (int q = beg + lane; q < end; q += TPR) {
                const float kv = v.val[q];
                d += kv * xb[v.ci[q]];
                s += fabsf(kv);
            }
            if constexpr (TPR > 1) {
                d = group_sum<TPR>(d);
                s = group_sum<TPR>(s);
            }
            const float sigma = inv_wgt * (0.999f / (s > 1e-4f ? s : 1e-4f));
            float yn = yl[i] + sigma * (v.b[i] - d);
            if (v.kind[i] == PDLP_ROW_LE && yn > 0.f) yn = 0.f;
            if (yn != 0.f)
                for (int q = beg + lane; q < end; q += TPR) atomicAdd(&aty[v.ci[q]], v.val[q] * yn);
            if (lane == 0) {
                yl[i] = yn;
                ypl[i] += yn;
            }
        }
        __syncthreads();
        ++it;
        ++window;
        if (window % ce != 0) continue;

        // ---- check: fold partial sums, KKT of the candidate(s), decision
        for (int j = threadIdx.x; j < v.n; j += NT) {
            v.x_sum[j] += (double)v.x_psum[j];
            v.x_psum[j] = 0.f;
        }
        for (int i = threadIdx.x; i < v.m; i += NT) {
            v.y_sum[i] += (double)ypl[i];
            ypl[i] = 0.f;
        }
        __syncthreads();
        const double w = (double)window;
        kkt_eval<NT, TPR>(v, R, yl, 1, w, acc, red);
        Kkt ka = kkt_of(acc, v);
        bool use_avg = true;
        if (v.adaptive) {                       // the better of average and current iterate
            kkt_eval<NT, TPR>(v, R, yl, 0, 1.0, acc2, red);
            const Kkt kc = kkt_of(acc2, v);
            if (!isfinite(ka.err) || kc.err < ka.err) {
                use_avg = false;
                ka = kc;
#pragma unroll
                for (int q = 0; q < NS; ++q) acc[q] = acc2[q];
            }
        }
        kk = ka;
        ++checks;
        // every thread computed the same values from the broadcast sums
        if (!isfinite(kk.err)) {
            status = PDLP_NUMERICAL;
            done_it = it;
            break;
        }
        if (kk.err < a.tol) {                   // pdhg.py:231-235
            status = PDLP_OPTIMAL;
            done_it = it;
            end_on_avg = use_avg;
            w_final = w;
            break;
        }
        bool restart;
        if (!v.adaptive) {
            restart = kk.err < 0.5 * ref_err;   // pdhg.py:238-244
        } else {
            restart = kk.err <= 0.2 * ref_err || (kk.err <= 0.8 * ref_err && kk.err > err_prev) ||
                      (double)(it - last_restart) >= 0.36 * (double)it;
        }
        if (restart) {
            if (v.adaptive && !v.fixed_weight && kk.dx > 1e-10 && kk.dy > 1e-10) {
                const double lw = 0.5 * log(kk.dx / kk.dy) + 0.5 * log((double)wgt);
                wgt = (float)clamp_weight(exp(lw), (double)w_start);
            }
            ref_err = kk.err;
            err_prev = INFINITY;
            last_restart = it;
            ++restarts;
            for (int j = threadIdx.x; j < v.n; j += NT) {
                if (use_avg) v.x[j] = (float)(v.x_sum[j] / w);
                v.x_sum[j] = 0.0;
                if (v.move_ref) v.x_last[j] = v.x[j];
            }
            for (int i = threadIdx.x; i < v.m; i += NT) {
                if (use_avg) yl[i] = (float)(v.y_sum[i] / w);
                v.y_sum[i] = 0.0;
                if (v.move_ref) v.y_last[i] = yl[i];
            }
            window = 0;
            __syncthreads();
        } else {
            err_prev = kk.err;
        }
        rebuild_aty<NT, TPR>(v, R, yl);
    }

    // ---- outputs: FP64 solution in x_sum / y_sum
    if (status == PDLP_OPTIMAL && end_on_avg) {
        for (int j = threadIdx.x; j < v.n; j += NT) v.x_sum[j] = v.x_sum[j] / w_final;
        for (int i = threadIdx.x; i < v.m; i += NT) v.y_sum[i] = v.y_sum[i] / w_final;
    } else {
        for (int j = threadIdx.x; j < v.n; j += NT) v.x_sum[j] = (double)v.x[j];
        for (int i = threadIdx.x; i < v.m; i += NT) v.y_sum[i] = (double)yl[i];
        if (status == PDLP_RUNNING) done_it = it;
    }
    for (int i = threadIdx.x; i < v.m; i += NT) v.y[i] = yl[i];
    if (threadIdx.x == 0) {
        PdlpBatchOut o;
        o.status = status;
        o.iterations = done_it;
        o.restarts = restarts;
        o.checks = checks;
        o.rel_kkt = kk.err;
        o.rp = kk.rp;
        o.rd = kk.rd;
        o.gap = kk.gap;
        o.p_obj = kk.po;
        o.d_obj = kk.dobj;
        o.weight = (double)wgt;
        a.out[k] = o;
    }
    __syncthreads();
}

template <int NT>
__global__ void __launch_bounds__(NT) batch_kernel(PdlpBatch a) {
    extern __shared__ float lds[];
    __shared__ double red[NS * (NT / kWave)];
    __shared__ int s_q;
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(a.queue, 1);
        __syncthreads();
        // uniform values: the LP's descriptor and pointers then live in scalar registers
        const int q = __builtin_amdgcn_readfirstlane(s_q);
        __syncthreads();
        if (q >= a.count) return;                 // every wave of the workgroup leaves here
        const int k = __builtin_amdgcn_readfirstlane(a.order[q]);
        switch (a.lps[k].tpr) {
            case 1: solve_one<NT, 1>(a, k, lds, red); break;
            case 2: solve_one<NT, 2>(a, k, lds, red); break;
            case 4: solve_one<NT, 4>(a, k, lds, red); break;
            case 8: solve_one<NT, 8>(a, k, lds, red); break;
            case 16: solve_one<NT, 16>(a, k, lds, red); break;
            case 32: solve_one<NT, 32>(a, k, lds, red); break;
            default: solve_one<NT, 64>(a, k, lds, red); break;
        }
    }
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" int pdlp_batch_solve(const PdlpBatch* a, void* stream) {
    if (a == nullptr) return fail(PDLP_E_ARG, "PdlpBatch is NULL");
    if (a->count < 0) return fail(PDLP_E_ARG, "count must be >= 0");
    if (a->count == 0) return 0;
    if (!a->lps || !a->order || !a->queue || !a->row_ptr || !a->col_idx || !a->val || !a->c || !a->lo ||
        !a->hi || !a->c64 || !a->lo_shift || !a->ub_span || !a->b || !a->row_kind || !a->b64 || !a->b_shift ||
        !a->x || !a->x_psum || !a->x_sum || !a->tau || !a->y || !a->y_sum || !a->out || !a->x_last || !a->y_last)
        return fail(PDLP_E_ARG, "a required PdlpBatch pointer is NULL");
    if (a->check_every <= 0 || a->max_iters < 0) return fail(PDLP_E_ARG, "check_every > 0 and max_iters >= 0");
    if (a->lds_bytes <= 0 || a->lds_bytes > 160 * 1024)
        return fail(PDLP_E_LIMIT, "lds_bytes must be in (0, 160 KiB]: an LP needs 8*(m+n) bytes");
    hipStream_t s = as_stream(stream);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    const unsigned grid = (unsigned)std::min(a->count, a->grid > 0 ? a->grid : 2 * cus);
    hipError_t e;
    switch (a->block) {
        case 256:
            e = hipFuncSetAttribute((const void*)batch_kernel<256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    a->lds_bytes);
            if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
            batch_kernel<256><<<grid, 256, a->lds_bytes, s>>>(*a);
            break;
        case 512:
            e = hipFuncSetAttribute((const void*)batch_kernel<512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    a->lds_bytes);
            if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
            batch_kernel<512><<<grid, 512, a->lds_bytes, s>>>(*a);
            break;
        case 1024:
            e = hipFuncSetAttribute((const void*)batch_kernel<1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    a->lds_bytes);
            if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
            batch_kernel<1024><<<grid, 1024, a->lds_bytes, s>>>(*a);
            break;
        default:
            return fail(PDLP_E_TUNE, "block must be 256, 512 or 1024");
    }
    return hip_fail(hipGetLastError(), "batch_kernel launch");
}
