// Batched restarted PDHG: persistent launches, one workgroup per LP at a time.
//
// Per LP (box form, rows '=' or '<=', lo <= x <= hi), LDS holds
//   X   : n doubles — xbar during iterations, the FP64 primal point of a KKT check
//   yl  : m — the dual iterate
//   rpl : m+1 row pointers, cpl : n+1 column pointers (of the transpose)
// in two parts of the batch, one launch each:
//   FP32 part: FP32 iterates, x in global memory, 12n + 8m + 8 bytes of LDS,
//              or, when every LP of the part fits, x, its partial sum, tau_hat
//              and c in LDS as well (28n + 8m + 8 bytes: one 512-thread
//              workgroup per CU has the CU's LDS to itself);
//   FP64 part: FP64 iterates, x in LDS as well, 20n + 12m + 8 bytes.
// Small LPs with many empty columns drift in FP32 (their restarts see the
// rounding of the residuals and shrink the primal weight without end) and
// finish in FP64; for the large LPs, whose size sets the LDS per workgroup,
// FP32 halves the LDS so two workgroups share a CU.  Each iteration is
//   column pass  aty_j = sum_i K_ij y_i over column j's entries (CSR of K^T),
//                x+ = clip(x - tau (c - aty)), xbar = 2x+ - x
//   row pass     d_i = K_i xbar; y_i+ = proj(y_i + sigma_i (b_i - d_i))
// Both passes gather from LDS and write one value per row or column, so there
// are no atomics: the float atomics of a scatter formulation serialise in the
// LDS and cost as much as the rest of the row pass together.  Every sum runs
// in a fixed order, so results are the same from run to run.
//
// A group of 2^lw lanes (lw per LP and pass, from the line lengths) owns a
// line (row or column); each lane takes every 2^lw-th entry, the first EL of
// them loaded at once.  Each group works on NL lines per round (2-4 in the
// FP32 part, two in the FP64 part), so the loads of all are in flight
// together.  Line pointers come from LDS and the operands of the step are
// loaded with the entries, so a round waits for global memory once.
//
// Every check_every iterations the partial sums are folded into the FP64
// window sums and the reference KKT of the window average (and, for the
// adaptive scheme, of the current iterate) is evaluated in FP64 (pdhg.py:184-196
// on the lifted form), followed by the termination/restart rule of
// pdhg.py:225-244 or its adaptive variant.

#include "pdlp_common.hpp"

#include <algorithm>
#include <map>
#include <string>
#include <mutex>
#include <type_traits>

namespace pdlp {
namespace {

// Global-memory qualifier for an LP's arrays.  The LP's pointers are formed in
// the kernel from the descriptor, so without it the compiler treats them as
// generic and emits flat loads: 64-bit addresses in VGPR pairs, and every flat
// access also counts against the LDS wait counter.
#define GMEM __attribute__((address_space(1)))

constexpr int EL = 8;   // entries per lane loaded together
constexpr int KEL = 4;  // entries per lane loaded together in a KKT evaluation (FP64 sums: registers)
constexpr int KNL = 2;  // lines per group and round in a KKT evaluation
constexpr int NS = 6;   // KKT sums: rp^2, rd^2, primal obj, dual obj, |dx|^2, |dy|^2

struct LPView {
    int m, n;
    const GMEM int32_t* rp;     // m+1, relative to the nnz base
    const GMEM uint16_t* ci;    // relative to the column base
    const GMEM float* val;
    const GMEM int32_t* cp;     // n+1 column pointers of K^T, relative to the nnz base
    const GMEM uint16_t* ri;    // row index of each K^T entry, relative to the row base
    const GMEM float* tval;
    const GMEM float *c, *lo, *hi;
    const GMEM double *c64, *lo_shift, *ub_span;
    const GMEM float* b;
    const GMEM int8_t* kind;
    const GMEM double *b64, *b_shift;
    GMEM float *x, *x_psum, *tau;
    GMEM double* x_sum;
    GMEM float *y, *y_psum;
    GMEM double* y_sum;
    GMEM float *x_last, *y_last;   // last restart point (adaptive scheme)
    const GMEM double *rs, *cs;    // row / column scales of a scaled LP (rs == nullptr: unscaled)
    double norm_b, norm_c;
    float weight;             // starting primal weight w: tau = w tau_hat, sigma = sigma_hat / w
    int lr, lc;               // log2 lanes per row / per column
    bool adaptive;            // adaptive restarts (flags bit 0)
    bool fixed_weight;        // keep the primal weight (flags bit 1)
    bool move_ref;            // distances since the last restart (flags bit 2 clear) or from the start
};

__device__ __forceinline__ int log2_lanes(int t) {
    return __builtin_amdgcn_readfirstlane(t > 1 ? 31 - __builtin_clz(t) : 0);
}

__device__ __forceinline__ LPView view_of(const PdlpBatch& a, const PdlpBatchLP& lp) {
    LPView v;
    v.m = lp.m;
    v.n = lp.n;
    v.rp = (decltype(v.rp))(a.row_ptr + lp.rp_off);
    v.ci = (decltype(v.ci))(a.col_idx + lp.nnz_off);
    v.val = (decltype(v.val))(a.val + lp.nnz_off);
    v.cp = (decltype(v.cp))(a.t_col_ptr + lp.cp_off);
    v.ri = (decltype(v.ri))(a.t_row_idx + lp.nnz_off);
    v.tval = (decltype(v.tval))(a.t_val + lp.nnz_off);
    v.c = (decltype(v.c))(a.c + lp.col_off);
    v.lo = (decltype(v.lo))(a.lo + lp.col_off);
    v.hi = (decltype(v.hi))(a.hi + lp.col_off);
    v.c64 = (decltype(v.c64))(a.c64 + lp.col_off);
    v.lo_shift = (decltype(v.lo_shift))(a.lo_shift + lp.col_off);
    v.ub_span = (decltype(v.ub_span))(a.ub_span + lp.col_off);
    v.b = (decltype(v.b))(a.b + lp.row_off);
    v.kind = (decltype(v.kind))(a.row_kind + lp.row_off);
    v.b64 = (decltype(v.b64))(a.b64 + lp.row_off);
    v.b_shift = (decltype(v.b_shift))(a.b_shift + lp.row_off);
    v.x = (decltype(v.x))(a.x + lp.col_off);
    v.x_psum = (decltype(v.x_psum))(a.x_psum + lp.col_off);
    v.x_sum = (decltype(v.x_sum))(a.x_sum + lp.col_off);
    v.tau = (decltype(v.tau))(a.tau + lp.col_off);
    v.y = (decltype(v.y))(a.y + lp.row_off);
    v.y_sum = (decltype(v.y_sum))(a.y_sum + lp.row_off);
    v.y_psum = (decltype(v.y_psum))(a.y_psum + lp.row_off);
    v.x_last = (decltype(v.x_last))(a.x_last + lp.col_off);
    v.y_last = (decltype(v.y_last))(a.y_last + lp.row_off);
    v.rs = a.row_scale != nullptr ? (decltype(v.rs))(a.row_scale + lp.row_off) : nullptr;
    v.cs = a.col_scale != nullptr ? (decltype(v.cs))(a.col_scale + lp.col_off) : nullptr;
    v.norm_b = lp.norm_b;
    v.norm_c = lp.norm_c;
    v.weight = (float)lp.tau_scale;
    v.lr = log2_lanes(lp.tpr);
    v.lc = log2_lanes(lp.tpc);
    v.adaptive = (lp.flags & 1) != 0;
    v.fixed_weight = (lp.flags & 2) != 0;
    v.move_ref = (lp.flags & 4) == 0;
    return v;
}

// Cross-lane move inside a row of 16 lanes (DPP: no LDS traffic).
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over aligned groups of 2^lw lanes (lw uniform, 0..6); every lane of a
// group ends with the same sum.  Groups are entirely active or inactive.
__device__ __forceinline__ float group_sum_rt(float v, int lw) {
    if (lw >= 1) v += dpp<0xB1>(v);    // quad_perm [1,0,3,2]: lane ^ 1
    if (lw >= 2) v += dpp<0x4E>(v);    // quad_perm [2,3,0,1]: lane ^ 2
    if (lw >= 3) v += dpp<0x141>(v);   // row_half_mirror: lane i <-> 7 - i in each 8
    if (lw >= 4) v += dpp<0x140>(v);   // row_mirror: lane i <-> 15 - i in each 16
    if (lw >= 5) v += __shfl_xor(v, 16);
    if (lw >= 6) v += __shfl_xor(v, 32);
    return v;
}

__device__ __forceinline__ double group_sum_rt(double v, int lw) {
    for (int o = 1; o < (1 << lw); o <<= 1) v += __shfl_xor(v, o);
    return v;
}

// NL lines at once: s[k] += val[q] * vec[idx[q]] and a[k] += |val[q]| over the
// lane's entries of [b[k], e[k]) (entries lane, lane + 2^lw, ...), sums in T.
// The first E entries per lane of all lines are loaded before any is used.
template <int NL, int E, typename T, typename V>
__device__ __forceinline__ void dotn(const GMEM float* val, const GMEM uint16_t* idx, const V* vec, const int (&b)[NL],
                                     const int (&e)[NL], int lane, int lw, T (&s)[NL], T (&a)[NL]) {
    float v[NL][E];
    int c[NL][E];
#pragma unroll
    for (int k = 0; k < NL; ++k)
#pragma unroll
        for (int t = 0; t < E; ++t) {
            const int q = b[k] + lane + (t << lw);
            v[k][t] = q < e[k] ? val[q] : 0.f;
            c[k][t] = q < e[k] ? idx[q] : 0;
        }
#pragma unroll
    for (int k = 0; k < NL; ++k)
#pragma unroll
        for (int t = 0; t < E; ++t) {
            s[k] += (T)v[k][t] * (T)vec[c[k][t]];
            a[k] += (T)fabsf(v[k][t]);
        }
    // lines longer than E entries per lane (rare: the lane count covers most lines)
#pragma unroll
    for (int k = 0; k < NL; ++k)
        for (int q = b[k] + lane + (E << lw); q < e[k]; q += 1 << lw) {
            const T kv = (T)val[q];
            s[k] += kv * (T)vec[idx[q]];
            a[k] += fabs(kv);
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

// Reference KKT of (x, y): src 0 = current iterate, 1 = window average (w).
// Uses X as n doubles.  acc (all threads): rp^2, rd^2, primal and dual
// objective, and for the adaptive scheme the squared distances of the point
// from the last restart point in the step-size metric (dx^2/tau_hat, dy^2/sigma_hat).
template <int NT, typename T, bool XIN, bool SC>
__device__ __forceinline__ void kkt_eval(const LPView& v, double* X, const T* xl, const T* yl, const int* rpl,
                                         const int* cpl, int src, double w, double (&acc)[NS], double* red) {
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;
    // C1: primal point into LDS (FP64), primal objective, primal distance
    for (int j = threadIdx.x; j < v.n; j += NT) {
        double xj;
        if (src) xj = v.x_sum[j] / w;
        else if constexpr (XIN) xj = (double)xl[j];
        else xj = (double)v.x[j];
        X[j] = xj;
        acc[2] += v.c64[j] * (xj - v.lo_shift[j]);
        if (v.adaptive) {
            const double dx = xj - (double)v.x_last[j];
            acc[4] += dx * dx / (double)v.tau[j];
        }
    }
    __syncthreads();
    // C2: rows — primal residual, dual objective, slack dual residual, dual distance
    {
        const int lw = v.lr, lane = threadIdx.x & ((1 << lw) - 1), st = NT >> lw;
        for (int i0 = threadIdx.x >> lw; i0 < v.m; i0 += KNL * st) {
            int bb[KNL], ee[KNL];
            double d[KNL], as[KNL];
#pragma unroll
            for (int k = 0; k < KNL; ++k) {
                const int i = i0 + k * st;
                bb[k] = i < v.m ? rpl[i] : 0;
                ee[k] = i < v.m ? rpl[i + 1] : 0;
                d[k] = 0.0;
                as[k] = 0.0;
            }
            dotn<KNL, KEL, double, double>(v.val, v.ci, X, bb, ee, lane, lw, d, as);
#pragma unroll
            for (int k = 0; k < KNL; ++k) {
                d[k] = group_sum_rt(d[k], lw);
                as[k] = group_sum_rt(as[k], lw);
            }
#pragma unroll
            for (int k = 0; k < KNL; ++k) {
                const int i = i0 + k * st;
                if (lane == 0 && i < v.m) {
                    const bool le = v.kind[i] == PDLP_ROW_LE;
                    double res = d[k] - v.b64[i];
                    if (le && res < 0.0) res = 0.0;
                    const double yi = src ? v.y_sum[i] / w : (double)yl[i];
                    double yo = yi;               // residual and dual of the original row
                    if constexpr (SC) {
                        res /= v.rs[i];
                        yo = yi * v.rs[i];
                    }
                    acc[0] += res * res;
                    if (le && yo > 0.0) acc[1] += yo * yo;
                    acc[3] += v.b_shift[i] * yi;
                    if (v.adaptive) {
                        const double dy = yi - (double)v.y_last[i];
                        acc[5] += dy * dy * (as[k] > 1e-4 ? as[k] : 1e-4) / 0.999;
                    }
                }
            }
        }
    }
    // C3: columns — reduced costs c - K^T y of the point: dual residual / upper-bound-row duals.
    // The averaged y goes to LDS (in X, free after C2) when it fits (m <= n).
    const bool staged = src && v.m <= v.n;
    if (staged) {
        __syncthreads();
        for (int i = threadIdx.x; i < v.m; i += NT) X[i] = v.y_sum[i] / w;
        __syncthreads();
    }
    {
        const int lw = v.lc, lane = threadIdx.x & ((1 << lw) - 1), st = NT >> lw;
        for (int j0 = threadIdx.x >> lw; j0 < v.n; j0 += KNL * st) {
            int bb[KNL], ee[KNL];
            double s[KNL], u[KNL];
#pragma unroll
            for (int k = 0; k < KNL; ++k) {
                const int j = j0 + k * st;
                bb[k] = j < v.n ? cpl[j] : 0;
                ee[k] = j < v.n ? cpl[j + 1] : 0;
                s[k] = 0.0;
                u[k] = 0.0;
            }
            if (!src) {
                dotn<KNL, KEL, double, T>(v.tval, v.ri, yl, bb, ee, lane, lw, s, u);
            } else if (staged) {
                dotn<KNL, KEL, double, double>(v.tval, v.ri, X, bb, ee, lane, lw, s, u);
            } else {                                // averaged y gathered from global memory
#pragma unroll
                for (int k = 0; k < KNL; ++k)
                    for (int q = bb[k] + lane; q < ee[k]; q += 1 << lw) s[k] += (double)v.tval[q] * (v.y_sum[v.ri[q]] / w);
            }
#pragma unroll
            for (int k = 0; k < KNL; ++k) s[k] = group_sum_rt(s[k], lw);
#pragma unroll
            for (int k = 0; k < KNL; ++k) {
                const int j = j0 + k * st;
                if (lane == 0 && j < v.n) {
                    const double r = v.c64[j] - s[k];
                    const double rneg = r < 0.0 ? r : 0.0;
                    const double span = v.ub_span[j];
                    if (isfinite(span)) {
                        acc[3] += span * rneg;
                    } else {                      // reduced cost of the original column
                        double ro = rneg;
                        if constexpr (SC) ro /= v.cs[j];
                        acc[1] += ro * ro;
                    }
                }
            }
        }
    }
    block_sum_all<NT, NS>(acc, red);
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

// One LP from start to finish.  T = float: FP32 iterates, x in global memory,
// or with XR x, its partial sum, tau_hat and c in LDS as well; T = double:
// FP64 iterates, x in LDS.  NLF: lines per group and round of the FP32 part.
template <int NT, typename T, bool XR, bool SC, int NLF>
__device__ __forceinline__ void solve_one(const PdlpBatch& a, int k, float* lds, double* red) {
    constexpr bool RES = std::is_same<T, double>::value;
    constexpr bool XIN = RES || XR;    // x in LDS
    static_assert(!(RES && XR), "the FP64 part keeps tau_hat and c in global memory");
    // lines per group and round: all their loads are in flight together (FP64: two, for registers)
    constexpr int NL = RES ? 2 : NLF;
    const PdlpBatchLP lp = a.lps[k];
    const LPView v = view_of(a, lp);
    char* base = reinterpret_cast<char*>(lds);
    double* X = reinterpret_cast<double*>(base + (RES ? 8 * v.n : 0));
    T* xb = reinterpret_cast<T*>(X);
    T* yl = reinterpret_cast<T*>(base + (RES ? 16 * v.n : 8 * v.n));
    int* rpl = reinterpret_cast<int*>(yl + v.m);
    int* cpl = rpl + v.m + 1;
    // x in LDS: at the start (FP64 part) or after the line pointers (FP32 part with XR),
    // followed there by its partial sum, tau_hat and c
    T* xl = RES ? reinterpret_cast<T*>(base) : reinterpret_cast<T*>(cpl + v.n + 1);
    float* pl = reinterpret_cast<float*>(xl) + v.n;
    float* tl = pl + v.n;
    float* cl = tl + v.n;
    const int lr = v.lr, lc = v.lc;
    const int lane_r = threadIdx.x & ((1 << lr) - 1), lane_c = threadIdx.x & ((1 << lc) - 1);
    auto x_get = [&](int j) -> T {
        if constexpr (XIN) return xl[j];
        else return v.x[j];
    };
    auto x_set = [&](int j, T x) {
        if constexpr (XIN) xl[j] = x;
        else v.x[j] = x;
    };

    // ---- set-up: line pointers, x (resident) and y into LDS, sums cleared, restart point = start
    for (int i = threadIdx.x; i <= v.m; i += NT) rpl[i] = v.rp[i];
    for (int j = threadIdx.x; j <= v.n; j += NT) cpl[j] = v.cp[j];
    for (int i = threadIdx.x; i < v.m; i += NT) {
        yl[i] = (T)v.y[i];
        v.y_psum[i] = 0.f;
        v.y_sum[i] = 0.0;
        v.y_last[i] = v.y[i];
    }
    for (int j = threadIdx.x; j < v.n; j += NT) {
        if constexpr (XIN) xl[j] = (T)v.x[j];
        if constexpr (XR) {
            pl[j] = 0.f;
            cl[j] = v.c[j];
        }
        v.x_sum[j] = 0.0;
        v.x_psum[j] = 0.f;
        v.x_last[j] = v.x[j];
    }
    __syncthreads();
    // tau_hat_j = 0.999 / max(sum_i |K_ij|, 1e-4) from the column's entries
    for (int j = threadIdx.x >> lc; j < v.n; j += NT >> lc) {
        float s = 0.f;
        for (int q = cpl[j] + lane_c; q < cpl[j + 1]; q += 1 << lc) s += fabsf(v.tval[q]);
        s = group_sum_rt(s, lc);
        if (lane_c == 0) {
            const float t = 0.999f / (s > 1e-4f ? s : 1e-4f);
            v.tau[j] = t;
            if constexpr (XR) tl[j] = t;
        }
    }
    __syncthreads();

    // starting primal weight: given, or (NaN) ||sqrt(sigma_hat) b|| / ||sqrt(tau_hat) c||
    float w_start = v.weight;
    if (!(w_start > 0.f)) {
        double nw[2] = {0.0, 0.0};
        for (int j = threadIdx.x; j < v.n; j += NT) nw[1] += v.c64[j] * v.c64[j] * (double)v.tau[j];
        for (int i = threadIdx.x >> lr; i < v.m; i += NT >> lr) {
            double s = 0.0;
            for (int q = rpl[i] + lane_r; q < rpl[i + 1]; q += 1 << lr) s += fabs((double)v.val[q]);
            s = group_sum_rt(s, lr);
            if (lane_r == 0) nw[0] += v.b64[i] * v.b64[i] * 0.999 / (s > 1e-4 ? s : 1e-4);
        }
        block_sum_all<NT, 2>(nw, red);
        const double nb = sqrt(nw[0]), nc = sqrt(nw[1]);
        w_start = (nb > 1e-12 && nc > 1e-12) ? (float)fmin(fmax(nb / nc, 1e-6), 1e6) : 1.f;
    }

    double acc[NS], acc2[NS];
    kkt_eval<NT, T, XIN, SC>(v, X, xl, yl, rpl, cpl, 0, 1.0, acc, red);
    Kkt kk = kkt_of(acc, v);
    double ref_err = kk.err, err_prev = INFINITY;
    int status = isfinite(kk.err) ? PDLP_RUNNING : PDLP_NUMERICAL;
    int window = 0, restarts = 0, checks = 0, it = 0, done_it = 0, last_restart = 0;
    bool end_on_avg = false;
    double w_final = 1.0;
    T wgt = (T)w_start;

    const int ce = a.check_every;
    const int sr = NT >> lr, sc = NT >> lc;   // line groups per workgroup
    const T one = 1, two_ = 2, step_num = (T)0.999, floor_ = (T)1e-4;
    // running sums of the iterates: FP32 partial sums (folded at each check) or the FP64 window sums
    GMEM T* xps;
    GMEM T* yps;
    if constexpr (RES) {
        xps = v.x_sum;
        yps = v.y_sum;
    } else {
        xps = v.x_psum;
        yps = v.y_psum;
    }
    while (status == PDLP_RUNNING && it < a.max_iters) {
        const T inv_wgt = one / wgt;
        // ---- column pass: aty = K^T y (gathered from LDS), primal step, xbar
        for (int j0 = threadIdx.x >> lc; j0 < v.n; j0 += NL * sc) {
            int jj[NL], bb[NL], ee[NL];
            T xo[NL], tt[NL], lo[NL], hi[NL], cc[NL], ps[NL], s[NL], u[NL];
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const int j = j0 + k * sc;
                const bool in = j < v.n;
                jj[k] = in ? j : j0;
                bb[k] = in ? cpl[j] : 0;
                ee[k] = in ? cpl[j + 1] : 0;
            }
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const int j = jj[k];
                xo[k] = x_get(j);
                lo[k] = (T)v.lo[j];
                hi[k] = (T)v.hi[j];
                if constexpr (XR) {
                    tt[k] = tl[j];
                    cc[k] = cl[j];
                    ps[k] = pl[j];
                } else {
                    tt[k] = (T)v.tau[j];
                    if constexpr (RES) cc[k] = v.c64[j];
                    else cc[k] = v.c[j];
                    // the sums this round adds to, loaded with the other operands (not after the step)
                    ps[k] = lane_c == 0 ? xps[j] : (T)0;
                }
                s[k] = 0;
                u[k] = 0;
            }
            dotn<NL, EL, T, T>(v.tval, v.ri, yl, bb, ee, lane_c, lc, s, u);
#pragma unroll
            for (int k = 0; k < NL; ++k) s[k] = group_sum_rt(s[k], lc);
            if (lane_c == 0) {
#pragma unroll
                for (int k = 0; k < NL; ++k) {
                    const int j = j0 + k * sc;
                    if (j < v.n) {
                        T xn = xo[k] - (wgt * tt[k]) * (cc[k] - s[k]);
                        xn = xn < lo[k] ? lo[k] : xn;
                        xn = xn > hi[k] ? hi[k] : xn;
                        xb[j] = two_ * xn - xo[k];
                        x_set(j, xn);
                        if constexpr (XR) pl[j] = ps[k] + xn;
                        else xps[j] = ps[k] + xn;
                    }
                }
            }
        }
        __syncthreads();
        // ---- row pass: K xbar (gathered from LDS), dual step
        for (int i0 = threadIdx.x >> lr; i0 < v.m; i0 += NL * sr) {
            int ii[NL], bb[NL], ee[NL];
            T bv[NL], qs[NL], d[NL], u[NL];
            bool le[NL];
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const int i = i0 + k * sr;
                const bool in = i < v.m;
                ii[k] = in ? i : i0;
                bb[k] = in ? rpl[i] : 0;
                ee[k] = in ? rpl[i + 1] : 0;
            }
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const int i = ii[k];
                if constexpr (RES) bv[k] = v.b64[i];
                else bv[k] = v.b[i];
                le[k] = v.kind[i] == PDLP_ROW_LE;
                qs[k] = lane_r == 0 ? yps[i] : (T)0;
                d[k] = 0;
                u[k] = 0;
            }
            dotn<NL, EL, T, T>(v.val, v.ci, xb, bb, ee, lane_r, lr, d, u);
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                d[k] = group_sum_rt(d[k], lr);
                u[k] = group_sum_rt(u[k], lr);
            }
            if (lane_r == 0) {
#pragma unroll
                for (int k = 0; k < NL; ++k) {
                    const int i = i0 + k * sr;
                    if (i < v.m) {
                        T yn = yl[i] + inv_wgt * (step_num / (u[k] > floor_ ? u[k] : floor_)) * (bv[k] - d[k]);
                        if (le[k] && yn > (T)0) yn = 0;
                        yl[i] = yn;
                        yps[i] = qs[k] + yn;
                    }
                }
            }
        }
        __syncthreads();
        ++it;
        ++window;
        if (window % ce != 0) continue;

        // ---- check: fold partial sums (FP32 iterates), KKT of the candidate(s), decision
        if constexpr (!RES) {
            for (int j = threadIdx.x; j < v.n; j += NT) {
                if constexpr (XR) {
                    v.x_sum[j] += (double)pl[j];
                    pl[j] = 0.f;
                } else {
                    v.x_sum[j] += (double)v.x_psum[j];
                    v.x_psum[j] = 0.f;
                }
            }
            for (int i = threadIdx.x; i < v.m; i += NT) {
                v.y_sum[i] += (double)v.y_psum[i];
                v.y_psum[i] = 0.f;
            }
            __syncthreads();
        }
        const double w = (double)window;
        kkt_eval<NT, T, XIN, SC>(v, X, xl, yl, rpl, cpl, 1, w, acc, red);
        Kkt ka = kkt_of(acc, v);
        bool use_avg = true;
        if (v.adaptive) {                       // the better of average and current iterate
            kkt_eval<NT, T, XIN, SC>(v, X, xl, yl, rpl, cpl, 0, 1.0, acc2, red);
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
                wgt = (T)clamp_weight(exp(lw), (double)w_start);
            }
            ref_err = kk.err;
            err_prev = INFINITY;
            last_restart = it;
            ++restarts;
            for (int j = threadIdx.x; j < v.n; j += NT) {
                if (use_avg) x_set(j, (T)(v.x_sum[j] / w));
                v.x_sum[j] = 0.0;
                if (v.move_ref) v.x_last[j] = (float)x_get(j);
            }
            for (int i = threadIdx.x; i < v.m; i += NT) {
                if (use_avg) yl[i] = (T)(v.y_sum[i] / w);
                v.y_sum[i] = 0.0;
                if (v.move_ref) v.y_last[i] = (float)yl[i];
            }
            window = 0;
        } else {
            err_prev = kk.err;
        }
        __syncthreads();
    }

    // ---- outputs: FP64 solution in x_sum / y_sum, last iterate in x / y
    if (status == PDLP_OPTIMAL && end_on_avg) {
        for (int j = threadIdx.x; j < v.n; j += NT) v.x_sum[j] = v.x_sum[j] / w_final;
        for (int i = threadIdx.x; i < v.m; i += NT) v.y_sum[i] = v.y_sum[i] / w_final;
    } else {
        for (int j = threadIdx.x; j < v.n; j += NT) v.x_sum[j] = (double)x_get(j);
        for (int i = threadIdx.x; i < v.m; i += NT) v.y_sum[i] = (double)yl[i];
        if (status == PDLP_RUNNING) done_it = it;
    }
    if constexpr (XIN)
        for (int j = threadIdx.x; j < v.n; j += NT) v.x[j] = (float)xl[j];
    for (int i = threadIdx.x; i < v.m; i += NT) v.y[i] = (float)yl[i];
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

// Persistent workgroups over one part of the work order: LPs order[first ..
// first + count), taken through the counter *queue.  T = float: FP32 part,
// T = double: FP64 part (each instantiation has its own register allocation).
template <int NT, typename T, bool XR, bool SC, int NLF>
__global__ void __launch_bounds__(NT) batch_kernel(PdlpBatch a, int first, int count, int* queue, int lds_limit) {
    extern __shared__ float lds[];
    __shared__ double red[NS * (NT / kWave)];
    __shared__ int s_q;
    constexpr long NB = std::is_same<T, double>::value ? 20 : (XR ? 28 : 12);
    constexpr long MB = std::is_same<T, double>::value ? 12 : 8;
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
        __syncthreads();
        // uniform values: the LP's descriptor and pointers then live in scalar registers
        const int q = __builtin_amdgcn_readfirstlane(s_q);
        __syncthreads();
        if (q >= count) return;                   // every wave of the workgroup leaves here
        const int k = __builtin_amdgcn_readfirstlane(a.order[first + q]);
        const int m = __builtin_amdgcn_readfirstlane(a.lps[k].m);
        const int n = __builtin_amdgcn_readfirstlane(a.lps[k].n);
        // an LP that does not fit the launch's LDS is not run and says so (the host
        // checked the part's largest need, so only inconsistent descriptors get here)
        if (NB * n + MB * m + 8 <= lds_limit) {
            solve_one<NT, T, XR, SC, NLF>(a, k, lds, red);
        } else if (threadIdx.x == 0) {
            PdlpBatchOut o = {};
            o.status = PDLP_SKIPPED;
            o.rel_kkt = NAN;
            a.out[k] = o;
        }
    }
}

template <int NT, typename T, bool XR, bool SC, int NLF>
int launch(const PdlpBatch* a, int first, int count, int* queue, int lds, int grid, hipStream_t s) {
    const hipError_t e = hipFuncSetAttribute((const void*)batch_kernel<NT, T, XR, SC, NLF>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
    batch_kernel<NT, T, XR, SC, NLF><<<grid, NT, lds, s>>>(*a, first, count, queue, lds);
    return hip_fail(hipGetLastError(), "batch_kernel launch");
}

// The FP32 part's lines per round (2, 3, 4); the FP64 part always takes two.
template <int NT, typename T, bool XR, bool SC>
int launch_lines(const PdlpBatch* a, int first, int count, int* queue, int lds, int grid, hipStream_t s) {
    if constexpr (std::is_same<T, double>::value) {
        return launch<NT, T, XR, SC, 2>(a, first, count, queue, lds, grid, s);
    } else {
        switch (a->lines ? a->lines : 3) {
            case 2: return launch<NT, T, XR, SC, 2>(a, first, count, queue, lds, grid, s);
            case 3: return launch<NT, T, XR, SC, 3>(a, first, count, queue, lds, grid, s);
            case 4: return launch<NT, T, XR, SC, 4>(a, first, count, queue, lds, grid, s);
            default: return fail(PDLP_E_TUNE, "lines must be 2, 3 or 4");
        }
    }
}

template <typename T, bool XR>
int launch_part(const PdlpBatch* a, int first, int count, int* queue, int block, int lds, int grid, int cus,
                hipStream_t s) {
    if (count == 0) return 0;
    if (lds <= 0 || lds > 160 * 1024) return fail(PDLP_E_LIMIT, "LDS per workgroup must be in (0, 160 KiB]");
    const int g = std::min(count, grid > 0 ? grid : 2 * cus);
    switch (block) {
        // scaled LPs (row_scale set) run the instantiation whose checks unscale
        case 256: return a->row_scale ? launch_lines<256, T, XR, true>(a, first, count, queue, lds, g, s)
                                      : launch_lines<256, T, XR, false>(a, first, count, queue, lds, g, s);
        case 512: return a->row_scale ? launch_lines<512, T, XR, true>(a, first, count, queue, lds, g, s)
                                      : launch_lines<512, T, XR, false>(a, first, count, queue, lds, g, s);
        default: return fail(PDLP_E_TUNE, "block must be 256 or 512");
    }
}

// A second stream per device (default priority, created on first use and kept
// until the process ends) on which the FP64 part runs beside the FP32 part:
// the FP64 part's time is set by a few small LPs' long, latency-bound chains,
// which overlap the FP32 part's streaming.  Fork and join through events, so
// the caller's stream still orders everything after the call.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
std::mutex g_side_mu;
std::map<std::pair<int, hipStream_t>, SideStream> g_side;

// One side stream per (device, caller stream): independent callers' FP64 parts do
// not queue behind each other.
int side_stream(int dev, hipStream_t caller, SideStream& out) {
    SideStream& ss = g_side[{dev, caller}];
    if (ss.s == nullptr) {
        if (int rc = hip_fail(hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking), "hipStreamCreate")) return rc;
        if (int rc = hip_fail(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming), "hipEventCreate")) return rc;
        if (int rc = hip_fail(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming), "hipEventCreate")) return rc;
    }
    out = ss;
    return 0;
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" int pdlp_batch_solve(const PdlpBatch* a, void* stream) {
    if (a == nullptr) return fail(PDLP_E_ARG, "PdlpBatch is NULL");
    if (a->count < 0) return fail(PDLP_E_ARG, "count must be >= 0");
    if (a->n_resident < 0 || a->n_resident > a->count) return fail(PDLP_E_ARG, "n_resident must be in [0, count]");
    if (a->parts < 0 || a->parts > 3) return fail(PDLP_E_ARG, "parts must be 0..3");
    if (a->count == 0) return 0;
    if (!a->lps || !a->order || !a->queue || !a->row_ptr || !a->col_idx || !a->val || !a->t_col_ptr ||
        !a->t_row_idx || !a->t_val || !a->c || !a->lo || !a->hi || !a->c64 || !a->lo_shift || !a->ub_span ||
        !a->b || !a->row_kind || !a->b64 || !a->b_shift || !a->x || !a->x_psum || !a->x_sum || !a->tau || !a->y ||
        !a->y_sum || !a->out || !a->x_last || !a->y_last || !a->y_psum)
        return fail(PDLP_E_ARG, "a required PdlpBatch pointer is NULL");
    if ((a->row_scale == nullptr) != (a->col_scale == nullptr))
        return fail(PDLP_E_ARG, "row_scale and col_scale must both be set or both be NULL");
    if (a->check_every <= 0 || a->max_iters < 0) return fail(PDLP_E_ARG, "check_every > 0 and max_iters >= 0");
    if (a->lines != 0 && (a->lines < 2 || a->lines > 4)) return fail(PDLP_E_TUNE, "lines must be 2, 3 or 4 (0 = 3)");
    const int n32 = a->count - a->n_resident;
    const int parts = a->parts ? a->parts : 3;
    if ((parts & 1) && n32 > 0) {
        if (a->lds_need <= 0) return fail(PDLP_E_ARG, "lds_need must be set for the FP32 part");
        if (a->lds_need > a->lds_bytes)
            return fail(PDLP_E_LIMIT, "an LP of the FP32 part needs " + std::to_string(a->lds_need) +
                                          " bytes of LDS, more than lds_bytes = " + std::to_string(a->lds_bytes));
    }
    if ((parts & 2) && a->n_resident > 0) {
        if (a->lds_need_resident <= 0) return fail(PDLP_E_ARG, "lds_need_resident must be set for the FP64 part");
        if (a->lds_need_resident > a->lds_resident)
            return fail(PDLP_E_LIMIT, "an LP of the FP64 part needs " + std::to_string(a->lds_need_resident) +
                                          " bytes of LDS, more than lds_resident = " +
                                          std::to_string(a->lds_resident));
    }
    hipStream_t s = as_stream(stream);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    auto part32 = [&](hipStream_t st) {
        return a->x_in_lds ? launch_part<float, true>(a, 0, n32, a->queue, a->block, a->lds_bytes, a->grid, cus, st)
                           : launch_part<float, false>(a, 0, n32, a->queue, a->block, a->lds_bytes, a->grid, cus, st);
    };
    auto part64 = [&](hipStream_t st) {
        return launch_part<double, false>(a, n32, a->n_resident, a->queue + 1,
                                          a->block_resident ? a->block_resident : 256, a->lds_resident,
                                          a->grid_resident, cus, st);
    };
    if (parts == 3 && n32 > 0 && a->n_resident > 0) {
        // both parts: the FP64 part on the side stream, concurrently.  The lock keeps
        // each call's record/wait pairs together when host threads share a device.
        std::lock_guard<std::mutex> lock(g_side_mu);
        SideStream ss;
        if (int rc = side_stream(dev, s, ss)) return rc;
        if (int rc = hip_fail(hipEventRecord(ss.fork, s), "hipEventRecord")) return rc;
        if (int rc = hip_fail(hipStreamWaitEvent(ss.s, ss.fork, 0), "hipStreamWaitEvent")) return rc;
        // once forked, the caller's stream always waits for the side stream, also after an
        // error, so the caller never frees buffers that a running FP64 part still reads
        int rc = part64(ss.s);
        const int rc_rec = hip_fail(hipEventRecord(ss.join, ss.s), "hipEventRecord");
        if (rc == 0 && rc_rec == 0) rc = part32(s);
        if (rc_rec != 0) {
            // no join event: wait here for the side stream's work instead
            (void)hipStreamSynchronize(ss.s);
            return rc ? rc : rc_rec;
        }
        const int rc_wait = hip_fail(hipStreamWaitEvent(s, ss.join, 0), "hipStreamWaitEvent");
        return rc ? rc : rc_wait;
    }
    if (parts & 1) {
        if (int rc = part32(s)) return rc;
    }
    if (parts & 2) return part64(s);
    return 0;
}
