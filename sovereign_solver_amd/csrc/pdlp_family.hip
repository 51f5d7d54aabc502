// Shared-matrix batched restarted PDHG ("family" engine).
//
// B box-form LPs share one matrix K; they differ in c, b, lo, hi.  The LPs
// step in lockstep and every per-LP vector is stored in tiles of 64 LPs with
// the LP index fastest, so a wavefront owns one tile: lane = LP, and one
// wavefront-wide step updates row i (or column j) of 64 LPs.  K's CSR entries
// are then the same for every lane (scalar loads) and each gathered value is
// one coalesced 256-byte read of 64 LPs.  Per iteration:
//
//   column pass  x+ = clip(x - w tau_hat (c - K^T y)),  xbar = 2x+ - x
//   row pass     y+ = proj(y + (sigma_hat / w) (b - K xbar))
//
// with FP32 iterates and FP32 partial sums of the window.  Every check_every
// iterations: the partial sums are folded into FP64 window sums, the
// reference KKT (solver/lp/pdhg.py:184-196 on the lifted form) of the window
// average and of the current iterate is reduced in FP64 for every LP, and each
// LP decides on termination / restart (pdhg.py:225-244, or the adaptive
// scheme) exactly as the persistent batched kernel does.  Finished LPs are
// frozen (their lane stops writing; a tile whose LPs are all finished is
// skipped) and hold their FP64 solution in x_sum / y_sum.

#include "pdlp_common.hpp"

#include <algorithm>

namespace pdlp {
namespace {

constexpr int T = PDLP_FAM_TILE;     // LPs per tile = lanes per wavefront
constexpr int NT = 256;              // threads per workgroup
constexpr int NWV = NT / kWave;      // wavefronts per workgroup
constexpr int NSC = PDLP_FAM_NSUM / 2;   // sums per candidate (window average at 0, current at NSC):
// 0 primal residual^2, 1 dual residual^2, 2 primal objective, 3 dual objective,
// 4 |dx|^2, 5 |dy|^2 (step-size metric), 6 b_shift'y, 7 max over the shifted box of
// (K'y)'x, 8 sum |b_shift_i y_i|, 9 sum of the |terms| of 7 (Farkas test of y)
constexpr int NS2 = PDLP_FAM_NSUM;
constexpr int KCH = PDLP_FAM_KKT_CHUNKS;

static_assert(T == kWave, "one tile per wavefront");

// ACT_COLD: restart from the cold start (x = clip(0, lo, hi), y = 0), window sums cleared.
// ACT_*_STOP: that action, then stop at the LP's iteration limit with the resulting
// iterate as solution (what ACT_DONE_CUR at pdlp_family_finish would leave).
enum Action : int32_t {
    ACT_NONE = 0, ACT_RESTART_AVG = 1, ACT_RESTART_CUR = 2, ACT_DONE_AVG = 3, ACT_DONE_CUR = 4, ACT_COLD = 5,
    ACT_RESTART_AVG_STOP = 6, ACT_RESTART_CUR_STOP = 7, ACT_COLD_STOP = 8
};

// The cold start's x_j = clip(0, lo_j, hi_j) for the lane's LP.
__device__ __forceinline__ float cold_x(const PdlpFamily& f, int tile, int j, int lane);

__device__ __forceinline__ size_t at(int tile, int len, int d, int lane) {
    return ((size_t)tile * len + d) * T + lane;
}

template <typename V>
__device__ __forceinline__ V ldv(const V* p, bool shared, int tile, int len, int d, int lane) {
    return shared ? p[d] : p[at(tile, len, d, lane)];
}

__device__ __forceinline__ float cold_x(const PdlpFamily& f, int tile, int j, int lane) {
    const float lo = ldv(f.lo, f.shared & PDLP_FAM_SHARED_LO, tile, f.n, j, lane);
    const float hi = ldv(f.hi, f.shared & PDLP_FAM_SHARED_HI, tile, f.n, j, lane);
    float v = 0.f < lo ? lo : 0.f;
    return v > hi ? hi : v;
}

// wavefront index inside the workgroup, as a uniform value
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// d += val[q] * v(LP lane, idx[q]) for the S entries q0 .. q0 + S - 1, in ascending q
template <int S>
__device__ __forceinline__ void tail_dot(const int32_t* __restrict__ idx, const float* __restrict__ val, int q0,
                                         const float* __restrict__ v, int tile, int len, int lane, float& d) {
    float g[S], w[S];
    int c[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
        c[u] = idx[q0 + u];
        w[u] = val[q0 + u];
    }
#pragma unroll
    for (int u = 0; u < S; ++u) g[u] = v[at(tile, len, c[u], lane)];
#pragma unroll
    for (int u = 0; u < S; ++u) d += w[u] * g[u];
}

// sum over q in [beg, end) of val[q] * v(LP lane, index idx[q]) in ascending q.
// The (uniform) index/value pairs are taken 8 at a time so that 8 gathers are
// in flight together; the summation order stays sequential.
__device__ __forceinline__ float gather_dot(const int32_t* __restrict__ idx, const float* __restrict__ val, int beg,
                                            int end, const float* __restrict__ v, int tile, int len, int lane) {
    float d = 0.f;
    int q = beg;
    // 16 entries' loads in flight at a time, then 8, then the rest (same summation order)
    for (; q + 16 <= end; q += 16) {
        float g[16], w[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            w[u] = val[q + u];
            g[u] = v[at(tile, len, idx[q + u], lane)];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) d += w[u] * g[u];
    }
    for (; q + 8 <= end; q += 8) {
        float g[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            w[u] = val[q + u];
            g[u] = v[at(tile, len, idx[q + u], lane)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) d += w[u] * g[u];
    }
    // the rest (fewer than 8): a path per count, so the index loads and then the
    // gathers of exactly those entries go out together (added in the same order)
    switch (end - q) {
        case 1: tail_dot<1>(idx, val, q, v, tile, len, lane, d); break;
        case 2: tail_dot<2>(idx, val, q, v, tile, len, lane, d); break;
        case 3: tail_dot<3>(idx, val, q, v, tile, len, lane, d); break;
        case 4: tail_dot<4>(idx, val, q, v, tile, len, lane, d); break;
        case 5: tail_dot<5>(idx, val, q, v, tile, len, lane, d); break;
        case 6: tail_dot<6>(idx, val, q, v, tile, len, lane, d); break;
        case 7: tail_dot<7>(idx, val, q, v, tile, len, lane, d); break;
        default: break;
    }
    return d;
}

// Workgroup -> (tile, chunk of the tile's rows / columns).  With XCD affinity
// the grid is 1-D and workgroup b runs on XCD b % 8 (round-robin dispatch):
// with 8 or more tiles XCD x owns tiles x, x+8, ...; with fewer, each tile
// owns 8 / tiles XCDs that share its chunks.  P = chunks per tile.
struct Place {
    int tile, chunk;
    bool ok;
};

__device__ __forceinline__ Place place(const PdlpFamily& f, int P) {
    Place r;
    if (!f.xcd_affinity) {
        r.tile = blockIdx.y;
        r.chunk = blockIdx.x;
        r.ok = true;
        return r;
    }
    const int b = blockIdx.x, x = b & 7, slot = b >> 3, T = f.tiles;
    if (T >= 8) {
        const int lt = slot / P;
        r.chunk = slot - lt * P;
        r.tile = x + 8 * lt;
        r.ok = r.tile < T;
    } else {
        const int g = 8 / T;
        r.tile = x / g;
        r.chunk = slot * g + (x - r.tile * g);
        r.ok = x < T * g && r.chunk < P;
    }
    return r;
}

// ---------------------------------------------------------------- iteration
// Column update of column j for the wavefront's 64 LPs, given K^T y.
__device__ __forceinline__ void column_update(const PdlpFamily& f, int tile, int lane, bool act, float wgt, int j,
                                              float aty) {
    const bool shc = f.shared & PDLP_FAM_SHARED_C, shl = f.shared & PDLP_FAM_SHARED_LO,
               shh = f.shared & PDLP_FAM_SHARED_HI;
    const size_t ix = at(tile, f.n, j, lane);
    const float xo = f.x[ix];
    float xn = xo - (wgt * f.tau_hat[j]) * (ldv(f.c, shc, tile, f.n, j, lane) - aty);
    const float lo = ldv(f.lo, shl, tile, f.n, j, lane), hi = ldv(f.hi, shh, tile, f.n, j, lane);
    xn = xn < lo ? lo : xn;
    xn = xn > hi ? hi : xn;
    if (act) {
        f.x[ix] = xn;
        f.xbar[ix] = 2.f * xn - xo;
        f.x_psum[ix] += xn;
    }
}

// Row update of row i for the wavefront's 64 LPs, given K xbar.
__device__ __forceinline__ void row_update(const PdlpFamily& f, int tile, int lane, bool act, float inv_wgt, int i,
                                           float d) {
    const bool shb = f.shared & PDLP_FAM_SHARED_B;
    const size_t ix = at(tile, f.m, i, lane);
    const float sigma = inv_wgt * f.sigma_hat[i];
    float yn = f.y[ix] + sigma * (ldv(f.b, shb, tile, f.m, i, lane) - d);
    if (f.row_kind[i] == PDLP_ROW_LE && yn > 0.f) yn = 0.f;
    if (act) {
        f.y[ix] = yn;
        f.y_psum[ix] += yn;
    }
}

// One long row (or column) per workgroup of NL wavefronts: each wavefront
// takes a contiguous share of the nonzeros, the shares meet in LDS in
// wavefront order, and wavefront 0 applies the update.
constexpr int NL = 8;

__device__ __forceinline__ float split_dot(const int32_t* __restrict__ idx, const float* __restrict__ val, int beg,
                                           int end, const float* __restrict__ v, int tile, int len, int lane, int w) {
    __shared__ float part[NL][T];
    const int per = (end - beg + NL - 1) / NL;
    const int b = min(beg + w * per, end), e = min(b + per, end);
    part[w][lane] = gather_dot(idx, val, b, e, v, tile, len, lane);
    __syncthreads();
    float d = 0.f;
    if (w == 0)
#pragma unroll
        for (int u = 0; u < NL; ++u) d += part[u][lane];
    return d;
}

// Column pass in one launch.  Chunks [0, long_cols) of a tile take one long
// column each (the last long_cols columns) on the whole workgroup, which then
// has NL wavefronts (NTH = NL * 64: the launcher's choice whenever there are
// long columns); the other chunks take NTH / 64 * items_per_wave short
// columns, items_per_wave per wavefront.  A long column's sum is split over
// NL wavefronts and a short column's is one lane's at every launch shape, so
// the shape changes no result.  Long and short columns run side by side.
template <int NTH>
__global__ void __launch_bounds__(NTH) primal_kernel(PdlpFamily f, int P) {
    const Place pl = place(f, P);
    if (!pl.ok) return;
    const int tile = pl.tile;
    const uint64_t mask = f.run_mask[tile];
    if (mask == 0) return;
    const int lane = lane_id(), w = wave_id();
    const bool act = (mask >> lane) & 1ull;
    const int k = tile * T + lane;
    const float wgt = act ? f.lp[k].weight : 1.f;
    const int nl = f.long_cols;
    if (pl.chunk < nl) {                               // uniform over the workgroup
        if constexpr (NTH == NL * kWave) {
            const int j = f.n - nl + pl.chunk;
            const float aty =
                split_dot(f.t_col_idx, f.t_val, f.t_row_ptr[j], f.t_row_ptr[j + 1], f.y, tile, f.m, lane, w);
            if (w == 0) column_update(f, tile, lane, act, wgt, j, aty);
        }
        return;
    }
    constexpr int NW = NTH / kWave;
    const int p0 = ((pl.chunk - nl) * NW + w) * f.items_per_wave;
    const int p1 = min(p0 + f.items_per_wave, f.n - nl);
    const bool shc = f.shared & PDLP_FAM_SHARED_C, shl = f.shared & PDLP_FAM_SHARED_LO,
               shh = f.shared & PDLP_FAM_SHARED_HI;
    for (int p = p0; p < p1; ++p) {
        const int j = p;
        // the column's own data first, so its loads are in flight with the gather's
        const size_t ix = at(tile, f.n, j, lane);
        const float xo = f.x[ix], th = f.tau_hat[j], cj = ldv(f.c, shc, tile, f.n, j, lane);
        const float lo = ldv(f.lo, shl, tile, f.n, j, lane), hi = ldv(f.hi, shh, tile, f.n, j, lane);
        const float ps = f.x_psum[ix];
        const float aty = gather_dot(f.t_col_idx, f.t_val, f.t_row_ptr[j], f.t_row_ptr[j + 1], f.y, tile, f.m, lane);
        float xn = xo - (wgt * th) * (cj - aty);
        xn = xn < lo ? lo : xn;
        xn = xn > hi ? hi : xn;
        if (act) {
            f.x[ix] = xn;
            f.xbar[ix] = 2.f * xn - xo;
            f.x_psum[ix] = ps + xn;
        }
    }
}

// Row pass in one launch, laid out as the column pass: chunks [0, long_rows)
// take one long row each (the last long_rows rows) on NL wavefronts, the
// others short rows, items_per_wave per wavefront.
template <int NTH>
__global__ void __launch_bounds__(NTH) dual_kernel(PdlpFamily f, int P) {
    const Place pl = place(f, P);
    if (!pl.ok) return;
    const int tile = pl.tile;
    const uint64_t mask = f.run_mask[tile];
    if (mask == 0) return;
    const int lane = lane_id(), w = wave_id();
    const bool act = (mask >> lane) & 1ull;
    const int k = tile * T + lane;
    const float inv_wgt = act ? 1.f / f.lp[k].weight : 1.f;
    const int nl = f.long_rows;
    if (pl.chunk < nl) {                               // uniform over the workgroup
        if constexpr (NTH == NL * kWave) {
            const int i = f.m - nl + pl.chunk;
            const float d = split_dot(f.col_idx, f.val, f.row_ptr[i], f.row_ptr[i + 1], f.xbar, tile, f.n, lane, w);
            if (w == 0) row_update(f, tile, lane, act, inv_wgt, i, d);
        }
        return;
    }
    constexpr int NW = NTH / kWave;
    const int p0 = ((pl.chunk - nl) * NW + w) * f.items_per_wave;
    const int p1 = min(p0 + f.items_per_wave, f.m - nl);
    const bool shb = f.shared & PDLP_FAM_SHARED_B;
    for (int p = p0; p < p1; ++p) {
        const int i = p;
        // the row's own data first, so its loads are in flight with the gather's
        const size_t ix = at(tile, f.m, i, lane);
        const float yo = f.y[ix], sh = f.sigma_hat[i], bi = ldv(f.b, shb, tile, f.m, i, lane);
        const bool le = f.row_kind[i] == PDLP_ROW_LE;
        const float ps = f.y_psum[ix];
        const float d = gather_dot(f.col_idx, f.val, f.row_ptr[i], f.row_ptr[i + 1], f.xbar, tile, f.n, lane);
        float yn = yo + (inv_wgt * sh) * (bi - d);
        if (le && yn > 0.f) yn = 0.f;
        if (act) {
            f.y[ix] = yn;
            f.y_psum[ix] = ps + yn;
        }
    }
}

// ------------------------------------------------------------------- checks
// The value is computed here, not inside a later branch that alone uses it (the
// compiler would otherwise move its load into that branch and wait on it there).
__device__ __forceinline__ void keep_loaded(double v) { asm volatile("" ::"v"(v)); }

// Per-lane sums of both candidates (window average a[0..6), current a[6..12)),
// reduced over the workgroup's wavefronts and stored as one chunk.
__device__ __forceinline__ void store_chunk(const PdlpFamily& f, int chunk, int tile, double (&a)[NS2]) {
    __shared__ double red[NWV - 1][NS2][T];
    const int lane = lane_id(), w = wave_id();
    if (w > 0)
#pragma unroll
        for (int s = 0; s < NS2; ++s) red[w - 1][s][lane] = a[s];
    __syncthreads();
    if (w == 0) {
        const size_t cpad = (size_t)f.tiles * T;
#pragma unroll
        for (int s = 0; s < NS2; ++s) {
            double t = a[s];
#pragma unroll
            for (int v = 0; v < NWV - 1; ++v) t += red[v][s][lane];
            f.partials[((size_t)chunk * NS2 + s) * cpad + (size_t)tile * T + lane] = t;
        }
    }
}

// max of g t over t in [lo - lo_shift, span]: lo - lo_shift is 0 for a finite
// lower bound, -inf for none; an unbounded maximum is +inf (y is then no
// certificate).
__device__ __forceinline__ double support(double g, double span, bool lo_inf) {
    if (g > 0.0) return isfinite(span) ? g * span : INFINITY;
    if (g < 0.0 && lo_inf) return INFINITY;
    return 0.0;
}

// Columns: fold the x partial sums; primal objective and distance of both
// candidates; K^T y of both candidates (y window sums not folded yet: the row
// pass folds them after this pass) for the dual residual / bound-row duals.
__global__ void __launch_bounds__(NT) kkt_cols_kernel(PdlpFamily f, int add, int P) {
    const Place pl = place(f, P);
    if (!pl.ok) return;
    const int tile = pl.tile;
    const uint64_t mask = f.run_mask[tile];
    if (mask == 0) return;                      // whole workgroup: no barrier is skipped by part of it
    const int lane = lane_id(), w = wave_id();
    const bool act = (mask >> lane) & 1ull;
    const int k = tile * T + lane;
    const int win = act ? f.lp[k].window + add : 0;   // window length including this check's iterations
    const double wd = (double)win;
    const bool adaptive = f.flags & 1;
    const bool shc = f.shared & PDLP_FAM_SHARED_C64, shl = f.shared & PDLP_FAM_SHARED_LOSHIFT,
               shu = f.shared & PDLP_FAM_SHARED_UBSPAN, shl2 = f.shared & PDLP_FAM_SHARED_LO;
    double a[NS2];
#pragma unroll
    for (int s = 0; s < NS2; ++s) a[s] = 0.0;
    for (int p = pl.chunk * NWV + w; p < f.n; p += P * NWV) {
        const int j = f.n - 1 - p;                       // long columns first
        const size_t ix = at(tile, f.n, j, lane);
        const double xs = f.x_sum[ix] + (double)f.x_psum[ix];
        if (act) {
            f.x_sum[ix] = xs;
            f.x_psum[ix] = 0.f;
        }
        const double xc = (double)f.x[ix];
        const double xa = win > 0 ? xs / wd : xc;
        const double c = ldv(f.c64, shc, tile, f.n, j, lane);
        const double ls = ldv(f.lo_shift, shl, tile, f.n, j, lane);
        a[2] += c * (xa - ls);
        a[NSC + 2] += c * (xc - ls);
        if (adaptive) {
            const double xl = (double)f.x_last[ix], th = (double)f.tau_hat[j];
            const double da = xa - xl, dc = xc - xl;
            a[4] += da * da / th;
            a[NSC + 4] += dc * dc / th;
        }
        double sa = 0.0, sc = 0.0;
        const int beg = f.t_row_ptr[j], end = f.t_row_ptr[j + 1];
        // every entry's y, y_sum and y_psum are loaded whatever the window (no branch
        // around a load), four entries' loads together, then summed in ascending q
        int q = beg;
        for (; q + 4 <= end; q += 4) {
            double kv[4], yc[4], ys[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                kv[u] = (double)f.t_val[q + u];
                const size_t iy = at(tile, f.m, f.t_col_idx[q + u], lane);
                yc[u] = (double)f.y[iy];
                ys[u] = f.y_sum[iy] + (double)f.y_psum[iy];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) keep_loaded(ys[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double ya = win > 0 ? ys[u] / wd : yc[u];
                sa += kv[u] * ya;
                sc += kv[u] * yc[u];
            }
        }
        for (; q < end; ++q) {
            const double kv = (double)f.t_val[q];
            const size_t iy = at(tile, f.m, f.t_col_idx[q], lane);
            const double yc = (double)f.y[iy];
            const double ys = f.y_sum[iy] + (double)f.y_psum[iy];
            keep_loaded(ys);
            const double ya = win > 0 ? ys / wd : yc;
            sa += kv * ya;
            sc += kv * yc;
        }
        const double span = ldv(f.ub_span, shu, tile, f.n, j, lane);
        const double ra = fmin(c - sa, 0.0), rc = fmin(c - sc, 0.0);
        if (isfinite(span)) {
            a[3] += span * ra;
            a[NSC + 3] += span * rc;
        } else {
            a[1] += ra * ra;
            a[NSC + 1] += rc * rc;
        }
        // Farkas test: max over x - lo_shift in [lo - lo_shift, span] of (K'y)_j (x - lo_shift)_j
        const bool lo_inf = isinf(ldv(f.lo, shl2, tile, f.n, j, lane));
        const double ua = support(sa, span, lo_inf), uc = support(sc, span, lo_inf);
        a[7] += ua;
        a[9] += fabs(ua);
        a[NSC + 7] += uc;
        a[NSC + 9] += fabs(uc);
    }
    store_chunk(f, pl.chunk, tile, a);
}

// Row i's share of the check sums, given da = K_i x_avg and dc = K_i x (FP64):
// fold the y partial sums; primal residual, dual objective, sign violation of
// y on <= rows, dual distance, for both candidates.
__device__ __forceinline__ void row_terms(const PdlpFamily& f, int tile, int lane, bool act, int win, double wd,
                                          bool adaptive, int i, double da, double dc, double (&a)[NS2]) {
    const bool shb = f.shared & PDLP_FAM_SHARED_B64, shs = f.shared & PDLP_FAM_SHARED_BSHIFT;
    const size_t iy = at(tile, f.m, i, lane);
    const double ys = f.y_sum[iy] + (double)f.y_psum[iy];
    if (act) {
        f.y_sum[iy] = ys;
        f.y_psum[iy] = 0.f;
    }
    const double yc = (double)f.y[iy];
    const double ya = win > 0 ? ys / wd : yc;
    const bool le = f.row_kind[i] == PDLP_ROW_LE;
    const double b = ldv(f.b64, shb, tile, f.m, i, lane);
    double ra = da - b, rc = dc - b;
    if (le && ra < 0.0) ra = 0.0;
    if (le && rc < 0.0) rc = 0.0;
    a[0] += ra * ra;
    a[NSC] += rc * rc;
    if (le && ya > 0.0) a[1] += ya * ya;
    if (le && yc > 0.0) a[NSC + 1] += yc * yc;
    const double bs = ldv(f.b_shift, shs, tile, f.m, i, lane);
    a[3] += bs * ya;
    a[NSC + 3] += bs * yc;
    a[6] += bs * ya;
    a[8] += fabs(bs * ya);
    a[NSC + 6] += bs * yc;
    a[NSC + 8] += fabs(bs * yc);
    if (adaptive) {
        const double yl = (double)f.y_last[iy], rm = f.row_metric[i];
        const double da2 = ya - yl, dc2 = yc - yl;
        a[5] += da2 * da2 * rm;
        a[NSC + 5] += dc2 * dc2 * rm;
    }
}

// K x of both candidates over entries [beg, end) of a row (x window sums
// already folded by the column pass).
__device__ __forceinline__ void row_products(const PdlpFamily& f, int tile, int lane, int win, double wd, int beg,
                                             int end, double& da, double& dc) {
    da = 0.0;
    dc = 0.0;
    // x and x_sum loaded whatever the window, four entries' loads together
    int q = beg;
    for (; q + 4 <= end; q += 4) {
        double kv[4], xc[4], xs[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            kv[u] = (double)f.val[q + u];
            const size_t jx = at(tile, f.n, f.col_idx[q + u], lane);
            xc[u] = (double)f.x[jx];
            xs[u] = f.x_sum[jx];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) keep_loaded(xs[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double xa = win > 0 ? xs[u] / wd : xc[u];
            da += kv[u] * xa;
            dc += kv[u] * xc[u];
        }
    }
    for (; q < end; ++q) {
        const double kv = (double)f.val[q];
        const size_t jx = at(tile, f.n, f.col_idx[q], lane);
        const double xc = (double)f.x[jx];
        const double xs = f.x_sum[jx];
        keep_loaded(xs);
        const double xa = win > 0 ? xs / wd : xc;
        da += kv * xa;
        dc += kv * xc;
    }
}

// Rows of the check.  The first nl chunks each take one long row (the last nl
// rows) on the whole workgroup, its entries split over the wavefronts; the
// other chunks stride over the short rows.
__global__ void __launch_bounds__(NT) kkt_rows_kernel(PdlpFamily f, int add, int P, int nl) {
    const Place pl = place(f, P);
    if (!pl.ok) return;
    const int tile = pl.tile;
    const uint64_t mask = f.run_mask[tile];
    if (mask == 0) return;
    const int lane = lane_id(), w = wave_id();
    const bool act = (mask >> lane) & 1ull;
    const int k = tile * T + lane;
    const int win = act ? f.lp[k].window + add : 0;   // window length including this check's iterations
    const double wd = (double)win;
    const bool adaptive = f.flags & 1;
    double a[NS2];
#pragma unroll
    for (int s = 0; s < NS2; ++s) a[s] = 0.0;
    if (pl.chunk < nl) {                               // uniform over the workgroup
        __shared__ double sp[2][NWV][T];
        const int i = f.m - nl + pl.chunk;
        const int beg = f.row_ptr[i], end = f.row_ptr[i + 1];
        const int per = (end - beg + NWV - 1) / NWV;
        const int b0 = min(beg + w * per, end), e0 = min(b0 + per, end);
        double da, dc;
        row_products(f, tile, lane, win, wd, b0, e0, da, dc);
        sp[0][w][lane] = da;
        sp[1][w][lane] = dc;
        __syncthreads();
        if (w == 0) {
            da = 0.0;
            dc = 0.0;
#pragma unroll
            for (int v = 0; v < NWV; ++v) {
                da += sp[0][v][lane];
                dc += sp[1][v][lane];
            }
            row_terms(f, tile, lane, act, win, wd, adaptive, i, da, dc, a);
        }
    } else {
        const int rows = f.m - nl;
        for (int p = (pl.chunk - nl) * NWV + w; p < rows; p += (P - nl) * NWV) {
            const int i = rows - 1 - p;                  // longer rows first when nl = 0
            double da, dc;
            row_products(f, tile, lane, win, wd, f.row_ptr[i], f.row_ptr[i + 1], da, dc);
            row_terms(f, tile, lane, act, win, wd, adaptive, i, da, dc, a);
        }
    }
    store_chunk(f, KCH + pl.chunk, tile, a);
}

struct Kkt {
    double rp, rd, gap, err, po, dobj, dx, dy;
    bool infeasible;     // the candidate's y is a Farkas certificate of primal infeasibility
};

// F(y) = b_shift'y - max_box (K'y)'(x - lo_shift) > 0 proves that no x in the box
// satisfies the rows (y <= 0 on <= rows holds by projection).  The margin 1e-6
// of the summed magnitudes exceeds the FP32 rounding of K's values (6e-8) and
// the FP64 rounding of the sums, so a feasible LP is never flagged.
__device__ __forceinline__ bool farkas(const double* s) {
    const double F = s[6] - s[7], scale = s[8] + s[9];
    return isfinite(F) && F > 0.0 && F > 1e-6 * scale;
}

__device__ __forceinline__ Kkt kkt_of(const double* s, double norm_b, double norm_c) {
    Kkt k;
    k.rp = sqrt(s[0]) / norm_b;
    k.rd = sqrt(s[1]) / norm_c;
    k.po = s[2];
    k.dobj = s[3];
    k.gap = fabs(k.po - k.dobj) / (1.0 + fabs(k.po) + fabs(k.dobj));
    k.err = fmax(k.rp, fmax(k.rd, k.gap));
    if (!isfinite(k.po) || !isfinite(k.dobj)) k.err = NAN;
    k.dx = sqrt(s[4]);
    k.dy = sqrt(s[5]);
    k.infeasible = farkas(s);
    return k;
}

__device__ __forceinline__ void write_out(PdlpBatchOut* o, const PdlpFamilyLP& s, const Kkt& kk, int iters) {
    PdlpBatchOut r;
    r.status = s.status;
    r.iterations = iters;
    r.restarts = s.restarts;
    r.checks = s.checks;
    r.rel_kkt = kk.err;
    r.rp = kk.rp;
    r.rd = kk.rd;
    r.gap = kk.gap;
    r.p_obj = kk.po;
    r.d_obj = kk.dobj;
    r.weight = (double)s.weight;
    *o = r;
}

// One workgroup of NSC wavefronts per tile: wavefront q adds up sums q and
// q + NSC over the chunks (lane = LP), then wavefront 0 takes one decision per
// LP.  mode 0: start (reference error of the starting point); mode 1: a check
// after check_every iterations.
__global__ void __launch_bounds__(NSC * T) decide_kernel(PdlpFamily f, int mode, int col_chunks, int row_chunks) {
    __shared__ double red[NS2][T];
    const int tile = blockIdx.x, lane = lane_id(), w = wave_id();
    const int k = tile * T + lane;
    const uint64_t mask = f.run_mask[tile];
    const bool act = (mask >> lane) & 1ull;
    for (int q = w; q < NS2; q += NSC) {
        const size_t cpad = (size_t)f.tiles * T;
        double t = 0.0;
        if (act) {
            // 32 chunks' loads in flight at a time, added in chunk order
#pragma unroll 32
            for (int c = 0; c < col_chunks; ++c) t += f.partials[((size_t)c * NS2 + q) * cpad + k];
#pragma unroll 32
            for (int c = 0; c < row_chunks; ++c) t += f.partials[((size_t)(KCH + c) * NS2 + q) * cpad + k];
        }
        red[q][lane] = t;
    }
    __syncthreads();
    if (w != 0) return;                                   // no barrier below
    bool running = false, stop = false;
    if (k < f.count) {
        PdlpFamilyLP s = f.lp[k];
        if (!act) {
            s.action = ACT_NONE;
        } else {
            double sum[NS2];
#pragma unroll
            for (int v = 0; v < NS2; ++v) sum[v] = red[v][lane];
            const bool adaptive = f.flags & 1, fixed_weight = f.flags & 2;
            Kkt ka = kkt_of(sum, s.norm_b, s.norm_c);
            const Kkt kc = kkt_of(sum + NSC, s.norm_b, s.norm_c);
            s.action = ACT_NONE;
            if (mode == 0) {
                Kkt k0 = kc;
                bool cold = false;
                if (s.warm && f.failover_iters > 0) {
                    // the average slot held the cold start (init_vec_kernel): keep the warm start
                    // unless the cold point is better (a NaN KKT loses; a tie, e.g. a start that
                    // only sets the primal weight, keeps it)
                    s.start_err = kc.err;
                    s.cold_err = ka.err;
                    cold = (s.warm & 2) && !(kc.err <= ka.err);
                    if (cold) {
                        k0 = ka;
                        s.warm = 0;
                        s.weight = s.weight_cold;
                        s.weight0 = s.weight_cold;
                    }
                    s.window = 0;
                    s.action = cold ? ACT_COLD : ACT_RESTART_CUR;   // clear the sums either way
                }
                s.ref_err = k0.err;
                s.err_prev = INFINITY;
                s.status = k0.infeasible ? PDLP_INFEASIBLE : isfinite(k0.err) ? PDLP_RUNNING : PDLP_NUMERICAL;
                if (s.status != PDLP_RUNNING) s.action = cold ? ACT_COLD : ACT_DONE_CUR;
                write_out(&f.out[k], s, k0, 0);
            } else {
                const int it = s.iter + f.check_every;
                s.iter = it;
                s.window += f.check_every;
                const bool inf_avg = ka.infeasible, inf_cur = kc.infeasible;
                bool use_avg = true;
                if (adaptive && (!isfinite(ka.err) || kc.err < ka.err)) {   // the better candidate
                    use_avg = false;
                    ka = kc;
                }
                ++s.checks;
                if (inf_cur || inf_avg) {                                    // certificate found
                    s.status = PDLP_INFEASIBLE;
                    s.done_iter = it;
                    s.w_div = (double)s.window;
                    s.action = inf_cur ? ACT_DONE_CUR : ACT_DONE_AVG;
                } else if (!isfinite(ka.err)) {
                    s.status = PDLP_NUMERICAL;
                    s.done_iter = it;
                    s.action = ACT_DONE_CUR;
                } else if (ka.err < f.tol) {                                 // pdhg.py:231-235
                    s.status = PDLP_OPTIMAL;
                    s.done_iter = it;
                    s.w_div = (double)s.window;
                    s.action = use_avg ? ACT_DONE_AVG : ACT_DONE_CUR;
                } else {
                    bool restart;
                    if (!adaptive) {
                        restart = ka.err < 0.5 * s.ref_err;                  // pdhg.py:238-244
                    } else {
                        restart = ka.err <= 0.2 * s.ref_err || (ka.err <= 0.8 * s.ref_err && ka.err > s.err_prev) ||
                                  (double)(it - s.last_restart) >= 0.36 * (double)(it - s.iter0);
                    }
                    if (restart) {
                        if (adaptive && !fixed_weight && ka.dx > 1e-10 && ka.dy > 1e-10) {
                            const double lw = 0.5 * log(ka.dx / ka.dy) + 0.5 * log((double)s.weight);
                            s.weight = (float)clamp_weight(exp(lw), (double)s.weight0);
                        }
                        s.ref_err = ka.err;
                        s.err_prev = INFINITY;
                        s.last_restart = it;
                        ++s.restarts;
                        s.w_div = (double)s.window;
                        s.window = 0;
                        s.action = use_avg ? ACT_RESTART_AVG : ACT_RESTART_CUR;
                    } else {
                        s.err_prev = ka.err;
                    }
                    // warm-start failover, judged once at the first check past failover_iters: a
                    // start that has made no progress and is still far out (KKT at least a tenth of
                    // the cold start's; one stalled near the tolerance is kept) solves cold from here
                    if (s.warm && f.failover_iters > 0 && it >= f.failover_iters) {
                        if (!(ka.err < s.start_err) && !(ka.err < 0.1 * s.cold_err)) {
                            s.weight = s.weight_cold;
                            s.weight0 = s.weight_cold;
                            s.ref_err = s.cold_err;
                            s.err_prev = INFINITY;
                            s.last_restart = it;
                            s.iter0 = it;
                            ++s.restarts;
                            s.window = 0;
                            s.action = ACT_COLD;
                        }
                        s.warm = 0;
                    }
                }
                if (s.status == PDLP_RUNNING && f.max_iters > 0 && it >= f.max_iters) {   // the LP's own limit
                    stop = true;
                    s.done_iter = it;
                    s.action = s.action == ACT_RESTART_AVG   ? ACT_RESTART_AVG_STOP
                               : s.action == ACT_RESTART_CUR ? ACT_RESTART_CUR_STOP
                               : s.action == ACT_COLD        ? ACT_COLD_STOP
                                                             : ACT_DONE_CUR;
                }
                write_out(&f.out[k], s, ka, s.status == PDLP_RUNNING ? it : s.done_iter);
            }
            running = s.status == PDLP_RUNNING && !stop;   // a stopped LP keeps PDLP_RUNNING
        }
        f.lp[k] = s;
    }
    const uint64_t run = __ballot(running);
    if (lane == 0) f.run_mask[tile] = run;
}

// Vector side of the decisions: restart to the average or the current point
// (clear the window sums, move the restart point), or store the solution.
__global__ void __launch_bounds__(NT) apply_kernel(PdlpFamily f) {
    const int tile = blockIdx.y;
    const int lane = lane_id(), w = wave_id();
    const int k = tile * T + lane;
    const int act = k < f.count ? f.lp[k].action : (int)ACT_NONE;
    if (__ballot(act != ACT_NONE) == 0) return;           // no barrier in this kernel
    const double wd = act != ACT_NONE ? f.lp[k].w_div : 1.0;
    const bool move_ref = (f.flags & 4) == 0;
    const int per = 16;
    const int d0 = (blockIdx.x * NWV + w) * per;
    const int d1 = min(d0 + per, f.n + f.m);
    for (int d = d0; d < d1; ++d) {
        const bool col = d < f.n;
        const size_t ix = col ? at(tile, f.n, d, lane) : at(tile, f.m, d - f.n, lane);
        float* v = col ? f.x : f.y;
        double* sum = col ? f.x_sum : f.y_sum;
        float* last = col ? f.x_last : f.y_last;
        switch (act) {
            case ACT_RESTART_AVG: {
                const float nv = (float)(sum[ix] / wd);
                v[ix] = nv;
                sum[ix] = 0.0;
                if (move_ref) last[ix] = nv;
                break;
            }
            case ACT_RESTART_CUR:
                sum[ix] = 0.0;
                if (move_ref) last[ix] = v[ix];
                break;
            case ACT_DONE_AVG:
                sum[ix] = sum[ix] / wd;
                break;
            case ACT_DONE_CUR:
                sum[ix] = (double)v[ix];
                break;
            case ACT_COLD: {
                const float nv = col ? cold_x(f, tile, d, lane) : 0.f;
                v[ix] = nv;
                last[ix] = nv;
                // a start that ends the solve at once (status set at init) keeps it as solution
                sum[ix] = f.lp[k].status == PDLP_RUNNING ? 0.0 : (double)nv;
                break;
            }
            case ACT_RESTART_AVG_STOP: {
                const float nv = (float)(sum[ix] / wd);
                v[ix] = nv;
                if (move_ref) last[ix] = nv;
                sum[ix] = (double)nv;
                break;
            }
            case ACT_RESTART_CUR_STOP:
                if (move_ref) last[ix] = v[ix];
                sum[ix] = (double)v[ix];
                break;
            case ACT_COLD_STOP: {
                const float nv = col ? cold_x(f, tile, d, lane) : 0.f;
                v[ix] = nv;
                last[ix] = nv;
                sum[ix] = (double)nv;
                break;
            }
            default:
                break;
        }
    }
}

__global__ void __launch_bounds__(NT) count_kernel(PdlpFamily f) {
    int c = 0;
    for (int t = threadIdx.x; t < f.tiles; t += NT) c += __popcll(f.run_mask[t]);
    c = group_sum<kWave>(c);
    __shared__ int red[NWV];
    if (lane_id() == 0) red[wave_id()] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int v = 0; v < NWV; ++v) t += red[v];
        *f.running = t;
    }
}

// ---------------------------------------------------------------- set-up / end
__global__ void __launch_bounds__(NT) init_vec_kernel(PdlpFamily f) {
    const int tile = blockIdx.y;
    const int lane = lane_id(), w = wave_id();
    const int per = 16;
    const int d0 = (blockIdx.x * NWV + w) * per;
    const int d1 = min(d0 + per, f.n + f.m);
    // LPs with a warm start and the failover on hold the cold start in their sums with a
    // window of 1 (init_lp_kernel), so the initial check evaluates it as the second candidate
    const int k = tile * T + lane;
    const bool judge = f.failover_iters > 0 && k < f.count && f.lp[k].warm;
    for (int d = d0; d < d1; ++d) {
        if (d < f.n) {
            const size_t ix = at(tile, f.n, d, lane);
            f.x_sum[ix] = judge ? (double)cold_x(f, tile, d, lane) : 0.0;
            f.x_psum[ix] = 0.f;
            f.x_last[ix] = f.x[ix];
            f.xbar[ix] = f.x[ix];
        } else {
            const size_t ix = at(tile, f.m, d - f.n, lane);
            f.y_sum[ix] = 0.0;
            f.y_psum[ix] = 0.f;
            f.y_last[ix] = f.y[ix];
        }
    }
}

__global__ void __launch_bounds__(T) init_lp_kernel(PdlpFamily f) {
    const int tile = blockIdx.x, lane = threadIdx.x;
    const int k = tile * T + lane;
    const bool valid = k < f.count;
    if (valid) {
        PdlpFamilyLP s = f.lp[k];
        s.status = PDLP_RUNNING;
        s.iter = s.window = s.restarts = s.checks = s.done_iter = s.last_restart = s.iter0 = 0;
        if (f.failover_iters > 0 && s.warm) s.window = 1;   // the cold start in the sums (init_vec_kernel)
        s.action = ACT_NONE;
        s.ref_err = INFINITY;
        s.err_prev = INFINITY;
        s.w_div = 1.0;
        f.lp[k] = s;
    }
    const uint64_t run = __ballot(valid);
    if (lane == 0) f.run_mask[tile] = run;
}

// Records of freshly loaded LPs from their norms (pdlp_family_records).
__global__ void __launch_bounds__(T) records_kernel(PdlpFamily f, const double* __restrict__ norms, int warm) {
    const int k = blockIdx.x * T + threadIdx.x;
    if (k >= f.count) return;
    PdlpFamilyLP s = {};
    s.norm_b = norms[3 * (size_t)k];
    s.norm_c = norms[3 * (size_t)k + 1];
    const float w = (f.flags & 1) ? (float)norms[3 * (size_t)k + 2] : 1.f;
    s.weight = s.weight0 = s.weight_cold = w;
    s.warm = warm;
    f.lp[k] = s;
}

// LPs still running at the end: iteration limit, current iterate as solution.
__global__ void __launch_bounds__(T) finish_lp_kernel(PdlpFamily f, int tail) {
    const int tile = blockIdx.x, lane = threadIdx.x;
    const int k = tile * T + lane;
    const bool act = (f.run_mask[tile] >> lane) & 1ull;
    if (k < f.count) {
        PdlpFamilyLP s = f.lp[k];
        if (act) {
            s.iter += tail;
            s.done_iter = s.iter;
            s.action = ACT_DONE_CUR;
            f.out[k].iterations = s.iter;
            f.out[k].weight = (double)s.weight;
        } else {
            s.action = ACT_NONE;
        }
        f.lp[k] = s;
    }
}

// Starting point from a row-major FP32 matrix (a network output, one row per
// LP): a 64 LP x 64 index block is read row by row (coalesced along the
// row), transposed through LDS and written to the tiled layout, with the
// affine de-normalisation and the projections on the way.
__global__ void __launch_bounds__(NT) set_start_kernel(PdlpFamily f, const float* __restrict__ src, int64_t ld,
                                                       int x_off, int y_off, const float* __restrict__ scale,
                                                       const float* __restrict__ shift) {
    __shared__ float blk[T][T + 1];
    const int tile = blockIdx.y, lane = lane_id(), w = wave_id();
    const int d0 = blockIdx.x * T, D = f.n + f.m;
    for (int r = w; r < T; r += NWV) {
        const int k = tile * T + r, d = d0 + lane;
        float v = 0.f;
        if (k < f.count && d < D) {
            const int col = d < f.n ? (x_off >= 0 ? x_off + f.col_orig[d] : -1)
                                    : (y_off >= 0 ? y_off + f.row_orig[d - f.n] : -1);
            if (col >= 0) {
                v = src[(int64_t)k * ld + col];
                if (scale) v *= scale[col];
                if (shift) v += shift[col];
                if (!isfinite(v)) v = 0.f;
            }
        }
        blk[r][lane] = v;
    }
    __syncthreads();
    const bool shl = f.shared & PDLP_FAM_SHARED_LO, shh = f.shared & PDLP_FAM_SHARED_HI;
    for (int dd = w; dd < T; dd += NWV) {
        const int d = d0 + dd;
        if (d >= D) break;
        float v = blk[lane][dd];
        if (d < f.n) {
            const float lo = ldv(f.lo, shl, tile, f.n, d, lane), hi = ldv(f.hi, shh, tile, f.n, d, lane);
            v = v < lo ? lo : v;
            v = v > hi ? hi : v;
            f.x[at(tile, f.n, d, lane)] = v;
        } else {
            const int i = d - f.n;
            if (f.row_kind[i] == PDLP_ROW_LE && v > 0.f) v = 0.f;
            f.y[at(tile, f.m, i, lane)] = v;
        }
    }
}

__global__ void __launch_bounds__(T) set_weight_kernel(PdlpFamily f, const float* __restrict__ src, int64_t ld,
                                                       int w_col, const float* __restrict__ scale,
                                                       const float* __restrict__ shift) {
    const int k = blockIdx.x * T + threadIdx.x;
    if (k >= f.count) return;
    float v = src[(int64_t)k * ld + w_col];
    if (scale) v *= scale[w_col];
    if (shift) v += shift[w_col];
    if (isfinite(v)) {
        const float w = (float)exp(fmin(fmax((double)v, -8.0), 8.0));
        f.lp[k].weight = w;
        f.lp[k].weight0 = w;
    }
}

// ------------------------------------------------------------ problem data
// Per-LP vectors from LP-major FP64 sources in the problems' own order (one row
// per LP), written tiled in device order.  Lane = LP; each wavefront takes
// LOAD_PER consecutive device indices, so consecutive loads of a lane fall in
// the same source cache lines.  Columns (which = 0): c -> c, c64; lo, hi -> lo,
// hi, lo_shift, ub_span.  Rows (which = 1): b -> b, b64.
constexpr int LOAD_PER = 32;

__global__ void __launch_bounds__(NT) load_vec_kernel(PdlpFamily f, int which, const double* __restrict__ sc,
                                                      const double* __restrict__ slo, const double* __restrict__ shi,
                                                      const double* __restrict__ sb) {
    const int tile = blockIdx.y, lane = lane_id(), w = wave_id();
    const int k = tile * T + lane;
    if (k >= f.count) return;                                // no barrier in this kernel
    const int L = which == 0 ? f.n : f.m;
    const int d0 = (blockIdx.x * NWV + w) * LOAD_PER, d1 = min(d0 + LOAD_PER, L);
    for (int d = d0; d < d1; ++d) {
        const size_t ix = at(tile, L, d, lane);
        if (which == 0) {
            const size_t src = (size_t)k * f.n + f.col_orig[d];
            if (sc) {
                const double c = sc[src];
                const_cast<float*>(f.c)[ix] = (float)c;
                const_cast<double*>(f.c64)[ix] = c;
            }
            if (slo) {
                const double lo = slo[src], hi = shi[src];
                const double ls = isfinite(lo) ? lo : 0.0;
                const_cast<float*>(f.lo)[ix] = (float)lo;
                const_cast<float*>(f.hi)[ix] = (float)hi;
                const_cast<double*>(f.lo_shift)[ix] = ls;
                const_cast<double*>(f.ub_span)[ix] = isfinite(hi) ? hi - ls : INFINITY;
            }
        } else {
            const double b = sb[(size_t)k * f.m + f.row_orig[d]];
            const_cast<float*>(f.b)[ix] = (float)b;
            const_cast<double*>(f.b64)[ix] = b;
        }
    }
}

// b_shift = b - K lo_shift (FP64), one row per wavefront, lane = LP.
__global__ void __launch_bounds__(NT) load_shift_kernel(PdlpFamily f) {
    const int tile = blockIdx.y, lane = lane_id(), w = wave_id();
    const int i = blockIdx.x * NWV + w;
    const int k = tile * T + lane;
    if (i >= f.m || k >= f.count) return;
    const bool shl = f.shared & PDLP_FAM_SHARED_LOSHIFT, shb = f.shared & PDLP_FAM_SHARED_B64;
    double a = 0.0;
    for (int q = f.row_ptr[i]; q < f.row_ptr[i + 1]; ++q)
        a += (double)f.val[q] * ldv(f.lo_shift, shl, tile, f.n, f.col_idx[q], lane);
    const_cast<double*>(f.b_shift)[at(tile, f.m, i, lane)] = ldv(f.b64, shb, tile, f.m, i, lane) - a;
}

// Per LP: norm_b = max(1, ||(b_shift, finite ub_span)||), norm_c = max(1, ||c||) (the
// reference-form normalisers, pdhg.py:181-182 on the lifted form) and the cold
// weight ||sqrt(sigma_hat) b|| / ||sqrt(tau_hat) c|| clamped to [1e-6, 1e6] (1 when a
// norm vanishes).  One workgroup of NWN wavefronts per tile: wavefront w sums the
// indices w, w + NWN, ...; the partial sums meet in a fixed order.
constexpr int NWN = 16;

__global__ void __launch_bounds__(NWN * T) load_norm_kernel(PdlpFamily f, double* __restrict__ norms) {
    __shared__ double red[NWN][4][T];
    const int tile = blockIdx.x, lane = lane_id(), w = wave_id();
    const int k = tile * T + lane;
    const bool shc = f.shared & PDLP_FAM_SHARED_C64, shu = f.shared & PDLP_FAM_SHARED_UBSPAN,
               shb = f.shared & PDLP_FAM_SHARED_B64, shs = f.shared & PDLP_FAM_SHARED_BSHIFT;
    double bs2 = 0.0, c2 = 0.0, bw2 = 0.0, cw2 = 0.0;
    // the wavefront's indices w, w + NWN, ... in order, four indices' loads together
    constexpr int U = 4;
    for (int d0 = w; d0 < f.n; d0 += U * NWN) {
        double c[U], sp[U], th[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + u * NWN < f.n ? d0 + u * NWN : d0;
            c[u] = ldv(f.c64, shc, tile, f.n, d, lane);
            sp[u] = ldv(f.ub_span, shu, tile, f.n, d, lane);
            th[u] = (double)f.tau_hat[d];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (d0 + u * NWN >= f.n) break;
            if (isfinite(sp[u])) bs2 += sp[u] * sp[u];
            c2 += c[u] * c[u];
            cw2 += c[u] * c[u] * th[u];
        }
    }
    for (int i0 = w; i0 < f.m; i0 += U * NWN) {
        double b[U], bs[U], rm[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * NWN < f.m ? i0 + u * NWN : i0;
            b[u] = ldv(f.b64, shb, tile, f.m, i, lane);
            bs[u] = ldv(f.b_shift, shs, tile, f.m, i, lane);
            rm[u] = f.row_metric[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i0 + u * NWN >= f.m) break;
            bs2 += bs[u] * bs[u];
            bw2 += b[u] * b[u] / rm[u];
        }
    }
    red[w][0][lane] = bs2;
    red[w][1][lane] = c2;
    red[w][2][lane] = bw2;
    red[w][3][lane] = cw2;
    __syncthreads();
    if (w != 0 || k >= f.count) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < NWN; ++v)
        for (int q = 0; q < 4; ++q) s[q] += red[v][q][lane];
    const double nb = sqrt(s[2]), nc = sqrt(s[3]);
    norms[3 * (size_t)k + 0] = fmax(1.0, sqrt(s[0]));
    norms[3 * (size_t)k + 1] = fmax(1.0, sqrt(s[1]));
    norms[3 * (size_t)k + 2] = (nb > 1e-12 && nc > 1e-12) ? fmin(fmax(nb / nc, 1e-6), 1e6) : 1.0;
}

__global__ void __launch_bounds__(T) clear_mask_kernel(PdlpFamily f) {
    if (threadIdx.x == 0) f.run_mask[blockIdx.x] = 0;
}

// --------------------------------------------------------------------- host
int check_family(const PdlpFamily* f) {
    if (f == nullptr) return fail(PDLP_E_ARG, "PdlpFamily is NULL");
    if (f->m <= 0 || f->n <= 0 || f->count <= 0) return fail(PDLP_E_SHAPE, "family needs m, n, count > 0");
    if (f->tiles != (f->count + T - 1) / T) return fail(PDLP_E_SHAPE, "tiles must be ceil(count / 64)");
    if (f->tiles > 65535) return fail(PDLP_E_LIMIT, "at most 65535 tiles of 64 LPs");
    if (f->check_every <= 0) return fail(PDLP_E_ARG, "check_every must be > 0");
    if (f->max_iters < 0) return fail(PDLP_E_ARG, "max_iters must be >= 0");
    if (f->items_per_wave < 1 || f->items_per_wave > 64) return fail(PDLP_E_TUNE, "items_per_wave in [1, 64]");
    if (f->long_rows < 0 || f->long_rows > f->m || f->long_cols < 0 || f->long_cols > f->n)
        return fail(PDLP_E_SHAPE, "long_rows in [0, m] and long_cols in [0, n]");
    if (!f->row_orig || !f->col_orig) return fail(PDLP_E_ARG, "row_orig / col_orig is NULL");
    if (f->xcd_affinity != 0 && f->xcd_affinity != 1) return fail(PDLP_E_TUNE, "xcd_affinity must be 0 or 1");
    if (!f->row_ptr || !f->col_idx || !f->val || !f->t_row_ptr || !f->t_col_idx || !f->t_val || !f->row_kind ||
        !f->tau_hat || !f->sigma_hat || !f->row_metric || !f->c || !f->lo || !f->hi || !f->b || !f->c64 ||
        !f->lo_shift || !f->ub_span || !f->b64 || !f->b_shift || !f->x || !f->xbar || !f->x_psum || !f->x_last ||
        !f->y || !f->y_psum || !f->y_last || !f->x_sum || !f->y_sum || !f->lp || !f->out || !f->run_mask ||
        !f->running || !f->partials)
        return fail(PDLP_E_ARG, "a required PdlpFamily pointer is NULL");
    return 0;
}

int chunks(int items) { return std::max(1, std::min(KCH, (items + NWV - 1) / NWV)); }

// Grid for P chunks per tile (see place()).
dim3 tile_grid(const PdlpFamily& f, int P) {
    if (!f.xcd_affinity) return dim3((unsigned)P, (unsigned)f.tiles);
    const int T = f.tiles;
    const int per_xcd = T >= 8 ? ((T + 7) / 8) * P : (P + (8 / T) - 1) / (8 / T);
    return dim3((unsigned)(8 * per_xcd), 1u);
}

// Chunks per tile of a pass over nl long items and `items` short ones, nw
// wavefronts per workgroup.
int pass_chunks(const PdlpFamily& f, int items, int nl, int nw) {
    const int per_block = nw * f.items_per_wave;
    return nl + (items + per_block - 1) / per_block;
}

dim3 vec_grid(const PdlpFamily& f) {
    const int per_block = NWV * 16;
    return dim3((unsigned)((f.n + f.m + per_block - 1) / per_block), (unsigned)f.tiles);
}

// With long items the pass runs on NL-wavefront workgroups, else on NT threads.
void launch_column_pass(const PdlpFamily& f, hipStream_t s) {
    const int nl = f.long_cols, sc = f.n - nl;
    if (nl) {
        const int P = pass_chunks(f, sc, nl, NL);
        primal_kernel<NL * kWave><<<tile_grid(f, P), NL * kWave, 0, s>>>(f, P);
    } else if (sc) {
        const int P = pass_chunks(f, sc, 0, NWV);
        primal_kernel<NT><<<tile_grid(f, P), NT, 0, s>>>(f, P);
    }
}

void launch_row_pass(const PdlpFamily& f, hipStream_t s) {
    const int nl = f.long_rows, sr = f.m - nl;
    if (nl) {
        const int P = pass_chunks(f, sr, nl, NL);
        dual_kernel<NL * kWave><<<tile_grid(f, P), NL * kWave, 0, s>>>(f, P);
    } else if (sr) {
        const int P = pass_chunks(f, sr, 0, NWV);
        dual_kernel<NT><<<tile_grid(f, P), NT, 0, s>>>(f, P);
    }
}

void launch_iterations(const PdlpFamily& f, int iters, hipStream_t s) {
    for (int it = 0; it < iters; ++it) {
        launch_column_pass(f, s);
        launch_row_pass(f, s);
    }
}

void launch_check(const PdlpFamily& f, int mode, hipStream_t s) {
    // long rows get a workgroup each while they take at most half the chunks
    const int nl = f.long_rows <= KCH / 2 ? f.long_rows : 0;
    const int cc = chunks(f.n), rc = nl + std::min(chunks(f.m - nl), KCH - nl);
    const int add = mode == 0 ? 0 : f.check_every;
    kkt_cols_kernel<<<tile_grid(f, cc), NT, 0, s>>>(f, add, cc);
    kkt_rows_kernel<<<tile_grid(f, rc), NT, 0, s>>>(f, add, rc, nl);
    decide_kernel<<<f.tiles, NSC * T, 0, s>>>(f, mode, cc, rc);
    apply_kernel<<<vec_grid(f), NT, 0, s>>>(f);
    count_kernel<<<1, NT, 0, s>>>(f);
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" int pdlp_family_init(const PdlpFamily* f, void* stream) {
    if (int e = check_family(f)) return e;
    hipStream_t s = as_stream(stream);
    init_lp_kernel<<<f->tiles, T, 0, s>>>(*f);
    init_vec_kernel<<<vec_grid(*f), NT, 0, s>>>(*f);
    launch_check(*f, 0, s);
    return hip_fail(hipGetLastError(), "pdlp_family_init launch");
}

extern "C" int pdlp_family_run(const PdlpFamily* f, int32_t blocks, void* stream) {
    if (int e = check_family(f)) return e;
    if (blocks < 0) return fail(PDLP_E_ARG, "blocks must be >= 0");
    hipStream_t s = as_stream(stream);
    for (int b = 0; b < blocks; ++b) {
        launch_iterations(*f, f->check_every, s);
        launch_check(*f, 1, s);
    }
    return hip_fail(hipGetLastError(), "pdlp_family_run launch");
}

extern "C" int pdlp_family_iterate(const PdlpFamily* f, int32_t iters, void* stream) {
    if (int e = check_family(f)) return e;
    if (iters < 0) return fail(PDLP_E_ARG, "iters must be >= 0");
    launch_iterations(*f, iters, as_stream(stream));
    return hip_fail(hipGetLastError(), "pdlp_family_iterate launch");
}

extern "C" int pdlp_family_set_start(const PdlpFamily* f, const float* src, int64_t ld, int32_t x_off,
                                     int32_t y_off, int32_t w_col, const float* scale, const float* shift,
                                     void* stream) {
    if (int e = check_family(f)) return e;
    if (src == nullptr && (x_off >= 0 || y_off >= 0 || w_col >= 0))
        return fail(PDLP_E_ARG, "src is NULL but a column offset is given");
    const int64_t need = std::max<int64_t>({(int64_t)x_off + f->n, (int64_t)y_off + f->m, (int64_t)w_col + 1});
    if (src != nullptr && (ld <= 0 || need > ld))
        return fail(PDLP_E_SHAPE, "src columns x_off+n, y_off+m, w_col must lie within ld");
    hipStream_t s = as_stream(stream);
    const dim3 g((unsigned)((f->n + f->m + T - 1) / T), (unsigned)f->tiles);
    set_start_kernel<<<g, NT, 0, s>>>(*f, src, ld, x_off, y_off, scale, shift);
    if (w_col >= 0) set_weight_kernel<<<f->tiles, T, 0, s>>>(*f, src, ld, w_col, scale, shift);
    return hip_fail(hipGetLastError(), "pdlp_family_set_start launch");
}

extern "C" int pdlp_family_load(const PdlpFamily* f, const double* c, const double* b, const double* lo,
                                const double* hi, double* norms, void* stream) {
    if (int e = check_family(f)) return e;
    if (norms == nullptr) return fail(PDLP_E_ARG, "norms is NULL");
    if ((lo == nullptr) != (hi == nullptr)) return fail(PDLP_E_ARG, "lo and hi are both given or both NULL");
    const int sh = f->shared;
    auto both = [&](int a, int b2) { return (sh & a) && (sh & b2); };
    auto none = [&](int a, int b2) { return !(sh & a) && !(sh & b2); };
    if (c ? !none(PDLP_FAM_SHARED_C, PDLP_FAM_SHARED_C64) : !both(PDLP_FAM_SHARED_C, PDLP_FAM_SHARED_C64))
        return fail(PDLP_E_ARG, "c is given exactly when c / c64 are not shared");
    if (b ? !none(PDLP_FAM_SHARED_B, PDLP_FAM_SHARED_B64) : !both(PDLP_FAM_SHARED_B, PDLP_FAM_SHARED_B64))
        return fail(PDLP_E_ARG, "b is given exactly when b / b64 are not shared");
    const int lo_bits = PDLP_FAM_SHARED_LO | PDLP_FAM_SHARED_HI | PDLP_FAM_SHARED_LOSHIFT | PDLP_FAM_SHARED_UBSPAN;
    if (lo ? (sh & lo_bits) != 0 : (sh & lo_bits) != lo_bits)
        return fail(PDLP_E_ARG, "lo / hi are given exactly when lo, hi, lo_shift and ub_span are not shared");
    const bool bshift_shared = (sh & PDLP_FAM_SHARED_BSHIFT) != 0;
    if (bshift_shared != (b == nullptr && lo == nullptr))
        return fail(PDLP_E_ARG, "b_shift is shared exactly when b and lo are");
    hipStream_t s = as_stream(stream);
    const int per_block = NWV * LOAD_PER;
    if (c || lo)
        load_vec_kernel<<<dim3((unsigned)((f->n + per_block - 1) / per_block), (unsigned)f->tiles), NT, 0, s>>>(
            *f, 0, c, lo, hi, nullptr);
    if (b)
        load_vec_kernel<<<dim3((unsigned)((f->m + per_block - 1) / per_block), (unsigned)f->tiles), NT, 0, s>>>(
            *f, 1, nullptr, nullptr, nullptr, b);
    if (!bshift_shared)
        load_shift_kernel<<<dim3((unsigned)((f->m + NWV - 1) / NWV), (unsigned)f->tiles), NT, 0, s>>>(*f);
    load_norm_kernel<<<f->tiles, NWN * T, 0, s>>>(*f, norms);
    return hip_fail(hipGetLastError(), "pdlp_family_load launch");
}

extern "C" int pdlp_family_records(const PdlpFamily* f, const double* norms, int32_t warm, void* stream) {
    if (int e = check_family(f)) return e;
    if (norms == nullptr) return fail(PDLP_E_ARG, "norms is NULL");
    if (warm < 0 || warm > 3) return fail(PDLP_E_ARG, "warm must be 0 .. 3");
    records_kernel<<<f->tiles, T, 0, as_stream(stream)>>>(*f, norms, warm);
    return hip_fail(hipGetLastError(), "records_kernel launch");
}

extern "C" int pdlp_family_finish(const PdlpFamily* f, int32_t tail, void* stream) {
    if (int e = check_family(f)) return e;
    if (tail < 0) return fail(PDLP_E_ARG, "tail must be >= 0");
    hipStream_t s = as_stream(stream);
    finish_lp_kernel<<<f->tiles, T, 0, s>>>(*f, tail);
    apply_kernel<<<vec_grid(*f), NT, 0, s>>>(*f);
    clear_mask_kernel<<<f->tiles, T, 0, s>>>(*f);
    count_kernel<<<1, NT, 0, s>>>(*f);
    return hip_fail(hipGetLastError(), "pdlp_family_finish launch");
}

namespace {

// Capture `reps` launches of `body` on `stream` into an instantiated graph.
template <typename F>
int capture(hipStream_t s, void** exec, F body) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    body();
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e, "family graph capture");
    }
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *exec = (void*)ge;
    return 0;
}

}  // namespace

extern "C" int pdlp_family_graph_capture(const PdlpFamily* f, int32_t blocks, void* stream, void** exec) {
    if (int e = check_family(f)) return e;
    if (blocks <= 0 || exec == nullptr) return fail(PDLP_E_ARG, "blocks must be positive and exec non-NULL");
    if (stream == nullptr) return fail(PDLP_E_ARG, "graph capture needs a non-default stream");
    hipStream_t s = as_stream(stream);
    return capture(s, exec, [&] {
        for (int b = 0; b < blocks; ++b) {
            launch_iterations(*f, f->check_every, s);
            launch_check(*f, 1, s);
        }
    });
}

extern "C" int pdlp_family_graph_capture_step(const PdlpFamily* f, int32_t which, int32_t reps, void* stream,
                                              void** exec) {
    if (int e = check_family(f)) return e;
    if (reps <= 0 || exec == nullptr || (which != 0 && which != 1))
        return fail(PDLP_E_ARG, "reps > 0, which in {0, 1} and exec non-NULL");
    if (stream == nullptr) return fail(PDLP_E_ARG, "graph capture needs a non-default stream");
    hipStream_t s = as_stream(stream);
    return capture(s, exec, [&] {
        for (int r = 0; r < reps; ++r) {
            if (which == 0) launch_column_pass(*f, s);
            else launch_row_pass(*f, s);
        }
    });
}
