/*
 * pdlp_hip.h — C ABI of libpdlp_hip.so, the MI355X (gfx950) restarted-PDHG engine.
 *
 * The reference has no native layer: its PDHG hot path is the Python loop of
 * solver/lp/pdhg.py:209-244 (CPU, SciPy csr_matvec + NumPy) and
 * solver/lp/pdhg_gpu.py:196-282, 378-427 (library kernels dispatched one by
 * one from Python, KKT on the host every 50 iterations).  Each entry point
 * below names the reference code it replaces.  The Python drop-in
 * (sovereign_solver_amd.lp.pdhg / pdhg_gpu / batch) binds these with ctypes.
 *
 * Conventions
 *   - every pointer is a device pointer owned by the caller; the library
 *     allocates no device memory (workspaces are passed in);
 *   - every call is asynchronous on the given stream (hipStream_t passed as
 *     void*; NULL = default stream);
 *   - return value: 0 on success, a positive hipError_t from the HIP runtime,
 *     or a negative PDLP_E* code for argument errors; pdlp_last_error()
 *     returns a thread-local message for the last failure;
 *   - no C++ exception crosses the ABI.
 */
#ifndef PDLP_HIP_H
#define PDLP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDLP_ABI_VERSION 15

/* precision of iterates, matrix values and problem vectors */
#define PDLP_F32 0
#define PDLP_F64 1

/* negative return codes */
#define PDLP_E_ARG    (-1)   /* bad argument (null pointer, size, enum) */
#define PDLP_E_TUNE   (-2)   /* unsupported tuning value */
#define PDLP_E_SHAPE  (-3)   /* inconsistent sizes */
#define PDLP_E_LIMIT  (-4)   /* size beyond a documented maximum */

/* run status kept in PdlpCtrl.status */
#define PDLP_RUNNING    0
#define PDLP_OPTIMAL    1
#define PDLP_NUMERICAL  3    /* NaN/Inf in a KKT evaluation */
#define PDLP_INFEASIBLE 4    /* primal infeasible (single-instance box form and family
                                engine): the dual iterate is a Farkas certificate,
                                b'y - max_{lo<=x<=hi} (K'y)'x > 0 with y <= 0 on
                                PDLP_ROW_LE rows */
#define PDLP_SKIPPED    6    /* persistent batch engine: the LP needs more LDS than its
                                part's launch has (descriptors inconsistent with
                                lds_need); not run */
#define PDLP_UNBOUNDED  5    /* dual infeasible (single-instance box form, adaptive
                                schemes): d = x - x_sum / apply_window is a ray with
                                c'd < 0 whose row and bound violations pass the
                                engine's relative tests, among them |viol(K d)| +
                                |viol_bounds(d)| <= 1e-6 (-c'd); the LP is unbounded
                                when it is feasible */

/* row kinds of the box form */
#define PDLP_ROW_EQ 0        /* K_i x  = b_i, dual free     */
#define PDLP_ROW_LE 1        /* K_i x <= b_i, dual y_i <= 0 */

/* restart schemes (PdlpTune.scheme, PdlpBatchLP.flags bit 0) */
#define PDLP_SCHEME_REFERENCE 0  /* pdhg.py:225-244: window average, restart when
                                    KKT < 0.5 x KKT at the last restart, fixed steps */
#define PDLP_SCHEME_ADAPTIVE  1  /* candidate = better of average / current iterate;
                                    restart on sufficient (0.2) or necessary (0.8)
                                    decay without progress, or when the window
                                    exceeds 0.36 x all iterations; at each restart
                                    the primal weight moves (geometric mean) toward
                                    the ratio |dx| / |dy| of the distances moved
                                    since the last restart, in the step-size metric */
#define PDLP_SCHEME_ADAPTIVE_FIXED 2  /* adaptive restarts, primal weight kept fixed */
#define PDLP_SCHEME_ADAPTIVE_START 3  /* adaptive restarts, weight toward |x - x0| / |y - y0|
                                         with (x0, y0) the starting point */

/* Number of doubles of PdlpDev.partials the engine needs:
 * 2 candidates x 2 passes x PDLP_PARTIAL_BLOCKS blocks x 16 sums. */
#define PDLP_PARTIAL_BLOCKS 1024
#define PDLP_PARTIALS_LEN   (4 * PDLP_PARTIAL_BLOCKS * 16)

/*
 * Device control block, one per instance, device-resident.  The engine
 * updates it; the host reads it back (a few dozen bytes) between blocks of
 * iterations.  Replaces the host-side scalars of pdhg.py:198-207
 * (window_len, current_kkt, status, iters).
 */
typedef struct PdlpCtrl {
    int32_t iter;            /* iterations executed                          */
    int32_t window;          /* iterations in the current averaging window   */
    int32_t status;          /* PDLP_RUNNING / PDLP_OPTIMAL / PDLP_NUMERICAL / PDLP_INFEASIBLE /
                                PDLP_UNBOUNDED */
    int32_t restarts;
    int32_t done_iter;       /* iteration at which status left RUNNING       */
    int32_t restart_pending; /* set by the check, consumed by the apply step */
    int32_t apply_window;    /* window length the pending average divides by */
    int32_t checks;
    double ref_err;          /* KKT at the last restart point (pdhg.py:203,244) */
    double last_err, rp, rd, gap;
    double tol;
    double norm_b, norm_c;   /* max(1,||b||), max(1,||c||) of the reference form */
    double tau_scale;        /* multiplies tau_j (primal weight w); 1 = reference */
    double sigma_scale;      /* multiplies sigma_i (1/w); 1 = reference          */
    double p_obj, d_obj;     /* objectives of the last check (reference form)   */
    /* adaptive scheme */
    double err_prev;         /* candidate KKT at the previous check (+inf after a restart) */
    int32_t cand;            /* candidate of the pending restart: 0 current, 1 average */
    int32_t last_restart;    /* iteration of the last restart                   */
    double weight0;          /* starting primal weight (tau_scale at pdlp_init); the adaptive
                                weight stays within [w0/1e2, w0*1e4]            */
} PdlpCtrl;

/*
 * One LP on the device.  Box form: min c'x, K_i x (= | <=) b_i, lo <= x <= hi.
 * The reference form of pdhg.py:62-125 is the special case with every row
 * '=', lo = 0, hi = +inf; the same kernels run both.
 * Arrays of type "prec" are float (PDLP_F32) or double (PDLP_F64).
 */
typedef struct PdlpDev {
    int32_t m, n, nnz;
    int32_t prec;
    /* K (m x n) CSR and K^T (n x m) CSR; column indices sorted per row */
    const int32_t *row_ptr, *col_idx; const void *val;
    const int32_t *t_row_ptr, *t_col_idx; const void *t_val;
    const void *c;            /* n, prec                                   */
    const void *lo, *hi;      /* n, prec; NULL means 0 and +inf            */
    const void *b;            /* m, prec                                   */
    const int8_t *row_kind;   /* m; NULL means every row PDLP_ROW_EQ       */
    /* FP64 data of the KKT yardstick (reference formula on the lifted form) */
    const double *c64;        /* n                                          */
    const double *b64;        /* m                                          */
    const double *b_shift;    /* m: b - K*lo_shift (right-hand side after the shift) */
    const double *lo_shift;   /* n: finite lower bound, else 0; NULL means 0 */
    const double *ub_span;    /* n: hi - lo_shift where hi is finite, else +inf; NULL means +inf */
    /* state */
    void *x, *xbar, *y;       /* n, n, m; prec */
    double *x_sum, *y_sum;    /* n, m: FP64 ergodic sums of the window */
    float  *x_psum, *y_psum;  /* n, m: PDLP_F32 only — FP32 sums of the iterations since
                                 the last check, folded into x_sum/y_sum at every check;
                                 NULL = add into x_sum/y_sum every iteration (PDLP_F64) */
    double *partials;         /* PDLP_PARTIALS_LEN doubles of workspace */
    PdlpCtrl *ctrl;
    void *x_last, *y_last;    /* n, m; prec: last restart point (adaptive scheme; may be
                                 NULL for the reference scheme) */
    /* optional SELL-64 copy of some rows of K for the dual step (PdlpTune.sell): slice s
       holds 64 rows, one per lane, its entries column-major (entry k of lane l at
       sell_ptr[s] + 64 k + l), so rows whose k-th columns are neighbours (the demand
       rows of a transportation LP) gather xbar in whole cache lines; the other rows
       stay with the CSR kernel through csr_rows.  Summation order = CSR order. */
    int32_t sell_slices;      /* 0: no SELL rows                                    */
    int32_t n_csr_rows;       /* entries of csr_rows                                */
    const int32_t *sell_ptr;  /* sell_slices + 1                                    */
    const int32_t *sell_rows; /* 64 * sell_slices: row of each lane, -1 = padding    */
    const int32_t *sell_col;  /* sell_ptr[sell_slices]: columns (padding 0)          */
    const void *sell_val;     /* same length, prec (padding 0)                      */
    const int32_t *csr_rows;  /* the rows outside the slices                        */
    int32_t sell_max_width;   /* largest slice width (entries per lane)              */
    int32_t k_const;          /* 0, or the value (an integer) every entry of K holds, when the
                                 SELL-64 copies (rows and K^T) hold no padding inside a row:
                                 the SELL primal and dual steps then read no values (the
                                 arrays still hold them; a transportation LP: 1) */
    void *sell_part;          /* prec, 64 * sell_slices * ceil(sell_max_width / 8): partial
                                 dot products of the split SELL step (PdlpTune.sell >= 8) */
    const void *sell_asum;    /* prec, 64 * sell_slices: sum_j |K_ij| of each lane's row */
    /* Diagonal scaling of a preconditioned LP: the kernels iterate on K = D_r K0 D_c
       (b = D_r b0, c = D_c c0, bounds / D_c) while the KKT yardstick is evaluated on the
       original LP: residual rows (K x - b)_i / D_r[i], duals D_r[i] y_i, reduced costs
       r_j / D_c[j] (objectives, b_shift . y and ub_span . min(r, 0) are invariant).
       The distances of the restart criteria stay in the scaled space.  NULL = unscaled. */
    const double *row_scale;  /* m: D_r */
    const double *col_scale;  /* n: D_c */
    /* optional SELL-64 copy of K^T for the primal step (PdlpTune.sell_primal): slice s
       holds columns 64 s .. 64 s + 63 in their natural order, one per lane, its entries
       column-major (entry k of lane l at t_sell_ptr[s] + 64 k + l; padding value 0,
       index 0), so no column pointer is read and every load of an entry is a whole
       line; summed in CSR order, it gives the one-lane CSR kernel's result
       (built by pdlp_csr_to_sell64 from the K^T CSR) */
    int32_t t_sell_slices;    /* ceil(n / 64), or 0: no SELL copy                   */
    int32_t t_sell_reserved;
    const int32_t *t_sell_ptr;  /* t_sell_slices + 1                                 */
    const int32_t *t_sell_idx;  /* t_sell_ptr[t_sell_slices]: rows of K (padding 0)   */
    const void *t_sell_val;     /* same length, prec (padding 0)                      */
    /* optional index runs for the one-launch SELL dual step (PdlpTune.sell 2..7): a slice
       whose rows all have its full width and whose every row's column indices step by one
       constant, col_k = col_0 + k * stride, has that stride in sell_stride (else 0), and
       a row of csr_rows whose indices step so has it in csr_stride (else 0); the step
       then reads only each row's first index.  NULL: every index is read. */
    const int32_t *sell_stride;   /* sell_slices                                     */
    const int32_t *csr_stride;    /* n_csr_rows, in the order of csr_rows            */
} PdlpDev;

/* Launch shape, chosen by the autotuner (sovereign_solver_amd.engine.tune). */
typedef struct PdlpTune {
    int32_t tpr_primal;   /* lanes per K^T row: 1,2,4,...,64, or 128/256/512/1024
                             = one row per workgroup of that many threads */
    int32_t tpr_dual;     /* lanes per K row, same choices */
    int32_t block;        /* threads per workgroup for tpr <= 64: multiple of 64 in [64, 1024] */
    int32_t check_every;  /* iterations between KKT checks; 50 = reference */
    int32_t xcd_remap;    /* bit 0: primal step, bit 1: dual step — give each XCD
                             a contiguous range of rows */
    int32_t exact;        /* 1: sequential row sums, no FMA contraction (bitwise
                             agreement with SciPy/NumPy in PDLP_F64); needs tpr 1 */
    int32_t block_dual;   /* threads per workgroup of the dual step for tpr_dual <= 64;
                             0 = same as block */
    int32_t scheme;       /* PDLP_SCHEME_* restart scheme */
    int32_t sell;         /* dual step on the SELL-64 slices of PdlpDev (if any): 0 off;
                             1 one wavefront per slice; 2..7 slices and the CSR rows in
                             one launch, a workgroup per slice piece (threads / rows of
                             a piece: 2: 1024/64, 3: 512/64, 4: 1024/16, 5: 1024/8,
                             6: 1024/32, 7: 512/16), the piece's entries split among
                             its wavefronts and their sums added in a fixed order, and
                             tpr_dual threads per CSR row (at least 64, at most the
                             workgroup); k >= 8 (multiple of 8) a wavefront per k
                             entries of a slice, partial sums added in order */
    int32_t no_ray;       /* 1: skip the unboundedness (ray of decrease) test of the
                             adaptive checks; 0: test at every check */
    int32_t sell_primal;  /* primal step over the SELL-64 copy of K^T (PdlpDev.t_sell_*), one
                             lane per column, `block` threads per workgroup: 1 one slice of
                             64 columns per wavefront, 2 two (a lane's two columns issued
                             together; same sums); 0: CSR */
} PdlpTune;

int         pdlp_abi_version(void);
const char* pdlp_last_error(void);

/* stride[k] = the column step s of CSR row rows[k] when its columns are col_0 + j s
 * for every entry j (1 for a single entry), else 0 (no entries, or steps that differ):
 * the PdlpDev.csr_stride of the listed rows.  No reference counterpart. */
int pdlp_row_strides(int32_t count, const int32_t* rows, const int32_t* row_ptr, const int32_t* col_idx,
                     int32_t* stride, void* stream);

/* width[s] = the longest row of rows 64 s .. 64 s + 63 of a CSR matrix (rows past
 * the end count 0), for s < ceil(rows / 64): the slice widths of its SELL-64 copy
 * (the host's sell_ptr is 64 times their running sum, as in pdlp_csr_to_sell64).
 * The reference builds no such copy (pdhg_gpu.py:137,161,186 keep CSR): no
 * counterpart. */
int pdlp_sell64_widths(int32_t rows, const int32_t* row_ptr, int32_t* width, void* stream);

/* SELL-64 copy of a CSR matrix's rows (the K^T of PdlpDev.t_sell_*): slice s holds
 * rows 64 s .. 64 s + 63, its width the longest of them; entry k of row 64 s + l goes
 * to sell_ptr[s] + 64 k + l.  sell_ptr (slices + 1, slices = ceil(rows / 64)) is the
 * caller's: sell_ptr[s + 1] - sell_ptr[s] = 64 * width.  sell_idx / sell_val must be
 * zero-filled beforehand (the padding).  No reference counterpart (the reference
 * keeps CSR; pdhg_gpu.py:137,161). */
int pdlp_csr_to_sell64(int32_t prec, int32_t rows, const int32_t* row_ptr, const int32_t* col_idx,
                       const void* val, const int32_t* sell_ptr, int32_t* sell_idx, void* sell_val, void* stream);

/* K^T of an arena of `mats` CSR matrices stored one after another (the persistent
 * batch's layout, PdlpBatchLP.rp_off / nnz_off / cp_off): matrix s has rows
 * row_off[s] .. row_off[s+1]-1 and columns col_off[s] .. col_off[s+1]-1 of the arena's
 * block-diagonal union (rows, cols in total), its entries at nz_off[s] .. nz_off[s+1]-1
 * (nnz in total) with 16-bit column indices local to the matrix, and its m + 1 row
 * pointers at row_ptr[rp_off[s]..] relative to its first entry.  Writes each matrix's
 * n + 1 column pointers at t_col_ptr[cp_off[s]..] (relative), and its K^T entries
 * (16-bit local row indices, values) at the same entry range — per matrix bitwise
 * SciPy's canonical transpose (the host restatement is one transpose of the union,
 * engine/batch.py).  All offset arrays have mats + 1 entries.  `work` holds at least
 * pdlp_arena_transpose_work_bytes(...) bytes.  Replaces the per-LP host transposes
 * of pdhg.py:179 for a batch; the reference has no batch: no direct counterpart. */
int pdlp_arena_transpose_work_bytes(int32_t mats, int32_t rows, int32_t cols, int32_t nnz, size_t* bytes);
int pdlp_arena_transpose(int32_t prec, int32_t mats, int32_t rows, int32_t cols, int32_t nnz,
                         const int32_t* row_off, const int32_t* col_off, const int32_t* nz_off,
                         const int32_t* rp_off, const int32_t* cp_off, const int32_t* row_ptr,
                         const uint16_t* col_idx, const void* val, int32_t* t_col_ptr, uint16_t* t_row_idx,
                         void* t_val, void* work, size_t work_bytes, void* stream);

/* Device CSR transpose for problem ingestion: K^T (cols x rows) from K
 * (rows x cols, CSR with ascending column indices per row), values of `prec`.
 * Replaces the host transposes `AT = A.T.tocsr()` of pdhg.py:179 and
 * pdhg_gpu.py:137,161,186.  The result equals SciPy's canonical transpose bit
 * for bit: rows of K^T in column order, entries in ascending row order.
 * t_row_ptr: cols + 1 entries; t_col_idx, t_val: nnz.  `work` holds at least
 * pdlp_csr_transpose_work_bytes(rows, cols, nnz) bytes of device memory (may
 * be NULL when nnz = 0).  Column indices must lie in [0, cols). */
int pdlp_csr_transpose_work_bytes(int32_t rows, int32_t cols, int32_t nnz, size_t* bytes);
int pdlp_csr_transpose(int32_t prec, int32_t rows, int32_t cols, int32_t nnz, const int32_t* row_ptr,
                       const int32_t* col_idx, const void* val, int32_t* t_row_ptr, int32_t* t_col_idx,
                       void* t_val, void* work, size_t work_bytes, void* stream);

/* y = A x, A in CSR (prec values), x,y prec.  The SpMV of pdhg.py:211,217
 * (SciPy csr_matvec) and pdhg_gpu.py:215,222,245,252 (cuSPARSE / torch.mv). */
int pdlp_spmv_csr(int32_t prec, const int32_t* row_ptr, const int32_t* col_idx,
                  const void* val, const void* x, void* y, int32_t nrows,
                  int32_t tpr, void* stream);

/* y = A x with A in SELL-64 slices of its rows in natural order (the layout of
 * pdlp_csr_to_sell64: slice s = rows 64 s .. 64 s + 63, entry k of row 64 s + l at
 * sell_ptr[s] + 64 k + l, padding 0): one lane per row, no row pointer read,
 * each row summed in CSR order, so y equals pdlp_spmv_csr's with one lane per
 * row.  For matrices of many short rows of similar length (K^T of a
 * transportation LP: 2 entries per row).  Replaces the same products as
 * pdlp_spmv_csr (pdhg.py:211; pdhg_gpu.py:215,245). */
int pdlp_spmv_sell64(int32_t prec, const int32_t* sell_ptr, const int32_t* sell_idx, const void* sell_val,
                     const void* x, void* y, int32_t nrows, void* stream);

/* y = A x with FP32 A, FP64 x and FP64 accumulation: the products of the KKT
 * check (pdhg.py:186,188) when the iterate engine runs in FP32. */
int pdlp_spmv_csr_f64acc(const int32_t* row_ptr, const int32_t* col_idx,
                         const float* val, const double* x, double* y,
                         int32_t nrows, int32_t tpr, void* stream);

/* Reset the state for a solve from (x, y) already in dev->x / dev->y
 * (zeros = the reference's cold start, pdhg.py:175-176): clears the sums and
 * counters and evaluates the KKT of the starting point into ctrl->ref_err
 * (pdhg.py:203).  ctrl->tol, norm_b, norm_c, tau_scale, sigma_scale must be
 * set by the caller beforehand. */
int pdlp_init(const PdlpDev* dev, const PdlpTune* tune, void* stream);

/* One primal step (pdhg.py:211; pdhg_gpu.py:215-216):
 *   x+ = clip(x - tau.(c - K^T y), lo, hi),  xbar = 2x+ - x,  x_sum += x+
 * tau_j = tau_scale * 0.999 / max(sum_i |K_ij|, 1e-4) is formed from the K^T
 * row in registers (pdhg.py:131-147), never stored. */
int pdlp_primal_step(const PdlpDev* dev, const PdlpTune* tune, void* stream);

/* One dual step (pdhg.py:217; pdhg_gpu.py:222-223):
 *   y+ = proj(y + sigma.(b - K xbar)),  y_sum += y+
 * proj keeps y_i <= 0 on PDLP_ROW_LE rows. */
int pdlp_dual_step(const PdlpDev* dev, const PdlpTune* tune, void* stream);

/* KKT of the window average, termination and restart (pdhg.py:225-244,
 * replacing pdhg_gpu.py:381-427 and its host round trip). */
int pdlp_kkt_check(const PdlpDev* dev, const PdlpTune* tune, void* stream);

/* y = K x (prec of dev) over the rows of dev: its SELL-64 slices (PdlpDev.sell_*) and
 * the remaining CSR rows (csr_rows), in the one-launch layout of tune->sell (2..7)
 * with tune->tpr_dual threads per CSR row — the product K xbar of the dual step
 * (pdhg.py:217 / pdhg_gpu.py:222), the same sums, without its update.  Needs
 * sell_slices > 0 (else pdlp_spmv_csr).  x: n entries, y: m entries. */
int pdlp_spmv_dev(const PdlpDev* dev, const PdlpTune* tune, const void* x, void* y, void* stream);

/* iters x (primal, dual), then a KKT check when check != 0. */
int pdlp_run(const PdlpDev* dev, const PdlpTune* tune, int32_t iters,
             int32_t check, void* stream);

/* pdlp_run without workspace: the same iterations and check as pdlp_run, but the
 * scratch of its kernels — the check's reduction partials (dev->partials) and the
 * split SELL dual step's partial dot products (dev->sell_part) — comes from one
 * static device array of the library, PDLP_NOALLOC_SCRATCH_BYTES per GPU (the
 * partials take PDLP_PARTIALS_LEN doubles of it, the SELL partials the rest), held
 * by each GPU the library runs on until the process ends.  dev->partials and
 * dev->sell_part are ignored and may be NULL; every other pointer is the caller's, as
 * for pdlp_run.  The call allocates nothing (no hipMalloc, no library allocation): it
 * is for callers that cannot allocate device memory, such as code under HIP graph
 * capture and serving loops.  Returns PDLP_E_LIMIT when the split SELL step
 * (tune->sell >= 8) needs more than the array's remaining bytes.  All calls on one
 * GPU share its array and it keeps the last call's contents (each call writes what
 * it reads), so calls that overlap in time on one GPU race: serialise them (one
 * stream, or events).  No counterpart in the reference. */
#define PDLP_NOALLOC_SCRATCH_BYTES (4u << 20)
int pdlp_run_noalloc(const PdlpDev* dev, const PdlpTune* tune, int32_t iters, int32_t check, void* stream);

/* Capture `blocks` x pdlp_run(dev, tune, tune->check_every, 1) in a HIP graph.
 * *exec receives an opaque handle for pdlp_graph_launch / pdlp_graph_destroy. */
int pdlp_graph_capture(const PdlpDev* dev, const PdlpTune* tune, int32_t blocks,
                       void* stream, void** exec);
int pdlp_graph_launch(void* exec, void* stream);
int pdlp_graph_destroy(void* exec);

/* Capture `reps` back-to-back launches of one step kernel (which: 0 = primal,
 * 1 = dual) in a HIP graph: the timing harness of the autotuner and of the
 * roofline figure (per-kernel time without host launch gaps). */
int pdlp_graph_capture_step(const PdlpDev* dev, const PdlpTune* tune, int32_t which, int32_t reps,
                            void* stream, void** exec);

/* ----------------------------------------------------------------------------
 * Batched engine: thousands of independent box-form LPs in persistent launches.
 *
 * The reference solves LPs one at a time (solver/lp/pdhg.py:152; B&B node LPs
 * one by one at solver/milp/branch_and_bound.py:367-374); this is the batched
 * replacement of that loop.  All LPs are packed into shared arrays ("arena");
 * each PdlpBatchLP holds one LP's offsets.  Each LP is stored as CSR of K
 * and CSR of K^T with the same number of entries: row pointers relative to
 * its nnz_off, column indices relative to its col_off; column pointers (of
 * K^T) relative to its nnz_off, row indices relative to its row_off.
 *
 * One workgroup owns one LP at a time and takes the next from the work queue
 * (`order`, typically largest first) when it finishes.  Two launches on the
 * stream: the FP32 part (iterates in FP32, x in global memory, 12*n+8*m+8
 * bytes of LDS per LP) and the FP64 part (iterates in FP64, x in LDS,
 * 20*n+12*m+8 bytes), for the LPs small enough.  Per iteration a pass over
 * the columns (K^T y gathered from LDS, primal step) and a pass over the rows
 * (K xbar gathered from LDS, dual step).  Termination and restarts follow
 * pdhg.py:225-244 per LP, with the relative KKT of the reference-form lift in
 * FP64.
 * ------------------------------------------------------------------------- */

typedef struct PdlpBatchLP {
    int32_t m, n;
    int32_t tpr;            /* lanes per row: 1,2,4,8,16,32,64 */
    int32_t flags;          /* bit 0: adaptive restarts (else the reference scheme);
                               bit 1: keep the primal weight fixed;
                               bit 2: weight from the distances to the starting point
                                      (else: moved since the last restart) */
    int32_t tpc;            /* lanes per column (of K^T): 1,2,4,8,16,32,64 */
    int32_t reserved;
    int64_t row_off;        /* first row of this LP in the row arrays        */
    int64_t col_off;        /* first column of this LP in the column arrays  */
    int64_t rp_off;         /* first of its m+1 row pointers                 */
    int64_t nnz_off;        /* first of its nonzeros (in K and in K^T)       */
    int64_t cp_off;         /* first of its n+1 column pointers (of K^T)     */
    double norm_b, norm_c;  /* reference-form normalisers                     */
    double tau_scale;       /* primal weight: tau *= w, sigma /= w; NaN: the kernel takes
                               ||sqrt(sigma_hat) b|| / ||sqrt(tau_hat) c|| (1 if either is 0) */
    double sigma_scale;
} PdlpBatchLP;

typedef struct PdlpBatchOut {
    int32_t status;         /* PDLP_RUNNING (= iteration limit), PDLP_OPTIMAL, PDLP_NUMERICAL */
    int32_t iterations;
    int32_t restarts;
    int32_t checks;
    double rel_kkt, rp, rd, gap, p_obj, d_obj;
    double weight;          /* primal weight at exit (to resume a related solve) */
} PdlpBatchOut;

typedef struct PdlpBatch {
    int32_t count;          /* number of LPs                                  */
    int32_t check_every;    /* 50 = reference cadence                          */
    int32_t max_iters;
    int32_t block;          /* FP32 part: threads per workgroup, 256 or 512   */
    int32_t lds_bytes;      /* FP32 part: dynamic LDS per workgroup, >= 12*n+8*m+8 of its LPs
                               (checked against lds_need; an LP whose own need still exceeds
                               it is not run and reports PDLP_SKIPPED) */
    int32_t grid;           /* FP32 part: persistent workgroups (0 = 2 per CU) */
    int32_t n_resident;     /* the last n_resident LPs of `order` run with FP64 iterates */
    int32_t block_resident; /* FP64 part: threads per workgroup, 256 or 512 (0 = 256) */
    int32_t lds_resident;   /* FP64 part: dynamic LDS, >= 20*n+12*m+8 of its LPs */
    int32_t grid_resident;  /* FP64 part: persistent workgroups (0 = 2 per CU) */
    int32_t parts;          /* parts to launch: bit 0 FP32 part, bit 1 FP64 part (0 = both) */
    int32_t x_in_lds;       /* FP32 part: x, its partial sum, tau_hat and c in LDS as well
                               (then lds_bytes >= 28*n+8*m+8 of its LPs) */
    int32_t lds_need;       /* FP32 part: the largest LDS need of its LPs (12*n+8*m+8, or
                               28*n+8*m+8 with x_in_lds); the call returns PDLP_E_LIMIT
                               when it exceeds lds_bytes */
    int32_t lds_need_resident; /* FP64 part: the largest 20*n+12*m+8 of its LPs; PDLP_E_LIMIT
                               when it exceeds lds_resident */
    int32_t lines;          /* FP32 part: lines (rows or columns) per lane group and round,
                               whose loads are in flight together: 2, 3 or 4 (0 = 3) */
    double tol;
    const PdlpBatchLP* lps; /* count descriptors                              */
    const int32_t* order;   /* count LP indices, processing order: FP32 part, then FP64 part */
    int32_t* queue;         /* two int32 work counters (FP32, FP64 part), the launched parts'
                               zero before the launch */
    /* matrices: K by rows, K^T by rows (= K by columns); 16-bit indices within
       the LP (every LP that fits in LDS has m, n < 32768) */
    const int32_t* row_ptr; const uint16_t* col_idx; const float* val;
    const int32_t* t_col_ptr; const uint16_t* t_row_idx; const float* t_val;
    /* per column: c, bounds (+-inf allowed), FP64 c, lower-bound shift, (hi - shift) or +inf */
    const float *c, *lo, *hi;
    const double *c64, *lo_shift, *ub_span;
    /* per row: b, kind, FP64 b, FP64 b - K*lo_shift */
    const float* b; const int8_t* row_kind; const double *b64, *b_shift;
    /* state: x and y hold the warm start on entry (zeros = cold start);
       on exit x_sum / y_sum hold each LP's solution in FP64 */
    float* x; float* x_psum; double* x_sum; float* tau;   /* per column */
    float* y; double* y_sum;                               /* per row    */
    PdlpBatchOut* out;
    float* x_last; float* y_last;   /* per column / row: last restart point (scratch) */
    float* y_psum;                  /* per row: sum of y since the last check (scratch) */
    /* Diagonal scaling of the stored LPs (K = D_r K0 D_c, b, c and bounds scaled to
       match): the checks evaluate the KKT yardstick of the original LPs, as PdlpDev's
       row_scale / col_scale do.  NULL = unscaled. */
    const double* row_scale;        /* per row: D_r */
    const double* col_scale;        /* per column: D_c */
} PdlpBatch;

/* Solve every LP of the batch: one persistent launch per part.  With both
 * parts the FP64 part runs concurrently on a second stream of the library's
 * (default priority, one per device and caller stream, kept until the process
 * ends), forked from and joined back into `stream` through events, so work
 * enqueued on `stream` afterwards sees both parts finished — also when the call
 * returns an error after the fork.  Returns PDLP_E_LIMIT when a part's lds_need
 * exceeds its LDS (the reference raises on bad input, pdhg_gpu.py:92). */
int pdlp_batch_solve(const PdlpBatch* batch, void* stream);

/* ----------------------------------------------------------------------------
 * Shared-matrix batch ("family"): B box-form LPs with one constraint matrix K
 * (m x n, row kinds shared) and their own c, b, lo, hi — parametric families
 * (varying costs and demands) and branch-and-bound nodes (varying bounds),
 * the node loop of solver/milp/branch_and_bound.py:367-374 and the per-instance
 * solve_lp_pdhg calls of solver/lp/pdhg.py:152.
 *
 * All LPs step in lockstep.  Per-LP vectors live in tiles of 64 LPs, LP index
 * fastest: element d of LP k in a vector of length L is at
 *     ((k / 64) * L + d) * 64 + k % 64
 * so one wavefront updates one row (or column) of 64 LPs with coalesced
 * loads, and the CSR of K is read with uniform (scalar) loads.  A vector
 * marked shared (PDLP_FAM_SHARED_* bit) is stored once, length L.
 *
 * Per check (every check_every iterations) the window sums are folded, the
 * reference KKT (pdhg.py:184-196 on the lifted form, FP64) of the window
 * average and of the current iterate is evaluated for every running LP, and
 * each LP takes its own termination / restart decision (pdhg.py:225-244, or
 * the adaptive scheme as in PdlpBatchLP.flags).  Finished LPs are frozen and
 * skipped by tiles; their FP64 solution is in x_sum / y_sum.
 * ------------------------------------------------------------------------- */

#define PDLP_FAM_TILE 64
#define PDLP_FAM_SHARED_C 1
#define PDLP_FAM_SHARED_B 2
#define PDLP_FAM_SHARED_LO 4
#define PDLP_FAM_SHARED_HI 8
#define PDLP_FAM_SHARED_C64 16
#define PDLP_FAM_SHARED_B64 32
#define PDLP_FAM_SHARED_BSHIFT 64
#define PDLP_FAM_SHARED_LOSHIFT 128
#define PDLP_FAM_SHARED_UBSPAN 256
#define PDLP_FAM_KKT_CHUNKS 128    /* partial-sum chunks per KKT pass (max)   */
#define PDLP_FAM_LONG 64            /* default split: rows / columns with more nonzeros run on a
                                       whole workgroup (the host numbers them last)          */
#define PDLP_FAM_NSUM 20            /* 10 sums x (window average, current)     */

typedef struct PdlpFamilyLP {
    int32_t status, iter, window, restarts, checks, done_iter, last_restart, action;
    double norm_b, norm_c;  /* reference-form normalisers                       */
    double ref_err, err_prev, w_div;
    float weight;           /* primal weight: tau = w tau_hat, sigma = sigma_hat / w */
    float weight0;          /* starting weight; the adaptive weight stays within [w0/1e2, w0*1e4] */
    float weight_cold;      /* weight of a cold start, taken when the warm start is dropped */
    int32_t warm;           /* host: bit 0 = the LP starts from a warm start (x, y or weight set by
                               the caller or pdlp_family_set_start), bit 1 = the start sets y, so
                               its KKT (whose gap term is near 1 for any x with y = 0) is compared
                               with the cold start's at init; the engine clears it once judged */
    int32_t iter0;          /* iteration of the last cold restart by the failover (0 otherwise): the
                               adaptive scheme's window-length rule counts from it */
    int32_t reserved;
    double start_err;       /* KKT of the warm start (set at pdlp_family_init) */
    double cold_err;        /* KKT of the cold start clip(0, lo, hi), y = 0 (set at init for warm LPs) */
} PdlpFamilyLP;

typedef struct PdlpFamily {
    int32_t m, n, count;    /* rows, columns, LPs                              */
    int32_t tiles;          /* ceil(count / 64)                                */
    int32_t check_every;
    int32_t flags;          /* as PdlpBatchLP.flags, for every LP              */
    int32_t shared;         /* PDLP_FAM_SHARED_* bits                          */
    int32_t items_per_wave; /* rows / columns per wavefront in a step, 1..64   */
    int32_t long_rows;      /* the last long_rows rows run on one workgroup each (the host
                               numbers the rows with more than a threshold of
                               nonzeros, PDLP_FAM_LONG by default, last)          */
    int32_t long_cols;      /* likewise for the last long_cols columns           */
    int32_t xcd_affinity;   /* 1: pin each tile's workgroups to its own XCD(s), so a tile's
                               gathered vectors stay in one L2 (workgroups are dispatched
                               round-robin over the 8 XCDs); 0: 2-D grid (chunk, tile) */
    int32_t failover_iters; /* warm-start failover (0 = off; the reference's
                               WARMSTART_FAILOVER_ITERS = 200, solver/config.py:75): for each LP
                               with lp.warm set, pdlp_family_init evaluates the KKT of the warm
                               start and of the cold start, and drops a start that sets y (warm
                               bit 1) when the cold start is better (or the warm KKT is NaN); at
                               the first check at or after
                               failover_iters iterations, a warm-started LP whose KKT has not
                               fallen below its starting value and is at least a tenth of the
                               cold start's restarts from the cold start (weight_cold, sums
                               cleared; iter0 = that iteration) */
    int32_t max_iters;      /* 0: none (the caller stops the run, pdlp_family_finish); > 0: an LP
                               still running at a check at or past max_iters of its own
                               iterations stops there with PDLP_RUNNING and its iterate after
                               that check's action as solution, as pdlp_family_finish with tail
                               0 would leave it (for streams whose LPs start at different checks;
                               a multiple of check_every gives the batch run's limit exactly) */
    int32_t reserved;
    double tol;
    /* K and K^T, CSR, FP32 values */
    const int32_t* row_ptr; const int32_t* col_idx; const float* val;
    const int32_t* t_row_ptr; const int32_t* t_col_idx; const float* t_val;
    const int8_t* row_kind;       /* m, PDLP_ROW_EQ / PDLP_ROW_LE            */
    const float* tau_hat;         /* n: 0.999 / max(sum_i |K_ij|, 1e-4)      */
    const float* sigma_hat;       /* m: 0.999 / max(sum_j |K_ij|, 1e-4)      */
    const double* row_metric;     /* m: max(sum_j |K_ij|, 1e-4) / 0.999      */
    const int32_t* row_orig;      /* m: caller's index of each row (pdlp_family_set_start) */
    const int32_t* col_orig;      /* n: caller's index of each column                     */
    /* problem data, tiled per LP or shared */
    const float *c, *lo, *hi, *b;
    const double *c64, *lo_shift, *ub_span, *b64, *b_shift;
    /* state, tiled; x and y hold the start on entry to pdlp_family_init */
    float *x, *xbar, *x_psum, *x_last;
    float *y, *y_psum, *y_last;
    double *x_sum, *y_sum;
    PdlpFamilyLP* lp;             /* count                                    */
    PdlpBatchOut* out;            /* count                                    */
    uint64_t* run_mask;           /* tiles: running LPs of the tile           */
    int32_t* running;             /* 1: number of running LPs after a check   */
    double* partials;             /* 2 * PDLP_FAM_KKT_CHUNKS * 12 * tiles * 64 */
} PdlpFamily;

/* Per-LP problem data from LP-major FP64 arrays already on the device, in the
 * problems' own row and column order: c [count][n], b [count][m] (rows
 * [A_ub; A_eq]), lo and hi [count][n] (both or neither).  A NULL source keeps
 * that vector shared: its single copy in f is used as it stands, and f->shared
 * carries the matching bits (c: C and C64; b: B and B64; lo / hi: LO, HI,
 * LOSHIFT and UBSPAN; BSHIFT exactly when b and lo are both shared).  Writes
 * the tiled per-LP vectors of f in device order (through row_orig / col_orig):
 * c, c64, b, b64, lo, hi, lo_shift = lo where finite else 0, ub_span = hi -
 * lo_shift where hi is finite else +inf, b_shift = b - K lo_shift (FP64), and
 * norms[3k .. 3k+2] = (norm_b, norm_c, cold weight) of LP k: the reference-form
 * normalisers max(1, ||(b_shift, finite ub_span)||) and max(1, ||c||)
 * (pdhg.py:181-182 on the lifted form) and ||sqrt(sigma_hat) b|| /
 * ||sqrt(tau_hat) c|| clamped to [1e-6, 1e6] (1 when a norm vanishes).  This
 * replaces the host-side construction of each LP's standard form
 * (pdhg.py:62-125) for a batch of LPs sharing K. */
int pdlp_family_load(const PdlpFamily* f, const double* c, const double* b, const double* lo, const double* hi,
                     double* norms, void* stream);
/* Set up every LP from (x, y), lp[k].norm_b/norm_c/weight as filled by the
 * host: sums cleared, restart point = start, KKT of the start. */
int pdlp_family_init(const PdlpFamily* f, void* stream);
/* The records lp[0 .. count) of LPs just loaded (pdlp_family_load's norms), on
 * the device: every field 0 except norm_b = norms[3k], norm_c = norms[3k+1],
 * weight = weight0 = weight_cold = the cold weight norms[3k+2] for the adaptive
 * schemes (flags bit 0) and 1 for the reference one, warm = `warm`.  With
 * pdlp_family_load, pdlp_family_set_start and pdlp_family_init on a view of f
 * whose per-LP pointers start at a tile t0 (tiled arrays advanced by
 * t0 * 64 * length, lp / out by t0 * 64 records, run_mask by t0 words), this
 * admits new LPs into free tiles while the LPs of the other tiles keep running
 * (a stream of LPs through one engine).  No reference counterpart. */
int pdlp_family_records(const PdlpFamily* f, const double* norms, int32_t warm, void* stream);
/* Starting points from a row-major FP32 matrix src[count][ld] (src may be
 * NULL when x_off, y_off and w_col are all negative: cold start) (e.g. the
 * warm-start network's output, already on the device): v = src[k*ld + col]
 * * scale[col] + shift[col] (scale / shift may be NULL); x_j from column
 * x_off + col_orig[j] clipped to [lo, hi], y_i from column y_off +
 * row_orig[i] (y_i <= 0 on
 * PDLP_ROW_LE rows); x_off < 0 / y_off < 0 give the cold values clip(0, lo,
 * hi) / 0; non-finite values count as 0.  If w_col >= 0 the primal weight of
 * LP k becomes exp(clamp(v, -8, 8)) of column w_col.  Call after the host has
 * written lp[] and before pdlp_family_init.  No counterpart in the
 * reference, whose PDHG always starts at 0 (pdhg.py:175-176). */
int pdlp_family_set_start(const PdlpFamily* f, const float* src, int64_t ld, int32_t x_off, int32_t y_off,
                          int32_t w_col, const float* scale, const float* shift, void* stream);
/* `blocks` times: check_every iterations of every running LP, then a check. */
int pdlp_family_run(const PdlpFamily* f, int32_t blocks, void* stream);
/* `iters` iterations without a check (the tail below check_every). */
int pdlp_family_iterate(const PdlpFamily* f, int32_t iters, void* stream);
/* End: LPs still running report PDLP_RUNNING with their current iterate as
 * the solution (x_sum / y_sum); `tail` iterations ran since the last check. */
int pdlp_family_finish(const PdlpFamily* f, int32_t tail, void* stream);
/* HIP graph of `blocks` x (check_every iterations + check), as pdlp_family_run;
 * launch with pdlp_graph_launch, free with pdlp_graph_destroy. */
int pdlp_family_graph_capture(const PdlpFamily* f, int32_t blocks, void* stream, void** exec);
/* HIP graph of `reps` column passes (which = 0) or row passes (which = 1),
 * one launch each (long and short items together), for timing. */
int pdlp_family_graph_capture_step(const PdlpFamily* f, int32_t which, int32_t reps, void* stream, void** exec);

/* ----------------------------------------------------------------------------
 * Learned warm start: a 3-layer MLP (d_in -> d_h -> d_h -> d_out, ReLU) whose
 * GEMMs run on BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.
 * pdlp_mlp_bwd_bf16 uses dout as the split-K work space of dz2 = dout w3^T
 * once the bias gradient has been taken from it.
 * The reference has no learned component (its PDHG always starts from 0,
 * pdhg.py:175-176); these kernels are new.  BF16 values are passed as
 * uint16_t bit patterns.
 * ------------------------------------------------------------------------- */

/* C[M,N] = epi(alpha * op(A)[M,K] op(B)[K,N] (+ C if accumulate))
 * op(A)(m,k) = trans_a ? A[k*lda + m] : A[m*lda + k]   (BF16)
 * op(B)(k,n) = trans_b ? B[n*ldb + k] : B[k*ldb + n]   (BF16)
 * epilogue, in order: + bias[n] (FP32, may be NULL); ReLU if relu;
 * zero where mask[m*ldc + n] <= 0 (BF16, may be NULL; the ReLU derivative);
 * stores to C (FP32, may be NULL) and/or Cb (BF16, may be NULL), both ldc.
 * Tiles of 128 x 128 per 256-thread workgroup, K in stages of 64; 16-byte
 * operand loads when lda and ldb are multiples of 8 and A, B are 16-byte
 * aligned, 2-byte loads otherwise.
 * split_k > 1 (a small M x N with a long K): the K range is cut into split_k
 * parts whose FP32 partial products go to work (split_k * M * N floats,
 * caller-owned), then summed in a fixed order (the same bits on every run)
 * before the epilogue; split_k <= 1 and work unused otherwise. */
typedef struct PdlpGemm {
    int32_t M, N, K;
    int32_t lda, ldb, ldc;
    int32_t trans_a, trans_b;
    int32_t relu, accumulate;
    float alpha;
    int32_t split_k;
    const uint16_t* A;
    const uint16_t* B;
    const float* bias;
    const uint16_t* mask;
    float* C;
    uint16_t* Cb;
    float* work;
} PdlpGemm;

int pdlp_gemm_bf16(const PdlpGemm* g, void* stream);

typedef struct PdlpMlp {
    int32_t batch, d_in, d_h, d_out;
    const uint16_t* x;                 /* batch x d_in                              */
    const uint16_t *w1, *w2, *w3;      /* BF16 weights [d_in][d_h], [d_h][d_h], [d_h][d_out] */
    const float *b1, *b2, *b3;         /* FP32 biases                                */
    uint16_t *h1, *h2;                 /* batch x d_h, post-ReLU (BF16)              */
    float* out;                        /* batch x d_out (FP32)                       */
    /* backward only */
    const float* target;               /* batch x d_out                              */
    const float* out_weight;           /* d_out loss weights, NULL = 1               */
    float* dout;                       /* batch x d_out scratch (FP32)               */
    uint16_t* dout_b;                  /* batch x d_out scratch (BF16)               */
    uint16_t *dz2, *dz1;               /* batch x d_h scratch (BF16)                 */
    float *gw1, *gw2, *gw3;            /* weight gradients (FP32, same layout as w)  */
    float *gb1, *gb2, *gb3;            /* bias gradients (FP32)                      */
    float* loss;                       /* 1 float: sum_b,j w_j (out - target)^2 / (batch*d_out) */
    float* work;                       /* backward: NULL, or work_floats FP32 of scratch for the
                                          K-split weight-gradient GEMMs (K = batch) and column sums,
                                          so a long batch with few output tiles still fills the chip */
    int64_t work_floats;
    /* backward: NULL, or W2 and W3 packed by pdlp_mlp_dz_pack (16-byte aligned); with
       d_h = 512 and d_out a multiple of 8 up to 32 the two masked products
       dz2 = (dout w3^T) * [h2 > 0] and dz1 = (dz2 w2^T) * [h1 > 0] then run as one fused
       kernel (same BF16 rounding points) */
    const uint16_t *w2b, *w3b;
} PdlpMlp;

/* h1 = relu(x w1 + b1), h2 = relu(h1 w2 + b2), out = h2 w3 + b3. */
int pdlp_mlp_fwd_bf16(const PdlpMlp* mlp, void* stream);
/* Loss of the last forward pass and all weight/bias gradients. */
int pdlp_mlp_bwd_bf16(const PdlpMlp* mlp, void* stream);

/* Fused inference of the same network for d_h = 512: the three layers of 128
 * rows per workgroup in one kernel, the hidden activations (rounded to BF16
 * at the same points as pdlp_mlp_fwd_bf16) kept in LDS and registers instead
 * of HBM.  The weights are first packed by pdlp_mlp_infer_pack (once per
 * weight update) into the kernel's MFMA operand-fragment order, so that every
 * operand load reads whole cache lines: w1p holds 512 * k1 values with k1 =
 * d_in rounded up to 16 (zeros beyond d_in), w2p 512 * 512, w3p 32 * 32 * 16
 * (zeros for channels >= d_out).
 * Replaces the layer-by-layer forward of the reference's cold start
 * (solver/lp/pdhg.py:175-176 starts from zeros; this predicts the start).
 * Limits: d_in a multiple of 8 up to 128, d_out up to 32 (PDLP_E_SHAPE). */
typedef struct PdlpMlpInfer {
    int32_t rows, d_in, d_h, d_out;
    const uint16_t* x;                 /* rows x d_in BF16, 16-byte aligned        */
    const uint16_t *w1p, *w2p, *w3p;   /* packed BF16 weights, 16-byte aligned     */
    const float *b1, *b2, *b3;         /* FP32 biases (b1, b2 16-byte aligned)     */
    float* out;                        /* rows x d_out FP32                         */
    uint16_t *h1, *h2;                 /* NULL, or rows x 512 BF16 each (16-byte aligned):
                                          the hidden activations, stored for a backward
                                          pass (pdlp_mlp_bwd_bf16 reads them)          */
} PdlpMlpInfer;
/* W2 (512 x 512) and W3 (512 x d_out) of a PdlpMlp in the fused backward kernel's
 * operand-fragment order: w2b 512 * 512 values, w3b 512 * 16 * ceil(d_out / 16). */
int pdlp_mlp_dz_pack(const uint16_t* w2, const uint16_t* w3, int32_t d_h, int32_t d_out, uint16_t* w2b,
                     uint16_t* w3b, void* stream);
int pdlp_mlp_infer_pack(const uint16_t* w1, const uint16_t* w2, const uint16_t* w3, int32_t d_in, int32_t d_h,
                        int32_t d_out, uint16_t* w1p, uint16_t* w2p, uint16_t* w3p, void* stream);
int pdlp_mlp_infer_bf16(const PdlpMlpInfer* p, void* stream);
/* Adam step on FP32 master weights w (n values) with gradient g, moments m, v;
 * refreshes the BF16 copy wb (may be NULL).  step is 1-based. */
int pdlp_adam_step(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr,
                   float beta1, float beta2, float eps, int32_t step, void* stream);
/* y = bf16(x), round to nearest even. */
int pdlp_cast_bf16(const float* x, uint16_t* y, int64_t n, void* stream);

/* ---------------------------------------------------------------- node model
 * A network applied to every node (variable) of every LP of a family with one
 * set of weights.  Its input rows are gathered from compact per-instance
 * vectors through a fixed table, and its outputs are mapped back to a start
 * [x, y, log primal weight] by a sparse assembly map.  No counterpart in the
 * reference, whose PDHG always starts at 0 (pdhg.py:175-176). */

/* Row r of the output holds node q = rows[r] (q = b * nodes + j: node j of
 * instance b); with rows NULL, q = hash(seed, r) mod (batch * nodes) when
 * seed != 0 (uniform sampling with replacement, the same rows for the same
 * seed and r), else q = r.  Entry f < width of the row is
 *   v = (g * scale[j][f] - mean[f]) * inv_std[f],
 *   g = 0 for table[j][f] < 0, src[b][t] for t = table[j][f] < src_len,
 *       shared[t - src_len] for t < src_len + shared_len, else
 *       src2[b][t - src_len - shared_len] (t < src_len + shared_len + src2_len:
 *       a second per-instance source, e.g. an earlier network's outputs);
 * entries width .. ld_out-1 are 0.  scale, mean, inv_std may be NULL (1, 0, 1).
 * Exactly one of out_bf16 / out_f32 is written. */
typedef struct PdlpNodeGather {
    int32_t batch, nodes, width, ld_out;
    int32_t src_len, shared_len;
    int64_t src2_len;
    int64_t count;                 /* output rows (rows != NULL or seed != 0), else batch * nodes */
    uint64_t seed;
    const float* src;              /* [batch][src_len]                  */
    const float* shared;           /* [shared_len] (NULL if 0)          */
    const float* src2;             /* [batch][src2_len] (NULL if 0)     */
    const int32_t* table;          /* [nodes][width]                    */
    const float* scale;            /* [nodes][width]                    */
    const float* mean;             /* [width]                           */
    const float* inv_std;          /* [width]                           */
    const int64_t* rows;           /* [count]                           */
    uint16_t* out_bf16;            /* [count][ld_out]                   */
    float* out_f32;                /* [count][ld_out]                   */
} PdlpNodeGather;

int pdlp_node_gather(const PdlpNodeGather* g, void* stream);

/* out[b][i] = bias[i] + sum_{e in indptr[i] .. indptr[i+1]-1} weight[e] *
 * net[(b * nodes + node[e]) * ld_net + chan[e]] for i < len (columns len ..
 * ld_out-1 untouched): net is the node network's FP32 output for `batch`
 * instances in order, out the start matrix read by pdlp_family_set_start. */
typedef struct PdlpNodeAssembly {
    int32_t batch, nodes, ld_net;
    int32_t len, ld_out, nnz;
    const int32_t* indptr;         /* [len + 1]   */
    const int32_t* node;           /* [nnz]       */
    const int32_t* chan;           /* [nnz]       */
    const float* weight;           /* [nnz]       */
    const float* bias;             /* [len] or NULL */
    const float* net;              /* [batch * nodes][ld_net] */
    float* out;                    /* [batch][ld_out]         */
} PdlpNodeAssembly;

int pdlp_node_assemble(const PdlpNodeAssembly* a, void* stream);

/* Network inputs of the economic-dispatch family (G generators, T periods) from
 * the LPs' own vectors, one row of ld_out floats per instance k: with c_g =
 * c[k*ld_c + g] (the first period's costs), d_t = -b[k*ld_b + t] (the demand rows'
 * right-hand sides), S = pmax_sum, the merit order = the stable ascending order of
 * c, cheaper_g = the capacity strictly before g in it (a sequential FP64 sum along
 * the order), marg_t = min(#{r : cum_r < d_t}, G - 1) (cum = the cumulated
 * capacity along the order) and price_t = the cost at position marg_t:
 *   out[off_d + t] = 4 d_t / S        out[off_c + g] = c_g / 60
 *   out[off_u + g] = c_g / cost_g - 1 out[off_cc + g] = 2 cheaper_g / S
 *   out[off_cc2 + g] = 2 (cheaper_g + pmax_g) / S
 *   out[off_p + t] = price_t / 60     out[off_m + t] = d_t / demand_t - 1
 *   out[off_x + t G + g] = clip(d_t - cheaper_g, 0, pmax_g) / pmax_g
 * each an FP64 expression rounded once to FP32 (bitwise the host definition);
 * other entries are left as they are.  G <= 1024, T <= 256.  The reference has
 * no learned component: no counterpart. */
typedef struct PdlpDispatchInputs {
    int32_t count, G, T;
    int32_t ld_c, ld_b, ld_out;
    int32_t off_d, off_c, off_u, off_cc, off_cc2, off_p, off_m, off_x;
    double pmax_sum;
    const double* c;               /* [count][ld_c] */
    const double* b;               /* [count][ld_b] */
    const double* pmax;            /* [G] */
    const double* cost;            /* [G] reference costs */
    const double* demand;          /* [T] reference demand */
    float* out;                    /* [count][ld_out] */
} PdlpDispatchInputs;

int pdlp_dispatch_inputs(const PdlpDispatchInputs* a, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PDLP_HIP_H */
