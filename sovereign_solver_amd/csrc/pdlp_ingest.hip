// Problem ingestion on the device: the CSR transpose K -> K^T that the
// single-instance engine's primal step gathers over.  The host restatement is
// SciPy's `(K.T).tocsr()` after `sort_indices()` (the canonical transpose that
// `solver/lp/pdhg.py:204-217` multiplies by through `A.T @ y`): rows of K^T in
// column order of K, entries of each in ascending row order.
//
// Method: a stable LSD radix sort (hipCUB) of the entries' column indices,
// carrying the entry ids, leaves the entries grouped by column and, within a
// column, in row order (the input is row-major and the sort is stable).  Then
// one kernel gathers row index and value per entry, and one writes the row
// pointers of K^T from the boundaries of the sorted keys.  Every step is
// deterministic, so the result equals SciPy's bit for bit.

#include "pdlp_common.hpp"

#include <hipcub/hipcub.hpp>

namespace pdlp {
namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int key_bits(int32_t cols) {
    int b = 1;
    while (b < 31 && (int64_t(1) << b) < (int64_t)cols) ++b;
    return b;
}

// Entry e of row r (found by binary search over the row pointers): its row
// index and its id, the value the sort carries.
__global__ void __launch_bounds__(256) entry_rows_kernel(const int32_t* __restrict__ rp, int32_t rows, int32_t nnz,
                                                         int32_t* __restrict__ row_of, int32_t* __restrict__ ids) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    int lo = 0, hi = rows - 1;                  // last row with rp[row] <= e
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rp[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    row_of[e] = lo;
    ids[e] = e;
}

template <typename V>
__global__ void __launch_bounds__(256) gather_kernel(const int32_t* __restrict__ sorted_ids,
                                                     const int32_t* __restrict__ row_of,
                                                     const V* __restrict__ val, int32_t nnz,
                                                     int32_t* __restrict__ t_col_idx, V* __restrict__ t_val) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int e = sorted_ids[k];
    t_col_idx[k] = row_of[e];
    t_val[k] = val[e];
}

// t_row_ptr[c] = first position of a key >= c.  Position k writes the pointers
// of the columns between the previous key and its own; position nnz closes the
// columns after the last key.  Indices are kept within [0, cols].
__global__ void __launch_bounds__(256) pointers_kernel(const int32_t* __restrict__ keys, int32_t nnz, int32_t cols,
                                                       int32_t* __restrict__ t_row_ptr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nnz) return;
    const int prev = k == 0 ? -1 : keys[k - 1];
    const int cur = k == nnz ? cols : min(keys[k], cols);
    for (int c = max(prev + 1, 0); c <= cur && c <= cols; ++c) t_row_ptr[c] = k;
}

struct Layout {
    size_t sorted_keys, ids, sorted_ids, row_of, temp, temp_bytes, total;
};

int layout(int32_t cols, int32_t nnz, Layout& l) {
    size_t temp = 0;
    if (nnz > 0) {
        const hipError_t e = hipcub::DeviceRadixSort::SortPairs(
            nullptr, temp, (const int32_t*)nullptr, (int32_t*)nullptr, (const int32_t*)nullptr, (int32_t*)nullptr,
            nnz, 0, key_bits(cols));
        if (e != hipSuccess) return hip_fail(e, "radix sort workspace query");
    }
    const size_t v = align_up((size_t)nnz * sizeof(int32_t));
    l.sorted_keys = 0;
    l.ids = v;
    l.sorted_ids = 2 * v;
    l.row_of = 3 * v;
    l.temp = 4 * v;
    l.temp_bytes = temp;
    l.total = 4 * v + align_up(temp);
    return 0;
}

// Row r's entries into SELL-64 slot order: entry k to sell_ptr[r / 64] + 64 k + r % 64.
template <typename T>
__global__ void __launch_bounds__(256) csr_to_sell64_kernel(int rows, const int32_t* __restrict__ row_ptr,
                                                            const int32_t* __restrict__ col_idx,
                                                            const T* __restrict__ val,
                                                            const int32_t* __restrict__ sell_ptr,
                                                            int32_t* __restrict__ sell_idx, T* __restrict__ sell_val) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int beg = row_ptr[r], end = row_ptr[r + 1];
    const int base = sell_ptr[r >> 6] + (r & 63), lim = sell_ptr[(r >> 6) + 1];
    for (int e = beg; e < end; ++e) {
        const int q = base + 64 * (e - beg);
        if (q >= lim) break;                 // a slice narrower than its rows: never past the slice
        sell_idx[q] = col_idx[e];
        sell_val[q] = val[e];
    }
}

// Arena transpose (many CSR matrices stored one after another, 16-bit local
// indices): the block-diagonal union's global row pointers and column indices,
// then the one transpose, then local row indices and per-matrix column pointers.
// In the union, matrix s's entries occupy [nz_off[s], nz_off[s+1]) both in K and in
// K^T, so K^T's values need no further move.

// index s with off[s] <= v < off[s+1] (off ascending, off[0] = 0, v < off[S]);
// empty ranges are skipped (the last s with off[s] <= v)
__device__ __forceinline__ int range_of(const int32_t* __restrict__ off, int S, int v) {
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) arena_rows_kernel(int S, int rows, const int32_t* __restrict__ row_off,
                                                         const int32_t* __restrict__ nz_off,
                                                         const int32_t* __restrict__ rp_off,
                                                         const int32_t* __restrict__ row_ptr,
                                                         int32_t* __restrict__ gptr) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > rows) return;
    if (g == rows) {
        gptr[rows] = nz_off[S];
        return;
    }
    const int s = range_of(row_off, S, g);
    gptr[g] = nz_off[s] + row_ptr[rp_off[s] + (g - row_off[s])];
}

__global__ void __launch_bounds__(256) arena_cols_kernel(int S, int nnz, const int32_t* __restrict__ nz_off,
                                                         const int32_t* __restrict__ col_off,
                                                         const uint16_t* __restrict__ col_idx,
                                                         int32_t* __restrict__ gcol) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    gcol[e] = (int32_t)col_idx[e] + col_off[range_of(nz_off, S, e)];
}

__global__ void __launch_bounds__(256) arena_local_rows_kernel(int S, int nnz, const int32_t* __restrict__ nz_off,
                                                               const int32_t* __restrict__ row_off,
                                                               const int32_t* __restrict__ grow,
                                                               uint16_t* __restrict__ t_row_idx) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    t_row_idx[q] = (uint16_t)(grow[q] - row_off[range_of(nz_off, S, q)]);
}

__global__ void __launch_bounds__(256) arena_col_ptr_kernel(int S, int slots, const int32_t* __restrict__ cp_off,
                                                            const int32_t* __restrict__ col_off,
                                                            const int32_t* __restrict__ nz_off,
                                                            const int32_t* __restrict__ btp,
                                                            int32_t* __restrict__ t_col_ptr) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= slots) return;
    const int s = range_of(cp_off, S, q);
    t_col_ptr[q] = btp[col_off[s] + (q - cp_off[s])] - nz_off[s];
}

struct ArenaLayout {
    size_t gptr, gcol, btp, btc, work, work_bytes, total;
};

int arena_layout(int32_t rows, int32_t cols, int32_t nnz, ArenaLayout& a) {
    size_t tw = 0;
    if (int rc = pdlp_csr_transpose_work_bytes(rows, cols, nnz, &tw)) return rc;
    a.gptr = 0;
    a.gcol = align_up((size_t)(rows + 1) * 4);
    a.btp = a.gcol + align_up((size_t)nnz * 4);
    a.btc = a.btp + align_up((size_t)(cols + 1) * 4);
    a.work = a.btc + align_up((size_t)nnz * 4);
    a.work_bytes = tw;
    a.total = a.work + align_up(tw);
    return 0;
}

// Column step of each listed row: one wavefront per row compares every step between
// neighbouring entries with the first; 1 for a single entry, 0 for none or no constant step.
__global__ void __launch_bounds__(256) row_strides_kernel(int count, const int32_t* __restrict__ rows,
                                                          const int32_t* __restrict__ row_ptr,
                                                          const int32_t* __restrict__ col_idx,
                                                          int32_t* __restrict__ stride) {
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= count) return;                         // whole wavefronts leave together
    const int lane = threadIdx.x & 63;
    const int r = rows[w];
    const int beg = row_ptr[r], end = row_ptr[r + 1];
    if (end - beg < 2) {
        if (lane == 0) stride[w] = end - beg;       // 0 entries: 0, one entry: 1
        return;
    }
    const int st = col_idx[beg + 1] - col_idx[beg];
    bool same = true;
    for (int e = beg + 1 + lane; e < end; e += 64) same = same && (col_idx[e] - col_idx[e - 1] == st);
    const bool all = __all(same);
    if (lane == 0) stride[w] = all ? st : 0;
}

// Width of each SELL-64 slice: the longest of its 64 rows (rows past the end count 0).
__global__ void __launch_bounds__(256) sell64_widths_kernel(int rows, const int32_t* __restrict__ row_ptr,
                                                            int32_t* __restrict__ width) {
    const int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int r = s * 64 + (threadIdx.x & 63);
    int len = r < rows ? row_ptr[r + 1] - row_ptr[r] : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, 64));
    if ((threadIdx.x & 63) == 0 && s * 64 < rows) width[s] = len;
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" {

int pdlp_csr_to_sell64(int32_t prec, int32_t rows, const int32_t* row_ptr, const int32_t* col_idx, const void* val,
                       const int32_t* sell_ptr, int32_t* sell_idx, void* sell_val, void* stream) {
    if (prec != PDLP_F32 && prec != PDLP_F64) return fail(PDLP_E_ARG, "csr_to_sell64: prec must be PDLP_F32 or PDLP_F64");
    if (rows < 0) return fail(PDLP_E_SHAPE, "csr_to_sell64: rows >= 0");
    if (rows == 0) return 0;
    if (!row_ptr || !col_idx || !val || !sell_ptr || !sell_idx || !sell_val)
        return fail(PDLP_E_ARG, "csr_to_sell64: NULL pointer");
    const unsigned g = (unsigned)((rows + 255) / 256);
    hipStream_t s = as_stream(stream);
    if (prec == PDLP_F32)
        csr_to_sell64_kernel<float><<<g, 256, 0, s>>>(rows, row_ptr, col_idx, (const float*)val, sell_ptr, sell_idx,
                                                      (float*)sell_val);
    else
        csr_to_sell64_kernel<double><<<g, 256, 0, s>>>(rows, row_ptr, col_idx, (const double*)val, sell_ptr,
                                                       sell_idx, (double*)sell_val);
    return hip_fail(hipGetLastError(), "csr_to_sell64_kernel launch");
}

int pdlp_row_strides(int32_t count, const int32_t* rows, const int32_t* row_ptr, const int32_t* col_idx,
                     int32_t* stride, void* stream) {
    if (count < 0) return fail(PDLP_E_SHAPE, "row_strides: count >= 0");
    if (count == 0) return 0;
    if (!rows || !row_ptr || !col_idx || !stride) return fail(PDLP_E_ARG, "row_strides: NULL pointer");
    row_strides_kernel<<<(unsigned)((count + 3) / 4), 256, 0, as_stream(stream)>>>(count, rows, row_ptr, col_idx,
                                                                                   stride);
    return hip_fail(hipGetLastError(), "row_strides_kernel launch");
}

int pdlp_sell64_widths(int32_t rows, const int32_t* row_ptr, int32_t* width, void* stream) {
    if (rows < 0) return fail(PDLP_E_SHAPE, "sell64_widths: rows >= 0");
    if (rows == 0) return 0;
    if (!row_ptr || !width) return fail(PDLP_E_ARG, "sell64_widths: NULL pointer");
    const int slices = (rows + 63) / 64;
    sell64_widths_kernel<<<(unsigned)((slices + 3) / 4), 256, 0, as_stream(stream)>>>(rows, row_ptr, width);
    return hip_fail(hipGetLastError(), "sell64_widths_kernel launch");
}

int pdlp_arena_transpose_work_bytes(int32_t mats, int32_t rows, int32_t cols, int32_t nnz, size_t* bytes) {
    if (bytes == nullptr) return fail(PDLP_E_ARG, "arena_transpose: bytes is NULL");
    if (mats < 1 || rows < 0 || cols < 0 || nnz < 0) return fail(PDLP_E_SHAPE, "arena_transpose: mats >= 1, sizes >= 0");
    ArenaLayout a;
    if (int rc = arena_layout(rows, cols, nnz, a)) return rc;
    *bytes = a.total;
    return 0;
}

int pdlp_arena_transpose(int32_t prec, int32_t mats, int32_t rows, int32_t cols, int32_t nnz,
                         const int32_t* row_off, const int32_t* col_off, const int32_t* nz_off,
                         const int32_t* rp_off, const int32_t* cp_off, const int32_t* row_ptr,
                         const uint16_t* col_idx, const void* val, int32_t* t_col_ptr, uint16_t* t_row_idx,
                         void* t_val, void* work, size_t work_bytes, void* stream) {
    if (prec != PDLP_F32 && prec != PDLP_F64) return fail(PDLP_E_ARG, "arena_transpose: prec must be PDLP_F32 or PDLP_F64");
    if (mats < 1 || rows < 0 || cols < 0 || nnz < 0) return fail(PDLP_E_SHAPE, "arena_transpose: mats >= 1, sizes >= 0");
    if (!row_off || !col_off || !nz_off || !rp_off || !cp_off || !row_ptr || !t_col_ptr ||
        (nnz > 0 && (!col_idx || !val || !t_row_idx || !t_val)))
        return fail(PDLP_E_ARG, "arena_transpose: NULL array");
    ArenaLayout a;
    if (int rc = arena_layout(rows, cols, nnz, a)) return rc;
    if (work == nullptr || work_bytes < a.total)
        return fail(PDLP_E_ARG, "arena_transpose: work is smaller than pdlp_arena_transpose_work_bytes");
    const hipStream_t s = as_stream(stream);
    char* w = (char*)work;
    int32_t* gptr = (int32_t*)(w + a.gptr);
    int32_t* gcol = (int32_t*)(w + a.gcol);
    int32_t* btp = (int32_t*)(w + a.btp);
    int32_t* btc = (int32_t*)(w + a.btc);
    constexpr int B = 256;
    arena_rows_kernel<<<(unsigned)((rows + 1 + B - 1) / B), B, 0, s>>>(mats, rows, row_off, nz_off, rp_off, row_ptr,
                                                                      gptr);
    if (nnz > 0)
        arena_cols_kernel<<<(unsigned)((nnz + B - 1) / B), B, 0, s>>>(mats, nnz, nz_off, col_off, col_idx, gcol);
    if (int rc = hip_fail(hipGetLastError(), "arena index kernels launch")) return rc;
    if (int rc = pdlp_csr_transpose(prec, rows, cols, nnz, gptr, gcol, val, btp, btc, t_val, w + a.work,
                                    a.work_bytes, stream))
        return rc;
    if (nnz > 0)
        arena_local_rows_kernel<<<(unsigned)((nnz + B - 1) / B), B, 0, s>>>(mats, nnz, nz_off, row_off, btc,
                                                                           t_row_idx);
    const int slots = cols + mats;                       // n + 1 pointers per matrix
    arena_col_ptr_kernel<<<(unsigned)((slots + B - 1) / B), B, 0, s>>>(mats, slots, cp_off, col_off, nz_off, btp,
                                                                       t_col_ptr);
    return hip_fail(hipGetLastError(), "arena pointer kernels launch");
}

int pdlp_csr_transpose_work_bytes(int32_t rows, int32_t cols, int32_t nnz, size_t* bytes) {
    if (bytes == nullptr) return fail(PDLP_E_ARG, "bytes is NULL");
    if (rows < 0 || cols < 0 || nnz < 0) return fail(PDLP_E_SHAPE, "negative size");
    Layout l;
    if (int rc = layout(cols, nnz, l)) return rc;
    *bytes = l.total;
    return 0;
}

int pdlp_csr_transpose(int32_t prec, int32_t rows, int32_t cols, int32_t nnz, const int32_t* row_ptr,
                       const int32_t* col_idx, const void* val, int32_t* t_row_ptr, int32_t* t_col_idx,
                       void* t_val, void* work, size_t work_bytes, void* stream) {
    if (prec != PDLP_F32 && prec != PDLP_F64) return fail(PDLP_E_ARG, "prec must be PDLP_F32 or PDLP_F64");
    if (rows < 0 || cols < 0 || nnz < 0) return fail(PDLP_E_SHAPE, "negative size");
    if (nnz > 0 && rows == 0) return fail(PDLP_E_SHAPE, "entries without rows");
    if (t_row_ptr == nullptr || (nnz > 0 && (row_ptr == nullptr || col_idx == nullptr || val == nullptr ||
                                             t_col_idx == nullptr || t_val == nullptr)))
        return fail(PDLP_E_ARG, "NULL array");
    Layout l;
    if (int rc = layout(cols, nnz, l)) return rc;
    if (nnz > 0 && (work == nullptr || work_bytes < l.total))
        return fail(PDLP_E_ARG, "work is smaller than pdlp_csr_transpose_work_bytes");
    const hipStream_t s = (hipStream_t)stream;
    constexpr int B = 256;
    char* w = (char*)work;
    int32_t* sorted_keys = (int32_t*)(w + l.sorted_keys);
    int32_t* ids = (int32_t*)(w + l.ids);
    int32_t* sorted_ids = (int32_t*)(w + l.sorted_ids);
    int32_t* row_of = (int32_t*)(w + l.row_of);
    if (nnz > 0) {
        const int g = (nnz + B - 1) / B;
        entry_rows_kernel<<<g, B, 0, s>>>(row_ptr, rows, nnz, row_of, ids);
        if (int rc = hip_fail(hipGetLastError(), "entry_rows_kernel launch")) return rc;
        size_t temp = l.temp_bytes;
        const hipError_t e = hipcub::DeviceRadixSort::SortPairs(w + l.temp, temp, col_idx, sorted_keys, ids,
                                                                sorted_ids, nnz, 0, key_bits(cols), s);
        if (e != hipSuccess) return hip_fail(e, "radix sort");
        if (prec == PDLP_F32)
            gather_kernel<float><<<g, B, 0, s>>>(sorted_ids, row_of, (const float*)val, nnz, t_col_idx,
                                                 (float*)t_val);
        else
            gather_kernel<double><<<g, B, 0, s>>>(sorted_ids, row_of, (const double*)val, nnz, t_col_idx,
                                                  (double*)t_val);
        if (int rc = hip_fail(hipGetLastError(), "gather_kernel launch")) return rc;
    }
    pointers_kernel<<<(nnz + 1 + B - 1) / B, B, 0, s>>>(sorted_keys, nnz, cols, t_row_ptr);
    return hip_fail(hipGetLastError(), "pointers_kernel launch");
}

}  // extern "C"
