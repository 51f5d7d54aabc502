// Node-model plumbing of the learned warm start: input/target rows gathered
// from compact per-instance vectors, and the sparse map from per-node network
// outputs back to a start [x, y, log weight] per LP.
//
// Gather: a group of `lanes` lanes per output row (the power of two covering the
// row, at most 64: one row per wavefront for 48 inputs, eight rows for 8
// outputs), lane f writing entries f, f + lanes, ... of the row, so a row's
// values go out as one contiguous store and the table row is one coalesced
// load; the instance's source vector is re-read by the rows of the same
// instance and stays in L2.
//
// Assembly: one thread per (instance, start entry); the entries of one
// instance read that instance's block of network rows.

#include "pdlp_common.hpp"

#include <algorithm>

namespace pdlp {
namespace {

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// splitmix64 finaliser: a bijective mix of (seed, r)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

constexpr int GATHER_THREADS = 256;

constexpr int GATHER_ROWS = 4;   // rows of a lane group whose loads go out together

// Which optional arrays are given is a template choice, so the element loop
// has no branch around a load: the table entry, scale, mean and 1/std of an
// element go out together, then the one source value the entry selects.
template <bool SCALE, bool MEAN, bool STD, bool BF16>
__global__ void __launch_bounds__(GATHER_THREADS) node_gather_kernel(PdlpNodeGather g, int64_t count, int lanes) {
    const int lane = threadIdx.x & (lanes - 1);
    const int per_block = GATHER_THREADS / lanes;
    const int64_t groups = (int64_t)gridDim.x * per_block;
    const int64_t total = (int64_t)g.batch * g.nodes;
    const int32_t* __restrict__ table = g.table;
    const float* __restrict__ shared = g.shared;
    const float* __restrict__ scale = g.scale;
    const float* __restrict__ mean = g.mean;
    const float* __restrict__ inv_std = g.inv_std;
    const int base2 = g.src_len + g.shared_len;
    // GATHER_ROWS rows per pass (r0, r0 + groups, ...), each element computed as alone
    for (int64_t r0 = blockIdx.x * (int64_t)per_block + threadIdx.x / lanes; r0 < count;
         r0 += GATHER_ROWS * groups) {
        int64_t rr[GATHER_ROWS];
        int j[GATHER_ROWS];
        const float* __restrict__ src[GATHER_ROWS];
        const float* __restrict__ src2[GATHER_ROWS];
#pragma unroll
        for (int u = 0; u < GATHER_ROWS; ++u) {
            rr[u] = r0 + u * groups;
            const int64_t r = rr[u] < count ? rr[u] : r0;          // a missing row repeats the first
            int64_t q;
            if (g.rows != nullptr) q = g.rows[r];
            else if (g.seed != 0) q = (int64_t)(mix64(g.seed ^ mix64((uint64_t)r)) % (uint64_t)total);
            else q = r;
            const int64_t b = q / g.nodes;
            j[u] = (int)(q - b * g.nodes);
            src[u] = g.src + b * g.src_len;
            src2[u] = g.src2 + b * g.src2_len;
        }
        for (int f = lane; f < g.ld_out; f += lanes) {
            float v[GATHER_ROWS];
#pragma unroll
            for (int u = 0; u < GATHER_ROWS; ++u) v[u] = 0.f;
            if (g.width > 0) {                          // uniform: the table exists
                const bool in = f < g.width;
                const int fc = in ? f : 0;
                const float mu = MEAN ? mean[fc] : 0.f;
                const float is = STD ? inv_std[fc] : 1.f;
                size_t e[GATHER_ROWS];
                int t[GATHER_ROWS];
                float sc[GATHER_ROWS];
#pragma unroll
                for (int u = 0; u < GATHER_ROWS; ++u) {
                    e[u] = (size_t)j[u] * g.width + fc;
                    t[u] = table[e[u]];
                    sc[u] = SCALE ? scale[e[u]] : 1.f;
                }
#pragma unroll
                for (int u = 0; u < GATHER_ROWS; ++u) {
                    // the value's address (an unused entry reads its own table word)
                    const int tu = t[u];
                    const float* at = tu < 0 ? reinterpret_cast<const float*>(table + e[u])
                                      : tu < g.src_len ? src[u] + tu
                                      : tu < base2     ? shared + (tu - g.src_len)
                                                       : src2[u] + (tu - base2);
                    float x = *at;
                    x = tu < 0 ? 0.f : x;
                    if (SCALE) x *= sc[u];
                    if (MEAN) x -= mu;
                    if (STD) x *= is;
                    v[u] = in ? x : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < GATHER_ROWS; ++u) {
                if (rr[u] >= count) continue;
                const size_t o = (size_t)rr[u] * g.ld_out + f;
                if (BF16) g.out_bf16[o] = to_bf16(v[u]);
                else g.out_f32[o] = v[u];
            }
        }
    }
}

template <bool SCALE, bool MEAN, bool STD>
void launch_gather3(const PdlpNodeGather& g, int64_t count, int lanes, unsigned blocks, hipStream_t s) {
    if (g.out_bf16 != nullptr) node_gather_kernel<SCALE, MEAN, STD, true><<<blocks, GATHER_THREADS, 0, s>>>(g, count, lanes);
    else node_gather_kernel<SCALE, MEAN, STD, false><<<blocks, GATHER_THREADS, 0, s>>>(g, count, lanes);
}

void launch_gather(const PdlpNodeGather& g, int64_t count, int lanes, unsigned blocks, hipStream_t s) {
    const bool sc = g.scale != nullptr, mu = g.mean != nullptr, is = g.inv_std != nullptr;
    if (sc && mu && is) launch_gather3<true, true, true>(g, count, lanes, blocks, s);
    else if (sc && mu) launch_gather3<true, true, false>(g, count, lanes, blocks, s);
    else if (sc && is) launch_gather3<true, false, true>(g, count, lanes, blocks, s);
    else if (sc) launch_gather3<true, false, false>(g, count, lanes, blocks, s);
    else if (mu && is) launch_gather3<false, true, true>(g, count, lanes, blocks, s);
    else if (mu) launch_gather3<false, true, false>(g, count, lanes, blocks, s);
    else if (is) launch_gather3<false, false, true>(g, count, lanes, blocks, s);
    else launch_gather3<false, false, false>(g, count, lanes, blocks, s);
}

__global__ void __launch_bounds__(256) node_assemble_kernel(PdlpNodeAssembly a) {
    const int b = blockIdx.y;
    const float* net = a.net + (size_t)b * a.nodes * a.ld_net;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.len; i += gridDim.x * blockDim.x) {
        float s = a.bias != nullptr ? a.bias[i] : 0.f;
        const int beg = a.indptr[i], end = a.indptr[i + 1];
        int e = beg;
        // four entries' loads together (a demand dual averages 200 predictions), added in order
        for (; e + 4 <= end; e += 4) {
            float wv[4], nv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                wv[u] = a.weight[e + u];
                nv[u] = net[(size_t)a.node[e + u] * a.ld_net + a.chan[e + u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) s += wv[u] * nv[u];
        }
        for (; e < end; ++e) s += a.weight[e] * net[(size_t)a.node[e] * a.ld_net + a.chan[e]];
        a.out[(size_t)b * a.ld_out + i] = s;
    }
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" {

int pdlp_node_gather(const PdlpNodeGather* g, void* stream) {
    if (g == nullptr) return fail(PDLP_E_ARG, "PdlpNodeGather is NULL");
    if (g->batch < 0 || g->nodes < 1 || g->width < 0 || g->ld_out < g->width || g->src_len < 0 ||
        g->shared_len < 0 || g->src2_len < 0 || g->count < 0)
        return fail(PDLP_E_SHAPE, "pdlp_node_gather: batch >= 0, nodes >= 1, ld_out >= width >= 0, lengths >= 0");
    if ((g->out_bf16 == nullptr) == (g->out_f32 == nullptr))
        return fail(PDLP_E_ARG, "pdlp_node_gather: exactly one of out_bf16 / out_f32");
    if (g->width > 0 && g->table == nullptr) return fail(PDLP_E_ARG, "pdlp_node_gather: table is NULL");
    if (g->src_len > 0 && g->batch > 0 && g->src == nullptr) return fail(PDLP_E_ARG, "pdlp_node_gather: src is NULL");
    if (g->shared_len > 0 && g->shared == nullptr) return fail(PDLP_E_ARG, "pdlp_node_gather: shared is NULL");
    if (g->src2_len > 0 && g->batch > 0 && g->src2 == nullptr)
        return fail(PDLP_E_ARG, "pdlp_node_gather: src2 is NULL");
    const int64_t total = (int64_t)g->batch * g->nodes;
    const bool listed = g->rows != nullptr || g->seed != 0;
    const int64_t count = listed ? g->count : total;
    if (listed && count > 0 && total == 0) return fail(PDLP_E_SHAPE, "pdlp_node_gather: rows drawn from an empty batch");
    if (count == 0 || g->ld_out == 0) return 0;
    int lanes = 1;
    while (lanes < 64 && lanes < g->ld_out) lanes *= 2;
    const int64_t rpb = GATHER_THREADS / lanes * GATHER_ROWS;    // rows per lane group and pass
    const unsigned blocks = (unsigned)std::min<int64_t>((count + rpb - 1) / rpb, 65536);
    launch_gather(*g, count, lanes, blocks, as_stream(stream));
    return hip_fail(hipGetLastError(), "node_gather_kernel launch");
}

int pdlp_node_assemble(const PdlpNodeAssembly* a, void* stream) {
    if (a == nullptr) return fail(PDLP_E_ARG, "PdlpNodeAssembly is NULL");
    if (a->batch < 0 || a->nodes < 1 || a->ld_net < 1 || a->len < 0 || a->ld_out < a->len || a->nnz < 0)
        return fail(PDLP_E_SHAPE, "pdlp_node_assemble: batch >= 0, nodes >= 1, ld_net >= 1, ld_out >= len >= 0");
    if (a->batch > 65535) return fail(PDLP_E_LIMIT, "pdlp_node_assemble: at most 65535 instances per call");
    if (a->batch == 0 || a->len == 0) return 0;
    if (!a->indptr || !a->net || !a->out || (a->nnz > 0 && (!a->node || !a->chan || !a->weight)))
        return fail(PDLP_E_ARG, "pdlp_node_assemble: NULL pointer");
    dim3 grid((unsigned)std::min((a->len + 255) / 256, 1024), (unsigned)a->batch);
    node_assemble_kernel<<<grid, 256, 0, as_stream(stream)>>>(*a);
    return hip_fail(hipGetLastError(), "node_assemble_kernel launch");
}

}  // extern "C"
