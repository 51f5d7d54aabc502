// Network inputs of the economic-dispatch family, computed on the device from
// the LPs' own vectors (the per-LP costs and right-hand sides the family
// engine has already staged in HBM), so that a learned start needs no host
// pass over the instances and no upload of its inputs.
//
// One workgroup per instance.  The generators' merit order is a stable sort of
// the first period's costs, done as ranks (rank g = #{h : c_h < c_g, or
// c_h = c_g and h < g}); the capacity cumulated along it is a sequential FP64
// sum in that order; every other term is one IEEE FP64 expression rounded to
// FP32 once, with contraction off, so the output is bitwise the host
// definition's (DispatchFamily.node_inputs_params).

#include "pdlp_common.hpp"

namespace pdlp {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxG = 1024;
constexpr int kMaxT = 256;

__global__ void __launch_bounds__(kThreads) dispatch_inputs_kernel(PdlpDispatchInputs a) {
#pragma clang fp contract(off)
    __shared__ double c[kMaxG], pm[kMaxG], cum[kMaxG], cheap[kMaxG];
    __shared__ int rank[kMaxG], order[kMaxG];
    __shared__ double dd[kMaxT];
    const int k = blockIdx.x;
    const int G = a.G, T = a.T, tid = threadIdx.x;
    const double* ck = a.c + (size_t)k * a.ld_c;
    const double* bk = a.b + (size_t)k * a.ld_b;
    float* v = a.out + (size_t)k * a.ld_out;
    for (int g = tid; g < G; g += kThreads) {
        c[g] = ck[g];
        pm[g] = a.pmax[g];
    }
    for (int t = tid; t < T; t += kThreads) dd[t] = -bk[t];
    __syncthreads();
    for (int g = tid; g < G; g += kThreads) {
        const double cg = c[g];
        int r = 0;
        for (int h = 0; h < G; ++h) r += (c[h] < cg) || (c[h] == cg && h < g);
        rank[g] = r;
        order[r] = g;
    }
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int r = 0; r < G; ++r) {
            s += pm[order[r]];
            cum[r] = s;
        }
    }
    __syncthreads();
    const double S = a.pmax_sum;
    for (int g = tid; g < G; g += kThreads) {
        const double ch = cum[rank[g]] - pm[g];
        cheap[g] = ch;
        v[a.off_c + g] = (float)(c[g] / 60.0);
        v[a.off_u + g] = (float)(c[g] / a.cost[g] - 1.0);
        v[a.off_cc + g] = (float)(2.0 * ch / S);
        v[a.off_cc2 + g] = (float)(2.0 * (ch + pm[g]) / S);
    }
    for (int t = tid; t < T; t += kThreads) {
        const double d = dd[t];
        int cnt = 0;
        for (int r = 0; r < G; ++r) cnt += cum[r] < d;
        const int mg = cnt < G - 1 ? cnt : G - 1;
        v[a.off_d + t] = (float)(4.0 * d / S);
        v[a.off_p + t] = (float)(c[order[mg]] / 60.0);
        v[a.off_m + t] = (float)(d / a.demand[t] - 1.0);
    }
    __syncthreads();
    // merit-order dispatch ignoring ramps, clip(d_t - cheaper_g, 0, Pmax_g) / Pmax_g, period-major
    for (int q = tid; q < T * G; q += kThreads) {
        const int t = q / G, g = q - t * G;
        const double x = fmin(fmax(dd[t] - cheap[g], 0.0), pm[g]);
        v[a.off_x + q] = (float)(x / pm[g]);
    }
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" int pdlp_dispatch_inputs(const PdlpDispatchInputs* a, void* stream) {
    if (a == nullptr) return fail(PDLP_E_ARG, "PdlpDispatchInputs is NULL");
    if (a->count < 0 || a->G < 1 || a->T < 1) return fail(PDLP_E_SHAPE, "dispatch_inputs: count >= 0, G >= 1, T >= 1");
    if (a->G > kMaxG || a->T > kMaxT)
        return fail(PDLP_E_LIMIT, "dispatch_inputs: at most 1024 generators and 256 periods");
    if (a->ld_c < a->G || a->ld_b < a->T) return fail(PDLP_E_SHAPE, "dispatch_inputs: ld_c >= G and ld_b >= T");
    const int64_t offs[8] = {a->off_d, a->off_c, a->off_u, a->off_cc, a->off_cc2, a->off_p, a->off_m, a->off_x};
    const int64_t lens[8] = {a->T, a->G, a->G, a->G, a->G, a->T, a->T, (int64_t)a->T * a->G};
    for (int i = 0; i < 8; ++i)
        if (offs[i] < 0 || offs[i] + lens[i] > a->ld_out)
            return fail(PDLP_E_SHAPE, "dispatch_inputs: an output block lies outside [0, ld_out)");
    if (!(a->pmax_sum > 0.0)) return fail(PDLP_E_ARG, "dispatch_inputs: pmax_sum must be positive");
    if (a->count == 0) return 0;
    if (!a->c || !a->b || !a->pmax || !a->cost || !a->demand || !a->out)
        return fail(PDLP_E_ARG, "dispatch_inputs: NULL array");
    dispatch_inputs_kernel<<<(unsigned)a->count, kThreads, 0, as_stream(stream)>>>(*a);
    return hip_fail(hipGetLastError(), "dispatch_inputs_kernel launch");
}
