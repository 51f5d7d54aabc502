// pdlp_run_noalloc: the single-instance iteration blocks of pdlp_run with the
// kernels' scratch in a static device array instead of caller workspace.
//
// pdlp_run's kernels need two scratch buffers beyond the LP's inputs and
// state: the KKT check's per-workgroup partial sums (PDLP_PARTIALS_LEN
// doubles, written and read within one check) and, for the split SELL-64 dual
// step (PdlpTune.sell >= 8), the partial dot products of each slice chunk
// (written and read within one dual step).  Neither carries data from one call
// to the next, so one array per GPU serves every call made one after another;
// calls that overlap in time on one GPU race on it.

#include "pdlp_common.hpp"

#include <mutex>
#include <string>

// The partials first, then the SELL partials.
static __device__ __align__(256) unsigned char scratch[PDLP_NOALLOC_SCRATCH_BYTES];

namespace pdlp {
namespace {

constexpr size_t kScratchPartialsBytes = (size_t)PDLP_PARTIALS_LEN * sizeof(double);
static_assert(kScratchPartialsBytes % 256 == 0 && kScratchPartialsBytes < PDLP_NOALLOC_SCRATCH_BYTES,
              "the partials must leave room for the SELL partials");
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device
// with that device current (each GPU holds its own copy at its own address).  In
// host code `scratch` names a host-side stand-in: only this address may reach a
// kernel.
int scratch_base(unsigned char** out) {
    int dev = 0;
    if (int rc = hip_fail(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(PDLP_E_LIMIT, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        if (int rc = hip_fail(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return rc;
        if (p == nullptr) return fail(PDLP_E_ARG, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace
}  // namespace pdlp

using namespace pdlp;

extern "C" int pdlp_run_noalloc(const PdlpDev* dev, const PdlpTune* tune, int32_t iters, int32_t check,
                                void* stream) {
    if (dev == nullptr || tune == nullptr) return fail(PDLP_E_ARG, "PdlpDev or PdlpTune is NULL");
    if (iters < 0) return fail(PDLP_E_ARG, "iters must be >= 0");
    PdlpDev d = *dev;
    size_t sell_need = 0;
    if (d.sell_slices > 0 && tune->sell >= 8 && d.sell_max_width > 0) {
        const size_t chunks = ((size_t)d.sell_max_width + (size_t)tune->sell - 1) / (size_t)tune->sell;
        sell_need = (size_t)d.sell_slices * chunks * 64 * (d.prec == PDLP_F64 ? sizeof(double) : sizeof(float));
    }
    const size_t sell_room = PDLP_NOALLOC_SCRATCH_BYTES - kScratchPartialsBytes;
    if (sell_need > sell_room)
        return fail(PDLP_E_LIMIT, "pdlp_run_noalloc: the split SELL step needs " + std::to_string(sell_need) +
                                      " bytes of scratch, more than the " + std::to_string(sell_room) +
                                      " left after the check's partials");
    unsigned char* base = nullptr;
    if (int rc = scratch_base(&base)) return rc;
    d.partials = (double*)base;
    d.sell_part = d.sell_slices > 0 ? (void*)(base + kScratchPartialsBytes) : nullptr;
    if (int rc = single_check_all(&d, tune)) return rc;
    return single_run(d, *tune, iters, check, as_stream(stream));
}
