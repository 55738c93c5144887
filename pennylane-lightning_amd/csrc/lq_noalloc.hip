// Probabilities of a state held in caller-owned device memory, with the
// per-block partial histograms in a fixed device array instead of memory
// the library allocates: for callers that cannot allocate (code under HIP
// graph capture, tight serving loops). The kernels are the ones lq_probs
// runs (LDS histogram per block, then a compensated sum of the rows).
//
// The array `scratch` is a file-scope __device__ variable: each GPU holds its
// own copy until the process ends. Its name in host code is not a device
// address; the address is taken once per device with hipGetSymbolAddress
// (with that device current) and only that pointer is passed to kernels.
// Every probs block writes its whole row of partials, so nothing needs
// clearing between calls; calls that overlap on one GPU share the array and
// race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lq_engine.hpp"
#include "lq_kernels.hpp"

namespace lqamd {
namespace {

constexpr uint32_t kRows = 512; // histogram blocks, as lq_probs uses at most
static __device__ __align__(256) double scratch[static_cast<size_t>(kRows) << kNoAllocProbsMaxWires];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

double *scratch_on_current_device() {
    int dev = -1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    LQ_ABORT_IF(dev < 0 || dev >= kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        LQ_ABORT_IF(p == nullptr, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch[dev] = static_cast<double *>(p);
    }
    return g_scratch[dev];
}

} // namespace

void probs_noalloc_init() { (void)scratch_on_current_device(); }

void probs_noalloc(const void *state, int n, int dtype, const uint32_t *wires, int nw, double *out, hipStream_t s) {
    LQ_ABORT_IF(state == nullptr || out == nullptr || wires == nullptr, "null argument");
    LQ_ABORT_IF(n < 1 || n > 40, "number of qubits must be in [1, 40]");
    LQ_ABORT_IF(dtype != 0 && dtype != 1, "dtype must be LQ_C64 or LQ_C128");
    LQ_ABORT_IF(nw < 1 || nw > kNoAllocProbsMaxWires || nw > n,
                "probabilities without allocation are limited to 1..10 wires (and at most the number of qubits)");
    int bits[kNoAllocProbsMaxWires];
    uint64_t seen = 0;
    for (int j = 0; j < nw; ++j) {
        LQ_ABORT_IF(wires[j] >= static_cast<uint32_t>(n), "Invalid wire index");
        const int b = n - 1 - static_cast<int>(wires[j]);
        LQ_ABORT_IF((seen >> b) & 1ull, "wires must be unique");
        seen |= 1ull << b;
        bits[j] = b;
    }
    double *part = scratch_on_current_device();
    // the block count lq_probs uses for this size (one row of partials each)
    const uint64_t len = 1ull << n;
    const uint32_t nb = static_cast<uint32_t>(len / 1024 < 1 ? 1 : (len / 1024 > kRows ? kRows : len / 1024));
    if (dtype) launch_probs_hist<double>(state, n, bits, nw, part, nb, s);
    else launch_probs_hist<float>(state, n, bits, nw, part, nb, s);
    launch_reduce_rows(part, nb, 1u << nw, out, s);
    hip_check(hipGetLastError(), "probs launch");
}

} // namespace lqamd
