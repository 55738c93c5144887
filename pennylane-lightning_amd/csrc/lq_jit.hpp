// Pass specialisation: a planned register pass becomes straight-line HIP
// source (register bits, control splits, CZ parities and tile geometry as
// compile-time constants; coefficients stay in the runtime buffer), compiled
// with hiprtc for gfx950 and cached by its source text. A circuit structure
// that repeats (variational loops, the adjoint sweep of a fixed ansatz, the
// benchmark) runs the specialised kernels; until a kernel is compiled the
// pass runs in the generic tile kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "lq_plan.hpp"

namespace lqamd {

// Source for a register pass of n qubits, dtype 0/1 (complex64/128), ns
// states; empty when the pass has stages the generator does not cover.
// With `geo`, the bit positions of the pass in the state index are not
// written into the source: the kernel reads them from a table placed right
// after the pass's coefficients and the deferred-preparation entry (the
// values are returned in *geo, in table order), so passes of the same shape
// at other positions or qubit counts share the kernel.
std::string jit_pass_source(const PassPlan &p, int n, int dtype, int ns, std::vector<uint64_t> *geo = nullptr);

struct JitConfig {
    bool enabled = true;
    bool sync = false;      // compile before first use instead of in the background
    int min_qubits = 16;    // smaller states always use the generic kernels
    int threads = 8;        // background compile threads
    int nice = 10;          // their scheduling niceness (LQ_JIT_NICE; 0: unchanged)
    bool nontemporal = true;  // state loads/stores marked nontemporal (streamed, not retained)
    bool split_rotations = true; // 2x2 gates as phase / real rotation / phase in specialised passes
    bool scale_rotations = true; // specialised passes: rotations as [[1, -t], [t, 1]]
    bool lds32 = true;        // transposes through 32 KiB of LDS (halves, one state at a time)
                              // so that more workgroups fit per CU
    int waves_per_eu = -1;    // > 0: ask the compiler to fit that many waves per SIMD; 0: no
                              // request; -1: by kind (one-state passes: 3 complex128, 4 complex64)
    bool lift = false;        // bit positions read at run time, so passes of one shape at other
                              // positions or qubit counts share a kernel (LQ_JIT_LIFT=1). Off:
                              // on the serving stream's 5315 passes (n 17..28, depth 4..16) it
                              // cuts distinct kernels only from 3104 to 2331
    int lane_swaps = 0;       // register-bit changes whose incoming bits are lane bits, without
                              // LDS: 1 only lane bits 4 and 5 (v_permlane16/32_swap), 2 any
                              // lane bit (bits 4, 5 as 1; lower bits by ds_bpermute shuffles),
                              // 0 none (LQ_JIT_LANES). Measured on the 30-qubit HEA: 1 no faster
                              // than the LDS transposes, 2 slower; off by default
    int direct_io = 1;        // loads/stores in the first/last stage's mapping (no entry/exit
                              // transpose): 0 never, 1 where measured to pay, 2 whenever they coalesce
};
JitConfig &jit_config();
// Waves per SIMD requested of a specialised pass kernel (0: none).
int jit_waves_per_eu(int dtype, int ns);
// Dynamic LDS bytes of a specialised register pass kernel.
size_t jit_lds_bytes(int kt, int dtype, int ns);

// Kernel for `src` if compiled and loaded on the current device; otherwise
// nullptr, after queueing the compile (or compiling now in sync mode).
hipFunction_t jit_lookup(const std::string &src);
// Same for a planned pass, found by its structure without generating the
// source again once the pass has been seen.
// *geo (may be null) receives the pass's position table (empty when the
// kernel has the positions in its source); it stays valid for the process.
hipFunction_t jit_lookup_plan(const PassPlan &p, int n, int dtype, int ns,
                              const std::vector<uint64_t> **geo = nullptr);
// Block until every queued compile has finished.
void jit_wait();
// Drop queued compiles and wait for those in flight; no new compiles start
// afterwards (kernels already compiled stay usable).
void jit_shutdown();
// Counters: kernels ready, compiles failed, compiles pending.
void jit_stats(int *ready, int *failed, int *pending);

// Launch a specialised pass (arguments as the generic register kernel).
void jit_launch(hipFunction_t f, int dtype, void *sv0, void *sv1, void *sv2, void *sv3, const void *coef,
                void *partials, uint64_t ntiles, size_t lds_bytes, hipStream_t s);

} // namespace lqamd
