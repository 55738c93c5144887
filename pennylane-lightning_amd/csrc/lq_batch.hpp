// Packed small-request execution: many independent energy + adjoint-gradient
// requests in one launch, one workgroup per request, with the request's two
// states (lambda and mu = H lambda) resident in LDS for the whole forward
// circuit, the Hamiltonian application, the energy and the backward sweep
// (lq/algorithms/AdjointJacobianLQubit.hpp:225-319 for one observable).
// No HBM traffic besides the program and the results.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "lq_engine.hpp"

namespace lqamd {

enum BOpKind : uint8_t {
    BOP_DENSE1 = 1, // 2x2 on bit t0
    BOP_DENSE2 = 2, // 4x4 on bits (t0 = local MSB, t1)
    BOP_DIAG = 3,   // 2^k diagonal, k = 0..2, on (t0, t1)
    BOP_PARITY = 4, // coef[popc(i & pmask) & 1]
    BOP_HAM = 5,    // state 1 = sum_t c_t P_t state 0 (terms [arg, arg + arg2))
    BOP_DOTID = 6,  // slot arg2 = sum conj(state 1) state 0
};
enum : uint8_t { BOP_F_DOT = 1 };

// One op (24 bytes). Gate ops act on the states in smask (bit 0 = lambda,
// bit 1 = mu); dot ops accumulate sum conj(mu) (G lambda) over the elements
// whose control bits match, into result slot arg2.
struct BOp {
    uint8_t kind, flags, smask, k;
    uint8_t t0, t1, pad0, pad1;
    uint32_t cmask, cval;
    uint32_t arg;  // coefficient offset (double2 units); HAM: first term
    uint32_t arg2; // dots: slot; PARITY: pmask; HAM: term count
};
static_assert(sizeof(BOp) == 24, "BOp layout");

struct BReq {
    uint32_t n;
    uint32_t op_begin, op_end;
    uint32_t pad;
    uint64_t out_off;   // first result slot (double2 units)
    uint64_t state_off; // states in global memory: first amplitude of lambda (mu follows)
};

// Largest state a workgroup keeps in LDS: 2 states x 2^n x element bytes
// within 128 KiB (complex128: 12 qubits, complex64: 13). Up to
// batch_max_qubits the workgroup works on its two states in global memory
// (L2 / Infinity Cache resident for these sizes) instead.
inline int batch_lds_qubits(int dtype) { return dtype ? 12 : 13; }
inline int batch_max_qubits(int) { return 16; }

// in_lds: every request has n <= batch_lds_qubits (max_n sizes the LDS);
// otherwise the states live in `scratch` at each request's state_off.
// Copy `bytes` (a multiple of 16) with a kernel on stream `s` (no copy engine
// involved); either side may be device-visible host memory.
void launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t s);

void launch_batch(const BReq *reqs, const BOp *ops, const double2 *coef, const PauliTerm *terms, double2 *out,
                  void *scratch, uint32_t nreq, int max_n, bool in_lds, int dtype, hipStream_t s);

// Host program of one request.
struct BatchProgram {
    int n = 0;
    std::vector<BOp> ops;
    std::vector<double2> coef;
    std::vector<PauliTerm> terms;
    std::vector<double> scales; // per trainable parameter (sorted order)
    uint32_t nslots = 0;        // slot 0 = <mu|lambda> (energy), 1 + t = gradient dot t
};

// Build the packed program of: forward `ops` from |0..0>, mu = H lambda,
// energy, backward sweep over the trainable parameters `tp`. Returns false
// (with a reason) when the request does not fit the packed kernel: too many
// qubits, a Hamiltonian with Hermitian factors, ops the kernel lacks.
bool build_batch_program(int n, int dtype, const OpsList &ops, const Obs &ham, const std::vector<uint64_t> &tp,
                         BatchProgram *out, std::string *why);

// Device buffers + stream for packed launches (one per concurrent launch).
class BatchRunner {
  public:
    BatchRunner();
    ~BatchRunner();
    BatchRunner(const BatchRunner &) = delete;
    BatchRunner &operator=(const BatchRunner &) = delete;

    // Upload and launch (asynchronous on this runner's stream); results are
    // read with finish().
    void launch(const std::vector<const BatchProgram *> &progs, int dtype);
    bool done();   // launch finished (non-blocking)
    // Wait, then energies[i] and grads[i][t] for the programs of the last launch.
    void finish(std::vector<double> *energies, std::vector<std::vector<double>> *grads);
    hipStream_t stream() const { return stream_; }
    // GPU time of the last finished launch's upload, kernels and download, ms
    // (device timestamps of events on the runner's stream)
    void segments(double out3[3]) const;
    // Bytes one program adds to a launch's device buffer.
    static size_t program_bytes(const BatchProgram &p);
    // Program bytes (program_bytes summed) a launch can take without growing
    // this runner's buffers: growing frees the old ones, and hipFree /
    // hipHostFree wait for the whole device.
    size_t capacity() const;

  private:
    hipStream_t stream_ = nullptr;
    hipEvent_t done_ = nullptr;
    hipEvent_t t0_ = nullptr, t_in_ = nullptr, t_k_ = nullptr, t_out_ = nullptr;
    bool kernel_upload_ = true;
    double seg_[3] = {0.0, 0.0, 0.0};
    DevBuf dbuf_, scratch_;
    HostPinned hin_, hout_;
    std::vector<const BatchProgram *> progs_;
    std::vector<uint64_t> out_off_;
    size_t out_slots_ = 0;
};

} // namespace lqamd
