// Device state vector, tile-program executor, measurements, observables and
// the adjoint Jacobian on one MI355X.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "lq_common.hpp"
#include "lq_gates.hpp"
#include "lq_kernels.hpp"
#include "lq_plan.hpp"

namespace lqamd {

void hip_check(hipError_t e, const char *what);

// Device and pinned host buffers that only grow. hipFree and hipHostFree
// wait for the whole device, so a buffer outgrown while other streams run
// (the concurrent requests of a server) is kept until the buffer itself is
// freed when it is small (kKeepOutgrown); larger ones are freed at once.
constexpr size_t kKeepOutgrown = 64u << 20;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    void ensure(size_t b);
    void release();
  private:
    std::vector<void *> outgrown_;
};

struct HostPinned {
    void *p = nullptr;
    size_t bytes = 0;
    unsigned flags = hipHostMallocDefault; // hipHostMalloc flags of the next allocation
    HostPinned() = default;
    HostPinned(const HostPinned &) = delete;
    HostPinned &operator=(const HostPinned &) = delete;
    ~HostPinned();
    void ensure(size_t b);
  private:
    std::vector<void *> outgrown_;
};

// Observables (core/observables/Observables.hpp:128-584)
struct Obs {
    enum Kind { Named, Hermitian, Tensor, Ham, Sparse } kind = Named;
    std::string name;
    std::vector<uint32_t> wires;
    std::vector<cd> matrix;                 // Hermitian, row-major complex128
    std::vector<std::shared_ptr<Obs>> terms; // Tensor factors / Hamiltonian terms
    std::vector<double> coeffs;             // Hamiltonian
    // Sparse: CSR over the full state index (lq/observables/ObservablesLQubit.hpp:372-436)
    std::vector<cd> sp_values;
    std::vector<int64_t> sp_indices, sp_indptr;
    std::vector<uint32_t> all_wires() const;
};

// CSR matrix over the state index (row i = amplitude index i)
struct CsrView {
    const int64_t *indptr;   // nrows + 1 entries
    uint64_t nrows;
    const int64_t *indices;  // nnz column indices
    const void *values;      // nnz complex values in the state's precision
    uint64_t nnz;
};

struct PauliWord {
    cd coeff;
    uint64_t x = 0, yz = 0; // physical bit masks
    int ny = 0;
};
// Expand into a Pauli sum on physical bits; false if a Hermitian factor occurs.
bool obs_to_pauli(const Obs &o, int n, std::vector<PauliWord> *out);

// One entry of OpsData (core/algorithms/JacobianData.hpp:39-253)
struct OpRec {
    std::string name;
    bool named = false;
    GateId gid = GateId::Identity;
    std::vector<double> params;
    std::vector<uint32_t> wires;
    bool inverse = false;
    std::vector<cd> matrix;
    std::vector<uint32_t> ctrls;
    std::vector<uint8_t> cvals;
};
struct OpsList {
    std::vector<OpRec> ops;
    size_t num_par_ops = 0;
};

// Dense operators on more targets than this run out of place through a
// global-memory kernel instead of as tile ops.
constexpr int kMaxFusedTargets = 10;
// From this many qubits, specialised programs are planned under several
// pass budgets and the plan with the fewest passes is kept.
constexpr int kPlanSearchQubits = 26;

struct EngineTuning {
    int kt_max = 0;      // 0: from the LDS budget
    int low_bits = -1;   // -1: 4 (complex128) / 5 (complex64); specialised passes 4
    double budget = 0.0; // 0: 24 per pass; specialised passes with split rotations 52
    size_t lds_bytes = 65536;
    bool reg = true;     // register-resident passes for large tiles
    bool plan_debug = false;
    bool plan_cache = true; // reuse plans of programs with the same structure
    bool lookahead = true; // lookahead tile-bit choice for specialised passes
    // while some specialised kernels of a split plan are missing: 0 run the
    // unsplit plan, 1 choose by the op-count estimate, 2 run the split plan
    // (its missing passes in the generic kernel)
    // (written through lq_set_split_fallback while other threads plan)
    std::atomic<int> split_fallback{1};
    // complex64 dense operators on exactly the five lowest index bits (no
    // controls) run as a 64x64 real GEMM on the matrix cores instead of as a
    // tile op (lq_set_mfma_dense5 / LQ_MFMA_DENSE5)
    // (measured on a 30-qubit complex64 state: 3.4 ms per operator against
    // 167 ms as a tile op; on by default)
    std::atomic<int> mfma_dense5{1};
    // Leading one-qubit gates after a reset fold into the product state the
    // first pass forms, and for large states several pass budgets are planned
    // (fewest passes kept). Both are correct and tested, and both measured
    // slower on the 30-qubit HEA (the greedy planner's passes came out
    // heavier: 124-141 ms against 119.5 ms per circuit), so they are off by
    // default (LQ_PRODUCT_FOLD=1, LQ_PLAN_SEARCH=1).
    // > 0: budget of a specialised program's first pass when that pass forms
    // a deferred state (LQ_FIRST_BUDGET)
    double first_budget = -1.0;
    std::atomic<int> product_fold{0};
    std::atomic<int> plan_search{0};
    // pass budget of specialised programs on three or four states (the
    // adjoint sweep of several observables; LQ_MULTI_BUDGET) and on two
    // (LQ_PAIR_BUDGET); <= 0: 26 (ns + 1) and 78
    double multi_budget = -1.0;
    double pair_budget = -1.0;
    // observables per adjoint sweep: 0 one; 1 groups of up to three when
    // their specialised kernels are ready and the estimated state traffic
    // is lower than one sweep per observable; 2 groups whenever ready
    std::atomic<int> adjoint_group{1};
    bool adjoint_debug = false; // LQ_ADJOINT_DEBUG: report each choice on stderr
};
EngineTuning &tuning();

class StateVector {
  public:
    StateVector(int n, int dtype);
    ~StateVector();
    StateVector(const StateVector &) = delete;
    StateVector &operator=(const StateVector &) = delete;

    int n() const { return n_; }
    int dtype() const { return dtype_; } // 0: complex64, 1: complex128
    size_t elem() const { return dtype_ ? 16 : 8; }
    uint64_t length() const { return 1ull << n_; }
    size_t bytes() const { return elem() << n_; }
    // device state; a deferred reset is written first
    void *data() {
        materialize();
        return d_;
    }
    hipStream_t stream() const { return stream_; }
    // Run this state's work on `s` from now on (nullptr: its own stream),
    // after waiting for the work queued so far. A server binds the states
    // it takes from its pool to the worker's stream, so its streams stay as
    // few as its workers (each one a hardware queue of its own).
    void bind_stream(hipStream_t s);

    // state preparation (lq/StateVectorLQubit.hpp:928-1036)
    void reset();
    void set_basis_index(uint64_t index);
    void set_basis_state(const uint32_t *bits, const uint32_t *wires, int nw);
    void set_state_vector(const void *host, const uint32_t *wires, int nw);
    void update_data(const void *host, uint64_t len);
    void get_data(void *host, uint64_t len);
    // the same from / into device memory of the caller (len amplitudes)
    void update_data_device(const void *src, uint64_t len);
    void get_data_device(void *dst, uint64_t len);
    void copy_from(StateVector &src);

    // gates (lq/StateVectorLQubit.hpp:378-710)
    void apply_named(const std::string &name, const uint32_t *wires, int nw, const uint32_t *ctrls,
                     const uint8_t *cvals, int nc, bool inverse, const double *params, int np);
    void apply_matrix(const void *mat, const uint32_t *wires, int nw, const uint32_t *ctrls, const uint8_t *cvals,
                      int nc, bool inverse);
    void apply_pauli_rot(const uint32_t *wires, int nw, bool inverse, double theta, const std::string &word);
    double apply_generator(const std::string &name, const uint32_t *wires, int nw, const uint32_t *ctrls,
                           const uint8_t *cvals, int nc);
    void apply_oprec(const OpRec &op, bool adj);
    // out = M in for a dense operator on physical bits tb (any width; out of
    // place, in != out); controls (cm, cv) keep other amplitudes unchanged
    void apply_wide(std::vector<cd> m, const std::vector<int> &tb, uint64_t cm, uint64_t cv, const void *in,
                    void *out);
    void normalize();
    void collapse(uint32_t wire, bool branch);

    // measurements (lq/measurements/MeasurementsLQubit.hpp)
    std::vector<double> probs(const uint32_t *wires, int nw, bool all);
    double expval_named(const std::string &name, const uint32_t *wires, int nw);
    std::vector<double> expval_named_batch(const std::vector<std::string> &names, const std::vector<uint32_t> &wires);
    double expval_matrix(const void *mat, const uint32_t *wires, int nw);
    double expval_obs(const Obs &o);
    double norm2();
    // variances (MeasurementsLQubit.hpp:420-536, 604-636): <O psi|O psi> - Re<psi|O psi>^2
    double var_named(const std::string &name, const uint32_t *wires, int nw);
    double var_matrix(const void *mat, const uint32_t *wires, int nw);
    double var_obs(const Obs &o);
    // sparse Hamiltonian over the state index (MeasurementsLQubit.hpp:323-337, 616-636)
    double expval_sparse(const CsrView &h);
    double var_sparse(const CsrView &h);
    // samples of `wires` (all wires when all) by the alias method over the
    // marginal probabilities, drawn with std::mt19937 seeded by `seed` when
    // seeded, else from std::random_device (MeasurementsLQubit.hpp:646-677,
    // MeasurementKernels.hpp:308-381); out is shots x nw, wires[0] first
    void generate_samples(const uint32_t *wires, int nw, bool all, uint64_t shots, bool seeded, uint64_t seed,
                          uint64_t *out);

    // queue
    bool has_queued() const { return !queue_.empty(); }
    void enqueue(LOp op);
    void flush();
    void synchronize();

    // Run an op program over `states` (device pointers of this size/dtype),
    // return the dot slot values (complex) indexed by slot id.
    std::vector<cd> run(const std::vector<LOp> &prog, const std::vector<void *> &states, int nslots);
    // The same in two steps, so a program can be planned on the host while
    // earlier work runs on the GPU: plan for `ns` states, then execute.
    struct Planned {
        std::vector<PassPlan> passes;
        std::vector<hipFunction_t> fns; // specialised kernel per pass, or null
        std::vector<const std::vector<uint64_t> *> geo; // its bit-position table (null or empty: none)
        int ns = 0;
        bool all_specialised() const {
            for (hipFunction_t f : fns)
                if (!f) return false;
            return !fns.empty();
        }
    };
    // forms: the program starts from a deferred state its first pass forms
    Planned plan(const std::vector<LOp> &prog, int ns, bool forms = false);
    std::vector<cd> execute(const Planned &pl, const std::vector<void *> &states, int nslots);

    // A recorded circuit: |0...0> followed by `ops`, captured once on this
    // state's stream as a HIP graph (program upload, the state's preparation,
    // every pass). A replay writes the pass coefficients of new parameters
    // into the graph's pinned staging buffer and launches the graph; when the
    // plan of the new parameters differs (kernels compiled since, a rotation
    // that changes form) the circuit is captured again.
    struct Recording {
        StateVector *owner = nullptr;
        std::vector<OpRec> ops;
        size_t nparams = 0;
        Planned planned;
        size_t coef_at = 0;
        std::vector<size_t> coef_off;
        HostPinned staging;
        DevBuf prog;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        uint64_t captures = 0, replays = 0;
        Recording() = default;
        Recording(const Recording &) = delete;
        Recording &operator=(const Recording &) = delete;
        ~Recording();
    };
    std::unique_ptr<Recording> record(const OpsList &ops);
    void replay(Recording &r, const double *params, size_t np);

    // out = O in (observable applied, Hamiltonians summed); in/out device states
    void apply_obs_to(const Obs &o, const void *in, void *out);
    cd inner(const void *bra, const void *ket);

    // Per-pass timing with HIP events on this handle's stream: when enabled,
    // every tile-pass launch is bracketed by events; read() totals them.
    struct PassStats {
        uint64_t passes = 0;
        double ms = 0.0;
        double bytes = 0.0; // algorithmic HBM bytes (amplitudes read + written)
        double max_ms = 0.0;
    };
    // <this|other> for states of the same size in either precision (FP64,
    // compensated sum of per-block partials)
    cd overlap(StateVector &other);

    void profile(bool on);
    PassStats profile_read(bool reset);
    // one record per profiled pass since the last profile_read(reset=true)
    struct PassRecord {
        double ms = 0.0, bytes = 0.0;
        uint32_t ops = 0, stages = 0, transposes = 0, specialised = 0;
        uint64_t tile_bits = 0; // physical bits of the pass's tile
    };
    std::vector<PassRecord> profile_passes();
    // passes launched as specialised kernels / generic tile kernels
    void pass_counts(bool reset, uint64_t *specialised, uint64_t *generic);

    // workspace states (kept between calls; released with release_workspace)
    void *workspace(int i);
    void release_workspace();

  private:
    int n_, dtype_;
    void *d_ = nullptr;
    // preparation deferred: the state is the product state with amplitude
    // prod_q pend_[2q + bit q of the index] (a basis state after a reset, with
    // the leading one-qubit gates of the next program folded in) but not yet
    // written; the first specialised pass of the next program forms it, any
    // other use writes it first (materialize)
    bool prep_pending_ = false;
    std::vector<cd> pend_;
    void materialize();
    Recording *capturing_ = nullptr; // execute() records into this recording's buffers
    std::vector<LOp> lower_recorded(const Recording &r);
    void capture(Recording &r, Planned pl);
    hipStream_t stream_ = nullptr;
    hipStream_t own_stream_ = nullptr; // created with the state; stream_ may be a caller's
    std::vector<LOp> queue_;
    DevBuf prog_, partials_, slots_, small_;
    HostPinned staging_;
    HostPinned readback_; // dot results, written by a copy kernel (kernel_copies)
    hipEvent_t staged_ = nullptr;
    hipEvent_t copied_ = nullptr; // a copy out of another state finished reading it
    hipEvent_t bounced_ = nullptr; // the last upload through bounce_ finished reading it
    HostPinned bounce_;            // small uploads (kernel_copies)
    // Small host<->device transfers on this state's stream: through a copy
    // kernel and fine-grained host memory when kernel copies are on (an
    // upload may write up to dev_cap bytes, rounded up to 16), else
    // hipMemcpyAsync. read_back returns once the data is in `host`.
    void upload_small(void *dev, size_t dev_cap, const void *host, size_t bytes);
    void read_back(void *host, const void *dev, size_t bytes);
    std::vector<std::unique_ptr<DevBuf>> ws_;
    bool prof_ = false;
    struct Pending {
        hipEvent_t a, b;
        double bytes;
        PassRecord rec;
    };
    std::vector<PassRecord> pass_log_;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> ev_free_;
    PassStats stats_;
    uint64_t jit_passes_ = 0, generic_passes_ = 0;
    hipEvent_t take_event();

    int kt_for(int ns) const;
    std::vector<double> reduce_rows_to_host(uint32_t rows, uint32_t cols);
    void apply_obs_inplace(const Obs &o, void *state, std::vector<LOp> &prog);
    // Re<psi|O psi> and <O psi|O psi> of a dense k-wire operator, one read pass
    void dense_moments(std::vector<cd> m, const uint32_t *wires, int nw, double *mean, double *mean_square);
    // sum conj(psi) H psi and sum |H psi|^2 of a CSR operator, one read pass
    void sparse_moments(const CsrView &h, double *mean, double *mean_square);
    // CSR operator on the device: offsets, then column indices, then values
    DevBuf sparse_;
    DevBuf wide_; // matrix of the last operator wider than a tile op
    DevBuf mfma_w_; // real 64x64 form of the last matrix-core operator
    bool apply_dense5_mfma(const std::vector<cd> &m, const std::vector<int> &tb, uint64_t cm);
    size_t csr_indices_off_ = 0, csr_values_off_ = 0;
    std::vector<char> sparse_host_;
    CsrView sparse_view(const Obs &o);
    void upload_csr(const CsrView &h);
};

// plan_passes through a bounded process-wide cache keyed by program
// structure (plan_structure_key); a hit refills the cached plan's
// coefficients from `prog`
std::vector<PassPlan> plan_cached(const std::vector<LOp> &prog, const PlanConfig &cfg);

// Lowered op of an OpsList entry: U, or U^dagger with `dagger`
// (core/algorithms/AdjointJacobianBase.hpp:108-127).
LOp lower_record(int n, const OpRec &op, bool dagger);
// Append the backward sweep on states {lambda, b_1..b_m} to prog: for each
// trainable op from the last, dots <b_j | G lambda> into slot
// slot0 + t*m + j, then U^dagger on all states. Returns the generator scales
// (negated for inverse ops) in sorted-trainable order.
std::vector<double> backward_sweep(int n, const OpsList &ops, const std::vector<uint64_t> &tp, size_t m, int slot0,
                                   std::vector<LOp> &prog);

// Backward sweeps run since the last reset (process-wide): shared by a
// group of observables, and for one observable.
void adjoint_sweep_counts(bool reset, uint64_t *grouped, uint64_t *single);
void adjoint_jacobian(StateVector &sv, const OpsList &ops, const std::vector<std::shared_ptr<Obs>> &obs,
                      const std::vector<uint64_t> &tp, bool apply_operations, double *jac, double *expvals);
// w_t = <dy | d psi / d theta_t> for the trainable parameters tp
// (lq/algorithms/VectorJacobianProduct.hpp:43-160); dy is a host array of
// len amplitudes in the state's precision.
void vector_jacobian_product(StateVector &sv, const OpsList &ops, const void *dy_host, uint64_t len,
                             const std::vector<uint64_t> &tp, bool apply_operations, cd *out);

} // namespace lqamd
