// Device state vector and executor (host side).
#include "lq_engine.hpp"
#include "lq_batch.hpp"
#include "lq_jit.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <random>
#include <set>
#include <mutex>
#include <unordered_map>

namespace lqamd {

void abort_with(const char *msg, const char *file, int line, const char *func) {
    std::string s = std::string("[") + file + "][Line:" + std::to_string(line) + "][Method:" + func +
                    "]: Error in PennyLane Lightning: " + msg;
    throw LightningException(s);
}

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        const std::string m = std::string(what) + ": " + hipGetErrorString(e);
        LQ_ABORT(m.c_str());
    }
}

DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
    if (p) (void)hipFree(p);
    for (void *o : outgrown_) (void)hipFree(o);
    outgrown_.clear();
    p = nullptr;
    bytes = 0;
}
// Buffers grow by at least half their size (rounded to 64 KiB) so that a
// run of slightly larger programs reallocates them a few times, not every
// time; the outgrown ones below kKeepOutgrown are kept (at most twice the
// final size in all), because hipFree and hipHostFree wait for the whole
// device: a server's worker freeing one waited for every request in flight
// (up to ~1 s) while the requests queued behind it waited too.
static size_t grown_size(size_t have, size_t need) {
    const size_t g = std::max(need, have + have / 2);
    return (g + 65535) & ~static_cast<size_t>(65535);
}

void DevBuf::ensure(size_t b) {
    if (b <= bytes) return;
    const size_t nb = grown_size(bytes, b);
    if (p && bytes <= kKeepOutgrown) outgrown_.push_back(p);
    else if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hip_check(hipMalloc(&p, nb), "hipMalloc");
    bytes = nb;
}

HostPinned::~HostPinned() {
    if (p) (void)hipHostFree(p);
    for (void *o : outgrown_) (void)hipHostFree(o);
}
void HostPinned::ensure(size_t b) {
    if (b <= bytes) return;
    const size_t nb = grown_size(bytes, b);
    if (p && bytes <= kKeepOutgrown) outgrown_.push_back(p);
    else if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hip_check(hipHostMalloc(&p, nb, flags), "hipHostMalloc");
    bytes = nb;
}

EngineTuning &tuning() {
    static EngineTuning &t = *[] {
        auto *p = new EngineTuning; // lives until exit
        EngineTuning &e = *p;
        if (const char *s = std::getenv("LQ_TILE_BITS")) e.kt_max = std::atoi(s);
        if (const char *s = std::getenv("LQ_LOW_BITS")) e.low_bits = std::atoi(s);
        if (const char *s = std::getenv("LQ_PASS_BUDGET")) e.budget = std::atof(s);
        if (const char *s = std::getenv("LQ_LDS_BYTES")) e.lds_bytes = static_cast<size_t>(std::atol(s));
        if (const char *s = std::getenv("LQ_REG")) e.reg = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_PLAN_DEBUG")) e.plan_debug = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_PLAN_CACHE")) e.plan_cache = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_PLAN_LOOKAHEAD")) e.lookahead = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_SPLIT_FALLBACK")) e.split_fallback.store(std::atoi(s));
        if (const char *s = std::getenv("LQ_MFMA_DENSE5")) e.mfma_dense5.store(std::atoi(s));
        if (const char *s = std::getenv("LQ_PRODUCT_FOLD")) e.product_fold.store(std::atoi(s) != 0);
        if (const char *s = std::getenv("LQ_FIRST_BUDGET")) e.first_budget = std::atof(s);
        if (const char *s = std::getenv("LQ_PLAN_SEARCH")) e.plan_search.store(std::atoi(s) != 0);
        if (const char *s = std::getenv("LQ_MULTI_BUDGET")) e.multi_budget = std::atof(s);
        if (const char *s = std::getenv("LQ_PAIR_BUDGET")) e.pair_budget = std::atof(s);
        if (const char *s = std::getenv("LQ_ADJOINT_GROUP")) e.adjoint_group.store(std::atoi(s));
        if (const char *s = std::getenv("LQ_ADJOINT_DEBUG")) e.adjoint_debug = std::atoi(s) != 0;
        return p;
    }();
    return t;
}

namespace {
template <typename F> void dispatch(int dtype, F &&f) {
    if (dtype) f(double{});
    else f(float{});
}

void check_wires(int n, const uint32_t *wires, int nw, const uint32_t *ctrls, int nc) {
    uint64_t seen = 0;
    for (int i = 0; i < nw + nc; ++i) {
        const uint32_t w = i < nw ? wires[i] : ctrls[i - nw];
        LQ_ABORT_IF(static_cast<int>(w) >= n, "wire index out of range");
        LQ_ABORT_IF((seen >> w) & 1ull, "`controlled_wires` and `target wires` must be disjoint.");
        seen |= 1ull << w;
    }
}

const std::vector<cd> &named_obs_matrix(const std::string &name) {
    static const std::vector<cd> X{0, 1, 1, 0}, Y{0, cd(0, -1), cd(0, 1), 0}, Z{1, 0, 0, -1},
        H{M_SQRT1_2, M_SQRT1_2, M_SQRT1_2, -M_SQRT1_2}, I{1, 0, 0, 1};
    if (name == "PauliX") return X;
    if (name == "PauliY") return Y;
    if (name == "PauliZ") return Z;
    if (name == "Hadamard") return H;
    if (name == "Identity") return I;
    LQ_ABORT((std::string("Expval does not exist for named observable ") + name).c_str());
}

LOp make_dot(LOp op, uint32_t sa, uint32_t sb, int slot) {
    op.flags = (op.flags & ~TF_ZERO_UNCTRL) | TF_DOT;
    op.sa = sa;
    op.sb = sb;
    op.slot = slot;
    return op;
}

LOp identity_dot(uint32_t sa, uint32_t sb, int slot) {
    LOp op;
    op.kind = TOP_DIAG;
    op.k = 0;
    op.coef = {1.0};
    return make_dot(op, sa, sb, slot);
}

// Pauli word as the cheapest op form for a dot: a parity op for Z-only
// words, a dense 2x2 / 4x4 (or diagonal) op for words on one or two qubits,
// the general Pauli-string op otherwise.
LOp pauli_dot_op(int n, const PauliWord &w) {
    (void)n;
    const uint64_t sup = w.x | w.yz;
    if (w.x == 0 && sup != 0) {
        LOp op;
        op.kind = TOP_PARITY;
        op.pmask = w.yz;
        op.coef = {1.0, -1.0};
        return op;
    }
    const int k = __builtin_popcountll(sup);
    if (k >= 1 && k <= 2) {
        std::vector<int> tb;
        std::vector<std::vector<cd>> fac;
        for (int b = 63; b >= 0; --b) {
            if (!((sup >> b) & 1ull)) continue;
            tb.push_back(b); // descending bit = wires ascending: tb[0] is the local MSB
            const bool x = (w.x >> b) & 1ull, z = (w.yz >> b) & 1ull;
            fac.push_back(named_obs_matrix(x && z ? "PauliY" : x ? "PauliX" : "PauliZ"));
        }
        std::vector<cd> m = fac[0];
        if (k == 2) {
            m.assign(16, 0.0);
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) m[r * 4 + c] = fac[0][(r >> 1) * 2 + (c >> 1)] * fac[1][(r & 1) * 2 + (c & 1)];
        }
        return lower_matrix(std::move(m), tb, 0, 0, true);
    }
    LOp op;
    op.kind = TOP_PAULI;
    op.xmask = w.x;
    op.yzmask = w.yz;
    op.ny = w.ny;
    cd ph = 1.0;
    for (int i = 0; i < (w.ny & 3); ++i) ph *= cd(0, -1);
    op.coef = {ph};
    return op;
}
} // namespace

std::vector<uint32_t> Obs::all_wires() const {
    std::set<uint32_t> s;
    if (kind == Named || kind == Hermitian || kind == Sparse) s.insert(wires.begin(), wires.end());
    for (const auto &t : terms)
        for (auto w : t->all_wires()) s.insert(w);
    return {s.begin(), s.end()};
}

bool obs_to_pauli(const Obs &o, int n, std::vector<PauliWord> *out) {
    auto check = [n](uint32_t w) { LQ_ABORT_IF(static_cast<int>(w) >= n, "observable wire out of range"); };
    switch (o.kind) {
    case Obs::Named: {
        check(o.wires.at(0));
        const uint64_t b = 1ull << phys_bit(n, o.wires[0]);
        if (o.name == "Identity") out->push_back({1.0, 0, 0, 0});
        else if (o.name == "PauliX") out->push_back({1.0, b, 0, 0});
        else if (o.name == "PauliY") out->push_back({1.0, b, b, 1});
        else if (o.name == "PauliZ") out->push_back({1.0, 0, b, 0});
        else if (o.name == "Hadamard") {
            out->push_back({M_SQRT1_2, b, 0, 0});
            out->push_back({M_SQRT1_2, 0, b, 0});
        } else return false;
        return true;
    }
    case Obs::Hermitian:
    case Obs::Sparse: return false;
    case Obs::Tensor: {
        std::vector<PauliWord> acc{{1.0, 0, 0, 0}};
        for (const auto &f : o.terms) {
            std::vector<PauliWord> fw;
            if (!obs_to_pauli(*f, n, &fw)) return false;
            std::vector<PauliWord> next;
            for (const auto &a : acc)
                for (const auto &b : fw) {
                    LQ_ABORT_IF((a.x | a.yz) & (b.x | b.yz), "All wires in observables must be disjoint.");
                    next.push_back({a.coeff * b.coeff, a.x | b.x, a.yz | b.yz, a.ny + b.ny});
                }
            acc.swap(next);
        }
        out->insert(out->end(), acc.begin(), acc.end());
        return true;
    }
    case Obs::Ham: {
        std::vector<PauliWord> all;
        for (size_t i = 0; i < o.terms.size(); ++i) {
            std::vector<PauliWord> tw;
            if (!obs_to_pauli(*o.terms[i], n, &tw)) return false;
            for (auto &w : tw) w.coeff *= o.coeffs[i];
            all.insert(all.end(), tw.begin(), tw.end());
        }
        out->insert(out->end(), all.begin(), all.end());
        return true;
    }
    }
    return false;
}

// ---------------------------------------------------------------------------
namespace {
// Program uploads and dot-result readbacks of execute() go through copy
// kernels on the state's stream and fine-grained host memory instead of the
// copy engines (LQ_KERNEL_COPIES=0: hipMemcpyAsync): the copy engines'
// queues are shared by every stream's transfers, so with concurrent states
// (a request server's workers) a transfer waited behind other streams'
// transfers, which wait for those streams' kernels.
bool kernel_copies() {
    static const bool v = [] {
        const char *e = std::getenv("LQ_KERNEL_COPIES");
        return !e || std::atoi(e) != 0;
    }();
    return v;
}

// Device address of fine-grained host memory, or nullptr.
void *host_dev_ptr(void *h) {
    void *d = nullptr;
    return hipHostGetDevicePointer(&d, h, 0) == hipSuccess ? d : nullptr;
}
} // namespace

StateVector::StateVector(int n, int dtype) : n_(n), dtype_(dtype ? 1 : 0) {
    LQ_ABORT_IF(n < 1 || n > 40, "number of qubits must be in [1, 40]");
    if (kernel_copies()) staging_.flags = readback_.flags = bounce_.flags = hipHostMallocCoherent;
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    own_stream_ = stream_;
    hip_check(hipEventCreateWithFlags(&staged_, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&copied_, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&bounced_, hipEventDisableTiming), "hipEventCreate");
    hipError_t e = hipMalloc(&d_, bytes());
    if (e != hipSuccess) {
        d_ = nullptr;
        (void)hipEventDestroy(bounced_);
        (void)hipEventDestroy(copied_);
        (void)hipEventDestroy(staged_);
        (void)hipStreamDestroy(stream_);
        own_stream_ = stream_ = nullptr;
        hip_check(e, "hipMalloc(state vector)");
    }
    reset();
}

void StateVector::upload_small(void *dev, size_t dev_cap, const void *host, size_t bytes) {
    if (bytes == 0) return;
    const size_t r = (bytes + 15) & ~static_cast<size_t>(15);
    void *bd = nullptr;
    if (kernel_copies() && r <= dev_cap) {
        hip_check(hipEventSynchronize(bounced_), "hipEventSynchronize"); // bounce_ is free again
        bounce_.ensure(r);
        bd = host_dev_ptr(bounce_.p);
    }
    if (bd) {
        std::memcpy(bounce_.p, host, bytes);
        launch_copy16(bd, dev, r, stream_);
        hip_check(hipGetLastError(), "upload launch");
        hip_check(hipEventRecord(bounced_, stream_), "hipEventRecord");
    } else {
        hip_check(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
    }
}

void StateVector::read_back(void *host, const void *dev, size_t bytes) {
    if (bytes == 0) return;
    void *rd = nullptr;
    if (kernel_copies() && bytes % 16 == 0) {
        readback_.ensure(bytes);
        rd = host_dev_ptr(readback_.p);
    }
    if (rd) {
        launch_copy16(dev, rd, bytes, stream_);
        hip_check(hipGetLastError(), "readback launch");
        hip_check(hipStreamSynchronize(stream_), "sync");
        std::memcpy(host, readback_.p, bytes);
    } else {
        hip_check(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream_), "memcpy");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }
}

StateVector::~StateVector() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto &p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ev_free_) (void)hipEventDestroy(e);
    ws_.clear();
    if (d_) (void)hipFree(d_);
    if (staged_) (void)hipEventDestroy(staged_);
    if (copied_) (void)hipEventDestroy(copied_);
    if (bounced_) (void)hipEventDestroy(bounced_);
    if (own_stream_) {
        (void)hipStreamSynchronize(own_stream_);
        (void)hipStreamDestroy(own_stream_);
    }
}

void StateVector::bind_stream(hipStream_t s) {
    flush();
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    stream_ = s ? s : own_stream_;
}

hipEvent_t StateVector::take_event() {
    if (!ev_free_.empty()) {
        hipEvent_t e = ev_free_.back();
        ev_free_.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

void StateVector::profile(bool on) {
    synchronize();
    prof_ = on;
}

void StateVector::pass_counts(bool reset, uint64_t *specialised, uint64_t *generic) {
    if (specialised) *specialised = jit_passes_;
    if (generic) *generic = generic_passes_;
    if (reset) jit_passes_ = generic_passes_ = 0;
}

StateVector::PassStats StateVector::profile_read(bool reset) {
    synchronize();
    for (auto &p : pending_) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        stats_.passes += 1;
        stats_.ms += ms;
        stats_.bytes += p.bytes;
        stats_.max_ms = std::max(stats_.max_ms, static_cast<double>(ms));
        p.rec.ms = ms;
        p.rec.bytes = p.bytes;
        pass_log_.push_back(p.rec);
        ev_free_.push_back(p.a);
        ev_free_.push_back(p.b);
    }
    pending_.clear();
    PassStats out = stats_;
    if (reset) {
        stats_ = PassStats{};
        pass_log_.clear();
    }
    return out;
}

std::vector<StateVector::PassRecord> StateVector::profile_passes() {
    profile_read(false);
    return pass_log_;
}

void *StateVector::workspace(int i) {
    while (static_cast<int>(ws_.size()) <= i) ws_.push_back(std::make_unique<DevBuf>());
    ws_[i]->ensure(bytes());
    return ws_[i]->p;
}

void StateVector::release_workspace() {
    synchronize();
    ws_.clear();
}

namespace {
// Dynamic LDS the tile kernel may request: the tuning budget, within the
// device's per-workgroup limit less the kernel's static reduction buffer.
size_t lds_budget() {
    static const size_t v = [] {
        int dev = 0, lim = 0;
        size_t b = tuning().lds_bytes;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lim > 0)
            b = std::min(b, static_cast<size_t>(lim) - 4096);
        return b;
    }();
    return v;
}
} // namespace

int StateVector::kt_for(int ns) const {
    const EngineTuning &t = tuning();
    const size_t budget = lds_budget();
    int kt = 0;
    while ((elem() * static_cast<size_t>(ns) << (kt + 1)) <= budget) ++kt;
    if (t.kt_max > 0) kt = std::min(kt, t.kt_max);
    return std::min(kt, 13);
}

void StateVector::reset() { set_basis_index(0); }

void StateVector::set_basis_index(uint64_t index) {
    LQ_ABORT_IF(index > length() - 1, "Invalid index");
    queue_.clear();
    // deferred: the next program's first specialised pass forms the state,
    // any other use writes it first
    pend_.assign(2 * static_cast<size_t>(n_), cd(0.0, 0.0));
    for (int q = 0; q < n_; ++q) pend_[2 * static_cast<size_t>(q) + ((index >> q) & 1ull)] = 1.0;
    prep_pending_ = true;
}

void StateVector::materialize() {
    if (!prep_pending_) return;
    prep_pending_ = false;
    ProductFactors f{};
    for (size_t i = 0; i < pend_.size(); ++i) f.f[i] = make_double2(pend_[i].real(), pend_[i].imag());
    dispatch(dtype_, [&](auto t) { launch_fill_product<decltype(t)>(d_, n_, f, stream_); });
}

void StateVector::set_basis_state(const uint32_t *bits, const uint32_t *wires, int nw) {
    uint64_t index = 0;
    for (int k = 0; k < nw; ++k) {
        LQ_ABORT_IF(static_cast<int>(wires[k]) >= n_, "wire index out of range");
        index |= static_cast<uint64_t>(bits[k] != 0) << (n_ - 1 - static_cast<int>(wires[k]));
    }
    set_basis_index(index);
}

void StateVector::set_state_vector(const void *host, const uint32_t *wires, int nw) {
    check_wires(n_, wires, nw, nullptr, 0);
    flush();
    const size_t b = elem() << nw;
    small_.ensure(b + 256 * sizeof(int));
    hip_check(hipMemcpyAsync(small_.p, host, b, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
    std::vector<int> bits(nw);
    for (int j = 0; j < nw; ++j) bits[j] = phys_bit(n_, wires[j]);
    dispatch(dtype_, [&](auto t) { launch_set_subspace<decltype(t)>(data(), n_, small_.p, bits.data(), nw, stream_); });
    synchronize();
}

void StateVector::update_data(const void *host, uint64_t len) {
    LQ_ABORT_IF(len != length(), "The size of provided data must be equal to the statevector size.");
    queue_.clear();
    prep_pending_ = false; // every amplitude is overwritten
    hip_check(hipMemcpyAsync(d_, host, bytes(), hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
    synchronize();
}

void StateVector::get_data(void *host, uint64_t len) {
    LQ_ABORT_IF(len != length(), "The size of provided data must be equal to the statevector size.");
    flush();
    hip_check(hipMemcpyAsync(host, data(), bytes(), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
    synchronize();
}

// device-to-device copies between the state and caller memory (a torch
// tensor's storage, another library's buffer): the streaming copy kernel
// when both addresses are 16-byte aligned, the runtime's copy otherwise
static void device_copy(const void *src, void *dst, size_t bytes, hipStream_t s) {
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u) == 0)
        launch_stream_copy(src, dst, bytes, s);
    else
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
}

void StateVector::update_data_device(const void *src, uint64_t len) {
    LQ_ABORT_IF(len != length(), "The size of provided data must be equal to the statevector size.");
    LQ_ABORT_IF(src == nullptr, "null device address");
    queue_.clear();
    prep_pending_ = false; // every amplitude is overwritten
    device_copy(src, d_, bytes(), stream_);
    synchronize();
}

void StateVector::get_data_device(void *dst, uint64_t len) {
    LQ_ABORT_IF(len != length(), "The size of provided data must be equal to the statevector size.");
    LQ_ABORT_IF(dst == nullptr, "null device address");
    flush();
    device_copy(data(), dst, bytes(), stream_);
    synchronize();
}

void StateVector::copy_from(StateVector &src) {
    LQ_ABORT_IF(src.n_ != n_ || src.dtype_ != dtype_, "state vectors differ in size or precision");
    if (&src == this) {
        flush();
        return;
    }
    src.flush();
    // data() may enqueue the write of a deferred reset on src's stream: take
    // it before waiting for that stream
    const void *sp = src.data();
    src.synchronize();
    queue_.clear();
    prep_pending_ = false; // every amplitude is overwritten
    launch_stream_copy(sp, d_, bytes(), stream_);
    // src's next write (a gate program, updateData, a deferred reset) runs on
    // src's stream: it waits until this copy has read the amplitudes
    hip_check(hipEventRecord(copied_, stream_), "hipEventRecord");
    hip_check(hipStreamWaitEvent(src.stream_, copied_, 0), "hipStreamWaitEvent");
}

void StateVector::enqueue(LOp op) {
    if (try_merge_1q(queue_, op)) return;
    queue_.push_back(std::move(op));
    if (queue_.size() >= 8192) flush();
}

void StateVector::flush() {
    if (queue_.empty()) return;
    std::vector<LOp> q;
    q.swap(queue_);
    run(q, {d_}, 0);
}

void StateVector::synchronize() {
    flush();
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

namespace {
bool better_plan(const std::vector<PassPlan> &a, const std::vector<PassPlan> &b);
}

std::vector<cd> StateVector::run(const std::vector<LOp> &prog, const std::vector<void *> &states, int nslots) {
    if (prog.empty()) return std::vector<cd>(static_cast<size_t>(nslots), 0.0);
    if (prep_pending_ && states.size() == 1 && states[0] == d_ && tuning().product_fold.load()) {
        // the state is a product state not yet written: its leading
        // one-qubit gates become factors of that product
        std::vector<LOp> rest = prog;
        std::vector<cd> f = pend_;
        if (absorb_into_product(rest, f) > 0) {
            if (rest.empty()) {
                pend_.swap(f);
                return std::vector<cd>(static_cast<size_t>(nslots), 0.0);
            }
            Planned folded = plan(rest, 1, true);
            if (n_ >= kPlanSearchQubits && tuning().plan_search.load()) {
                // both forms let the first pass form the state without reading
                // it; folding removes ops but can cost the greedy planner a pass
                Planned whole = plan(prog, 1, true);
                if (better_plan(whole.passes, folded.passes)) return execute(whole, states, nslots);
            }
            pend_.swap(f);
            return execute(folded, states, nslots);
        }
    }
    const bool forms = prep_pending_ && states.size() == 1 && states[0] == d_;
    return execute(plan(prog, static_cast<int>(states.size()), forms), states, nslots);
}

namespace {
// Plans by program structure (process-wide, bounded): a program seen before
// with other coefficients reuses its plan with the coefficients refilled.
struct PlanCache {
    std::mutex mu;
    std::unordered_map<std::string, std::vector<PassPlan>> map;
    std::deque<std::string> order;
    // a serving stream of 10..28-qubit requests at depths 4..16 has ~300
    // distinct forward and backward program structures
    static constexpr size_t kMaxEntries = 512;
    static PlanCache &get() {
        static PlanCache c;
        return c;
    }
};
} // namespace

std::vector<PassPlan> plan_cached(const std::vector<LOp> &prog, const PlanConfig &cfg) {
    if (!tuning().plan_cache) return plan_passes(prog, cfg);
    std::string key = plan_structure_key(prog, cfg);
    PlanCache &c = PlanCache::get();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.map.find(key);
        if (it != c.map.end()) {
            std::vector<PassPlan> passes = it->second;
            if (refresh_plan_coefs(passes, prog)) return passes;
        }
    }
    std::vector<PassPlan> passes = plan_passes(prog, cfg);
    std::lock_guard<std::mutex> lk(c.mu);
    // a refused refresh replaces the entry: the angles have moved to the new plan's side
    if (c.map.insert_or_assign(key, passes).second) {
        c.order.push_back(std::move(key));
        if (c.order.size() > PlanCache::kMaxEntries) {
            c.map.erase(c.order.front());
            c.order.pop_front();
        }
    }
    return passes;
}

namespace {
// Passes first (each reads and writes the states), then LDS transposes.
bool better_plan(const std::vector<PassPlan> &a, const std::vector<PassPlan> &b) {
    if (a.size() != b.size()) return a.size() < b.size();
    auto transposes = [](const std::vector<PassPlan> &v) {
        size_t t = 0;
        for (const PassPlan &pp : v) {
            t += pp.end_transpose ? 1 : 0;
            for (const TileStage &st : pp.stages) t += st.entry == ENTRY_TRANSPOSE;
        }
        return t;
    };
    return transposes(a) < transposes(b);
}
} // namespace

// The greedy planner's pass count moves irregularly with the per-pass budget
// (30-qubit HEA, depth 20: 17 to 19 passes over budgets 44..64, and which
// budget wins changes when the first layer folds into a product state), so
// for large states, where a pass costs milliseconds, several budgets are
// planned and the plan with the fewest passes kept; cached like single plans
// (under the key of budget -1), so each structure searches once.
std::vector<PassPlan> plan_search(const std::vector<LOp> &prog, const PlanConfig &cfg) {
    static const double kBudgets[] = {52.0, 44.0, 48.0, 56.0, 60.0, 64.0};
    auto search = [&] {
        std::vector<PassPlan> best;
        bool have = false;
        for (double b : kBudgets) {
            PlanConfig c = cfg;
            c.budget = b;
            std::vector<PassPlan> p = plan_passes(prog, c);
            if (!have || better_plan(p, best)) {
                best.swap(p);
                have = true;
            }
        }
        return best;
    };
    if (!tuning().plan_cache) return search();
    PlanConfig kc = cfg;
    kc.budget = -1.0;
    std::string key = plan_structure_key(prog, kc);
    PlanCache &c = PlanCache::get();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.map.find(key);
        if (it != c.map.end()) {
            std::vector<PassPlan> passes = it->second;
            if (refresh_plan_coefs(passes, prog)) return passes;
        }
    }
    std::vector<PassPlan> passes = search();
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.map.insert_or_assign(key, passes).second) {
        c.order.push_back(std::move(key));
        if (c.order.size() > PlanCache::kMaxEntries) {
            c.map.erase(c.order.front());
            c.order.pop_front();
        }
    }
    return passes;
}

StateVector::Planned StateVector::plan(const std::vector<LOp> &prog, int ns, bool forms) {
    LQ_ABORT_IF(ns < 1 || ns > kMaxStates, "too many states in one program");
    Planned out;
    out.ns = ns;
    if (prog.empty()) return out;
    const EngineTuning &t = tuning();
    PlanConfig cfg;
    cfg.n = n_;
    cfg.ns = ns;
    cfg.kt_max = kt_for(ns);
    cfg.low_bits = t.low_bits >= 0 ? t.low_bits : (dtype_ ? 4 : 5);
    cfg.budget = t.budget > 0 ? t.budget : 24.0;
    // register-resident passes for tiles of 11 or 12 bits with one or two states
    const int ktr = std::min(cfg.kt_max, 8 + kMaxRegBits);
    if (t.reg && ns <= 2 && n_ >= 11 && ktr >= 11) {
        cfg.kt_max = ktr;
        cfg.reg_rb = ktr - 8;
        cfg.max_dots = kMaxPassDots;
    }
    // register passes that will be specialised: 2x2 gates as phase / real
    // rotation / phase, so the rotations take real arithmetic and a stage's
    // phases apply as one table. That form suits only the specialised
    // kernels (it triples the ops the generic kernel interprets), so it runs
    // when its kernels are ready; meanwhile their compiles are queued and the
    // unsplit plan runs in the generic kernel.
    const JitConfig &jcfg = jit_config();
    // three or four states: register passes only as specialised kernels
    // (tiles of 11 bits; the generic register kernel holds two states), the
    // generic fallback being LDS passes
    const bool multi = ns > 2 && t.reg && n_ >= 11 && jcfg.lds32;
    const bool use_jit = (cfg.reg_rb > 0 || multi) && jcfg.enabled && n_ >= jcfg.min_qubits;
    std::vector<PassPlan> passes;
    std::vector<hipFunction_t> fns;
    std::vector<const std::vector<uint64_t> *> geo;
    auto lookup_all = [&](std::vector<hipFunction_t> &f) {
        size_t missing = 0;
        f.assign(passes.size(), nullptr);
        geo.assign(passes.size(), nullptr);
        for (size_t p = 0; p < passes.size(); ++p) {
            if (!passes[p].reg) continue;
            f[p] = jit_lookup_plan(passes[p], n_, dtype_, ns, &geo[p]);
            missing += f[p] == nullptr;
        }
        return missing;
    };
    if (use_jit && jcfg.split_rotations) {
        PlanConfig scfg = cfg;
        // measured on the 30-qubit HEA: budget 52 (complex128 18 passes,
        // 122 ms; 48 and 54: 19 passes, 126 and 131 ms; the 24-qubit adjoint
        // and 33-qubit QFT workloads equal or faster), and in complex64 4
        // contiguous low bits as in complex128 (one 128-byte line per 16
        // lanes: 18 passes, 79-80 ms, against 24 passes and 86 ms with 5 low
        // bits and budget 40)
        if (t.budget <= 0) scfg.budget = 52.0;
        if (t.low_bits < 0) scfg.low_bits = 4;
        scfg.scale_rotations = jcfg.scale_rotations;
        // complex64: keep the unscaled amplitudes (at most 1) well inside float range
        scfg.max_scale_growth = dtype_ ? 1e300 : 1e20;
        scfg.lookahead = t.lookahead;
        if (forms && ns == 1 && t.first_budget > 0) scfg.first_budget = t.first_budget;
        // two states (one observable's adjoint sweep, a VJP): budget 78,
        // measured on the 24-qubit / 432-parameter gradient (30 passes and
        // 18.0 ms against 33 and 18.8 ms with 52; 104 the same as 78); three
        // or four (observables sharing a sweep): 26 (ns + 1), since the
        // inverse gates act on every state and the dots per op grow with the
        // observables (three observables at 24 qubits: 130 gives 20 passes
        // and 27.1 ms, 104 24 passes and 33.6 ms, 78 32 passes and 40.1 ms)
        if (t.budget <= 0 && ns == 2) scfg.budget = t.pair_budget > 0 ? t.pair_budget : 78.0;
        if (t.budget <= 0 && ns > 2) scfg.budget = t.multi_budget > 0 ? t.multi_budget : 26.0 * (ns + 1);
        if (multi) {
            scfg.kt_max = 11;
            scfg.reg_rb = 3;
            scfg.max_dots = std::min(kMaxJitPassDots, kMaxPassDots * (ns - 1));
        }
        std::vector<LOp> split = prog;
        split_1q_rotations(split);
        passes = (t.budget <= 0 && t.plan_search.load() && n_ >= kPlanSearchQubits) ? plan_search(split, scfg)
                                                                              : plan_cached(split, scfg);
        const size_t missing = lookup_all(fns);
        const int fallback = t.split_fallback.load();
        if (missing && (fallback != 2 || multi)) {
            // The generic kernel interprets the op list, so its pass time grows
            // with the ops: measured on the 30-qubit HEA passes, about 0.28 of
            // a specialised pass per op of the split plan (phase tables, real
            // rotations) and 0.16 per op of the unsplit plan, while a
            // specialised pass costs 1 whatever its ops.
            std::vector<PassPlan> plain = plan_cached(prog, cfg);
            double split_cost = 0.0, plain_cost = 0.0;
            for (size_t p = 0; p < passes.size(); ++p)
                split_cost += (fns[p] || !passes[p].reg) ? 1.0 : 0.28 * static_cast<double>(passes[p].ops.size());
            for (const PassPlan &pp : plain) plain_cost += std::max(1.0, 0.16 * static_cast<double>(pp.ops.size()));
            if (fallback == 0 || multi || plain_cost < split_cost) {
                passes.swap(plain);
                fns.assign(passes.size(), nullptr);
                geo.assign(passes.size(), nullptr);
            }
        }
    } else {
        passes = plan_cached(prog, cfg);
        if (use_jit) lookup_all(fns);
        else fns.assign(passes.size(), nullptr);
        geo.resize(passes.size(), nullptr);
    }
    if (t.plan_debug) {
        for (size_t p = 0; p < passes.size(); ++p) {
            const PassPlan &pp = passes[p];
            int nreg = 0, nlds = 0, ntr = 0;
            for (const auto &st : pp.stages) {
                nreg += st.kind == STAGE_REG;
                nlds += st.kind == STAGE_LDS;
                ntr += st.kind == STAGE_REG && st.entry == ENTRY_TRANSPOSE;
            }
            std::fprintf(stderr, "[lq plan] pass %zu kt=%zu ops=%zu dots=%zu reg=%d stages=%zu (reg %d, lds %d, transposes %d, end %u)\n",
                         p, pp.tbits.size(), pp.ops.size(), pp.slot_ids.size(), pp.reg ? 1 : 0, pp.stages.size(), nreg,
                         nlds, ntr, pp.end_transpose);
        }
    }
    geo.resize(passes.size(), nullptr);
    out.passes.swap(passes);
    out.fns.swap(fns);
    out.geo.swap(geo);
    return out;
}

std::vector<cd> StateVector::execute(const Planned &pl, const std::vector<void *> &states, int nslots) {
    std::vector<cd> result(static_cast<size_t>(nslots), 0.0);
    const int ns = static_cast<int>(states.size());
    LQ_ABORT_IF(ns != pl.ns, "program planned for another number of states");
    const std::vector<PassPlan> &passes = pl.passes;
    const std::vector<hipFunction_t> &fns = pl.fns;
    if (passes.empty()) return result;

    auto align = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    std::vector<size_t> op_off(passes.size()), coef_off(passes.size()), part_off(passes.size()),
        st_off(passes.size()), geo_slots(passes.size(), 0);
    size_t ops_bytes = 0, coef_n = 0, part_n = 0, st_bytes = 0;
    std::vector<SlotRange> ranges(static_cast<size_t>(nslots), SlotRange{0, 0});
    std::vector<char> seen(static_cast<size_t>(nslots), 0);
    for (size_t p = 0; p < passes.size(); ++p) {
        op_off[p] = ops_bytes;
        ops_bytes += passes[p].ops.size() * sizeof(TileOp);
        st_off[p] = st_bytes;
        st_bytes += passes[p].stages.size() * sizeof(TileStage);
        coef_off[p] = coef_n;
        // + the deferred-preparation marker, the specialised kernel's table of
        // bit positions (two per entry) and the product factors read by
        // specialised one-state passes
        if (fns[p] && p < pl.geo.size() && pl.geo[p]) geo_slots[p] = (pl.geo[p]->size() + 1) / 2;
        coef_n += passes[p].coef.size() + 1 + geo_slots[p] + 2 * static_cast<size_t>(n_);
        part_off[p] = part_n;
        const uint64_t ntiles = 1ull << (n_ - static_cast<int>(passes[p].tbits.size()));
        for (size_t s = 0; s < passes[p].slot_ids.size(); ++s) {
            const int id = passes[p].slot_ids[s];
            LQ_ABORT_IF(id < 0 || id >= nslots || seen[id], "invalid dot slot");
            seen[id] = 1;
            ranges[id] = SlotRange{part_n + s * ntiles, ntiles};
        }
        part_n += passes[p].slot_ids.size() * ntiles;
    }
    const size_t stages_at = align(ops_bytes);
    const size_t coef_at = align(stages_at + st_bytes);
    const size_t ranges_at = align(coef_at + coef_n * sizeof(double2));
    const size_t total = align(ranges_at + ranges.size() * sizeof(SlotRange));

    // A deferred reset of this state is formed by the first pass when that is
    // a specialised one-state pass reading and writing it (a read-only pass
    // would leave the memory unwritten); otherwise it is written now.
    bool form_basis = false;
    if (prep_pending_) {
        bool reads = false;
        for (void *sp : states) reads |= sp == d_;
        if (reads) {
            form_basis = ns == 1 && states[0] == d_ && fns[0] != nullptr && passes[0].reg &&
                         (passes[0].load_mask & 1u) && (passes[0].store_mask & 1u);
            if (!form_basis) materialize();
        }
    }

    // a capture (record) stages into the recording's own buffers, which its
    // graph reads at every replay
    HostPinned &stg = capturing_ ? capturing_->staging : staging_;
    DevBuf &pbuf = capturing_ ? capturing_->prog : prog_;
    LQ_ABORT_IF(capturing_ && part_n != 0, "a recorded circuit carries no dots");
    // the staging buffer may still feed the previous upload
    if (!capturing_) hip_check(hipEventSynchronize(staged_), "hipEventSynchronize");
    stg.ensure(total);
    char *hs = static_cast<char *>(stg.p);
    for (size_t p = 0; p < passes.size(); ++p) {
        std::memcpy(hs + op_off[p], passes[p].ops.data(), passes[p].ops.size() * sizeof(TileOp));
        if (!passes[p].stages.empty())
            std::memcpy(hs + stages_at + st_off[p], passes[p].stages.data(), passes[p].stages.size() * sizeof(TileStage));
        double2 *c = reinterpret_cast<double2 *>(hs + coef_at) + coef_off[p];
        for (size_t i = 0; i < passes[p].coef.size(); ++i)
            c[i] = make_double2(passes[p].coef[i].real(), passes[p].coef[i].imag());
        const bool form = p == 0 && form_basis;
        c[passes[p].coef.size()] = make_double2(form ? 1.0 : -1.0, 0.0);
        if (geo_slots[p]) {
            double2 *g = c + passes[p].coef.size() + 1;
            g[geo_slots[p] - 1] = make_double2(0.0, 0.0);
            std::memcpy(g, pl.geo[p]->data(), pl.geo[p]->size() * sizeof(uint64_t));
        }
        double2 *f = c + passes[p].coef.size() + 1 + geo_slots[p];
        for (size_t i = 0; i < 2 * static_cast<size_t>(n_); ++i)
            f[i] = form ? make_double2(pend_[i].real(), pend_[i].imag()) : make_double2(0.0, 0.0);
    }
    if (form_basis) prep_pending_ = false;
    if (!ranges.empty()) std::memcpy(hs + ranges_at, ranges.data(), ranges.size() * sizeof(SlotRange));
    pbuf.ensure(total);
    // both buffers hold a multiple of 64 KiB (grown_size), so the copy may
    // round up to 16 bytes
    void *hs_dev = !capturing_ && kernel_copies() ? host_dev_ptr(hs) : nullptr;
    if (hs_dev) {
        launch_copy16(hs_dev, pbuf.p, (total + 15) & ~static_cast<size_t>(15), stream_);
        hip_check(hipGetLastError(), "program upload launch");
    } else {
        hip_check(hipMemcpyAsync(pbuf.p, hs, total, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync(program)");
    }
    if (!capturing_) hip_check(hipEventRecord(staged_, stream_), "hipEventRecord");
    if (part_n) partials_.ensure(part_n * sizeof(double2));
    if (capturing_) {
        capturing_->coef_at = coef_at;
        capturing_->coef_off = coef_off;
    }
    const bool prof = prof_ && !capturing_;

    char *dp = static_cast<char *>(pbuf.p);
    for (size_t p = 0; p < passes.size(); ++p) {
        const PassPlan &pp = passes[p];
        TilePassArgs a{};
        for (int s = 0; s < ns; ++s) a.sv[s] = states[s];
        a.ops = reinterpret_cast<const TileOp *>(dp + op_off[p]);
        a.coef = reinterpret_cast<const double2 *>(dp + coef_at) + coef_off[p];
        a.partials = static_cast<double2 *>(partials_.p) + part_off[p];
        a.nops = static_cast<uint32_t>(pp.ops.size());
        a.load_mask = pp.load_mask;
        a.store_mask = pp.store_mask;
        a.ns = static_cast<uint32_t>(ns);
        a.g = make_geom(n_, pp.tbits);
        a.ntiles = 1ull << a.g.nout;
        const size_t lds = static_cast<size_t>(ns) * (elem() << a.g.kt);
        TileRegArgs ra{};
        if (pp.reg) {
            LQ_ABORT_IF(a.g.kt < 11 || a.g.kt > 8 + kMaxRegBits || (ns > 2 && !fns[p]),
                        "register pass geometry out of range");
            LQ_ABORT_IF(pp.slot_ids.size() > static_cast<size_t>(fns[p] && ns > 2 ? kMaxJitPassDots : kMaxPassDots),
                        "too many dots in one pass");
            for (int s = 0; s < ns; ++s) ra.sv[s] = states[s];
            ra.ops = a.ops;
            ra.stages = reinterpret_cast<const TileStage *>(dp + stages_at + st_off[p]);
            ra.coef = a.coef;
            ra.partials = a.partials;
            ra.nstages = static_cast<uint32_t>(pp.stages.size());
            ra.nslots = static_cast<uint32_t>(pp.slot_ids.size());
            ra.load_mask = pp.load_mask;
            ra.store_mask = pp.store_mask;
            ra.end_transpose = pp.end_transpose;
            ra.ns = static_cast<uint32_t>(ns);
            ra.wide = 0;
            for (const TileOp &t : pp.ops) ra.wide |= t.kind == TOP_DENSE2 ? 1u : 0u;
            for (const TileStage &st : pp.stages) ra.wide |= st.kind == STAGE_LDS ? 1u : 0u;
            ra.ntiles = a.ntiles;
            ra.g = a.g;
        }
        Pending pe{};
        if (prof) {
            pe.a = take_event();
            pe.b = take_event();
            // a pass that forms a deferred reset reads nothing of that state
            const int reads = __builtin_popcount(pp.load_mask) - ((p == 0 && form_basis) ? 1 : 0);
            pe.bytes = static_cast<double>(bytes()) * (reads + __builtin_popcount(pp.store_mask));
            pe.rec.ops = static_cast<uint32_t>(pp.ops.size());
            pe.rec.stages = static_cast<uint32_t>(pp.stages.size());
            pe.rec.transposes = pp.end_transpose ? 1u : 0u;
            for (const TileStage &st : pp.stages) pe.rec.transposes += st.entry == ENTRY_TRANSPOSE;
            pe.rec.specialised = fns[p] ? 1u : 0u;
            for (int b : pp.tbits) pe.rec.tile_bits |= 1ull << b;
            hip_check(hipEventRecord(pe.a, stream_), "hipEventRecord");
        }
        if (fns[p])
            jit_launch(fns[p], dtype_, ra.sv[0], ra.sv[1], ra.sv[2], ra.sv[3], ra.coef, ra.partials, ra.ntiles,
                       jit_lds_bytes(a.g.kt, dtype_, ns), stream_);
        else if (pp.reg) dispatch(dtype_, [&](auto tt) { launch_tile_reg<decltype(tt)>(ra, lds, stream_); });
        else dispatch(dtype_, [&](auto tt) { launch_tile_pass<decltype(tt)>(a, lds, stream_); });
        ++(fns[p] ? jit_passes_ : generic_passes_);
        if (prof) {
            hip_check(hipEventRecord(pe.b, stream_), "hipEventRecord");
            pending_.push_back(pe);
        }
    }
    hip_check(hipGetLastError(), "tile pass launch");
    if (nslots > 0) {
        slots_.ensure(static_cast<size_t>(nslots) * sizeof(double2));
        launch_reduce_slots(static_cast<const double2 *>(partials_.p),
                            reinterpret_cast<const SlotRange *>(dp + ranges_at), static_cast<uint32_t>(nslots),
                            static_cast<double2 *>(slots_.p), stream_);
        const size_t rb = static_cast<size_t>(nslots) * sizeof(double2);
        void *rb_dev = nullptr;
        if (kernel_copies()) {
            readback_.ensure(rb);
            rb_dev = host_dev_ptr(readback_.p);
        }
        std::vector<double2> h(static_cast<size_t>(nslots));
        if (rb_dev) {
            launch_copy16(slots_.p, rb_dev, rb, stream_);
            hip_check(hipGetLastError(), "slot readback launch");
            hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
            std::memcpy(h.data(), readback_.p, rb);
        } else {
            hip_check(hipMemcpyAsync(h.data(), slots_.p, rb, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync(slots)");
            hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        }
        for (int i = 0; i < nslots; ++i) result[i] = cd(h[i].x, h[i].y);
    }
    return result;
}

// ---------------------------------------------------------------------------
// Recorded circuits (HIP graphs)

StateVector::Recording::~Recording() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

std::vector<LOp> StateVector::lower_recorded(const Recording &r) {
    queue_.clear();
    for (const OpRec &op : r.ops) {
        if (op.named && op.gid == GateId::Identity) continue;
        apply_oprec(op, false); // checks the op as applyOperation does, then queues it
    }
    std::vector<LOp> prog;
    prog.swap(queue_);
    return prog;
}

void StateVector::capture(Recording &r, Planned pl) {
    if (r.exec) (void)hipGraphExecDestroy(r.exec);
    if (r.graph) (void)hipGraphDestroy(r.graph);
    r.exec = nullptr;
    r.graph = nullptr;
    // the previous launch of this recording may still read its buffers
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    set_basis_index(0);
    // relaxed: the capture sizes the recording's own buffers (hipHostMalloc,
    // hipMalloc), which the stricter capture modes refuse
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    capturing_ = &r;
    try {
        (void)execute(pl, {d_}, 0);
    } catch (...) {
        capturing_ = nullptr;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    capturing_ = nullptr;
    hip_check(hipStreamEndCapture(stream_, &r.graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&r.exec, r.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    r.planned = std::move(pl);
    ++r.captures;
    // capturing ran nothing: this launch prepares the state
    hip_check(hipGraphLaunch(r.exec, stream_), "hipGraphLaunch");
    queue_.clear();
    prep_pending_ = false;
}

std::unique_ptr<StateVector::Recording> StateVector::record(const OpsList &ops) {
    auto r = std::make_unique<Recording>();
    r->owner = this;
    r->ops = ops.ops;
    for (const OpRec &op : r->ops) r->nparams += op.params.size();
    LQ_ABORT_IF(r->ops.size() > 8000, "too many operations to record in one graph");
    std::vector<LOp> prog = lower_recorded(*r);
    LQ_ABORT_IF(prog.empty(), "nothing to record");
    set_basis_index(0);
    capture(*r, plan(prog, 1, true));
    return r;
}

namespace {
// the same kernels, pass layout and coefficient counts: only coefficient
// values differ
bool same_shape(const StateVector::Planned &a, const StateVector::Planned &b) {
    if (a.passes.size() != b.passes.size() || a.fns != b.fns || a.geo != b.geo) return false;
    for (size_t p = 0; p < a.passes.size(); ++p) {
        const PassPlan &x = a.passes[p], &y = b.passes[p];
        if (x.reg != y.reg || x.tbits != y.tbits || x.coef.size() != y.coef.size() || x.load_mask != y.load_mask ||
            x.store_mask != y.store_mask || x.end_transpose != y.end_transpose || x.ops.size() != y.ops.size() ||
            x.stages.size() != y.stages.size() || x.slot_ids != y.slot_ids)
            return false;
        if (!x.ops.empty() && std::memcmp(x.ops.data(), y.ops.data(), x.ops.size() * sizeof(TileOp)) != 0) return false;
        if (!x.stages.empty() && std::memcmp(x.stages.data(), y.stages.data(), x.stages.size() * sizeof(TileStage)) != 0)
            return false;
    }
    return true;
}
} // namespace

void StateVector::replay(Recording &r, const double *params, size_t np) {
    LQ_ABORT_IF(r.owner != this, "the recording belongs to another state vector");
    LQ_ABORT_IF(np != r.nparams, "wrong number of parameters for the recorded circuit");
    LQ_ABORT_IF(np > 0 && params == nullptr, "null argument");
    size_t k = 0;
    for (OpRec &op : r.ops)
        for (double &v : op.params) v = params[k++];
    std::vector<LOp> prog = lower_recorded(r);
    set_basis_index(0);
    Planned pl = plan(prog, 1, true);
    if (!same_shape(pl, r.planned)) {
        capture(r, std::move(pl));
        return;
    }
    // the previous launch has read the staging buffer once this returns
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    char *hs = static_cast<char *>(r.staging.p);
    for (size_t p = 0; p < pl.passes.size(); ++p) {
        double2 *c = reinterpret_cast<double2 *>(hs + r.coef_at) + r.coef_off[p];
        const std::vector<cd> &cf = pl.passes[p].coef;
        for (size_t i = 0; i < cf.size(); ++i) c[i] = make_double2(cf[i].real(), cf[i].imag());
    }
    hip_check(hipGraphLaunch(r.exec, stream_), "hipGraphLaunch");
    ++r.replays;
    queue_.clear();
    prep_pending_ = false;
}

// ---------------------------------------------------------------------------
void StateVector::apply_named(const std::string &name, const uint32_t *wires, int nw, const uint32_t *ctrls,
                              const uint8_t *cvals, int nc, bool inverse, const double *params, int np) {
    GateId g;
    LQ_ABORT_IF(!gate_from_name(name, &g), (std::string("Operation does not exist for ") + name).c_str());
    if (g == GateId::Identity) return;
    const GateInfo &info = gate_info(g);
    if (nc > 0)
        LQ_ABORT_IF(!info.controllable, (std::string("Controlled operation does not exist for ") + name).c_str());
    if (info.nwires > 0)
        LQ_ABORT_IF(nw != info.nwires, (std::string("Wrong number of wires for ") + name).c_str());
    else
        LQ_ABORT_IF(nw < 1, "Number of wires must be larger than 0");
    LQ_ABORT_IF(np < info.nparams, (std::string("Wrong number of parameters for ") + name).c_str());
    check_wires(n_, wires, nw, ctrls, nc);
    enqueue(lower_gate(n_, g, wires, nw, ctrls, cvals, nc, inverse, params));
}

void StateVector::apply_matrix(const void *mat, const uint32_t *wires, int nw, const uint32_t *ctrls,
                               const uint8_t *cvals, int nc, bool inverse) {
    LQ_ABORT_IF(nw < 1, "Number of wires must be larger than 0");
    check_wires(n_, wires, nw, ctrls, nc);
    const size_t d = static_cast<size_t>(1) << nw;
    std::vector<cd> m(d * d);
    if (dtype_) {
        const auto *src = static_cast<const std::complex<double> *>(mat);
        for (size_t i = 0; i < d * d; ++i) m[i] = src[i];
    } else {
        const auto *src = static_cast<const std::complex<float> *>(mat);
        for (size_t i = 0; i < d * d; ++i) m[i] = cd(src[i].real(), src[i].imag());
    }
    if (inverse) {
        std::vector<cd> t(d * d);
        for (size_t r = 0; r < d; ++r)
            for (size_t c = 0; c < d; ++c) t[c * d + r] = std::conj(m[r * d + c]);
        m.swap(t);
    }
    std::vector<int> tb(nw);
    for (int j = 0; j < nw; ++j) tb[j] = phys_bit(n_, wires[j]);
    uint64_t cm = 0, cv = 0;
    for (int i = 0; i < nc; ++i) {
        cm |= 1ull << phys_bit(n_, ctrls[i]);
        if (!cvals || cvals[i]) cv |= 1ull << phys_bit(n_, ctrls[i]);
    }
    if (nw > kMaxFusedTargets) {
        // wider than a tile op: the pending gates, then the operator out of
        // place into a work state, copied back
        flush();
        void *tmp = workspace(5);
        apply_wide(std::move(m), tb, cm, cv, data(), tmp);
        launch_stream_copy(tmp, d_, bytes(), stream_);
        return;
    }
    if (apply_dense5_mfma(m, tb, cm)) return;
    enqueue(lower_matrix(std::move(m), tb, cm, cv, false));
}

bool StateVector::apply_dense5_mfma(const std::vector<cd> &m, const std::vector<int> &tb, uint64_t cm) {
    if (dtype_ != 0 || tb.size() != 5 || cm != 0 || n_ < 12 || !tuning().mfma_dense5.load()) return false;
    int pos[5]; // index bit of local row bit j (tb[0] is the local MSB)
    uint32_t seen = 0;
    for (int j = 0; j < 5; ++j) {
        if (tb[j] < 0 || tb[j] > 4) return false;
        pos[j] = tb[j];
        seen |= 1u << tb[j];
    }
    if (seen != 31u) return false;
    // U in the order of the index's low five bits, then its real 64x64 form
    // [[Re U, -Im U], [Im U, Re U]]
    auto to_bits = [&](uint32_t r) {
        uint32_t g = 0;
        for (int j = 0; j < 5; ++j)
            if ((r >> (4 - j)) & 1u) g |= 1u << pos[j];
        return g;
    };
    std::vector<float> w(64 * 64);
    for (uint32_t r = 0; r < 32; ++r)
        for (uint32_t c = 0; c < 32; ++c) {
            const cd u = m[r * 32 + c];
            const uint32_t gr = to_bits(r), gc = to_bits(c);
            w[gr * 64 + gc] = static_cast<float>(u.real());
            w[gr * 64 + 32 + gc] = static_cast<float>(-u.imag());
            w[(32 + gr) * 64 + gc] = static_cast<float>(u.imag());
            w[(32 + gr) * 64 + 32 + gc] = static_cast<float>(u.real());
        }
    flush();
    void *sp = data();
    // the previous operator may still be read by a kernel in flight
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    mfma_w_.ensure(w.size() * sizeof(float));
    hip_check(hipMemcpyAsync(mfma_w_.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, stream_),
              "hipMemcpyAsync(matrix)");
    launch_dense5_c64_mfma(sp, sp, static_cast<const float *>(mfma_w_.p), n_, stream_);
    hip_check(hipGetLastError(), "matrix-core operator launch");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); // w is freed on return
    return true;
}

void StateVector::apply_wide(std::vector<cd> m, const std::vector<int> &tb, uint64_t cm, uint64_t cv, const void *in,
                             void *out) {
    const int k = static_cast<int>(tb.size());
    LQ_ABORT_IF(m.size() != (static_cast<size_t>(1) << (2 * k)), "The size of matrix does not match with the given number of wires");
    std::vector<double2> h(m.size());
    for (size_t i = 0; i < m.size(); ++i) h[i] = make_double2(m[i].real(), m[i].imag());
    // the previous upload may still be read by a kernel in flight
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    wide_.ensure(h.size() * sizeof(double2));
    hip_check(hipMemcpyAsync(wide_.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice, stream_),
              "hipMemcpyAsync(matrix)");
    dispatch(dtype_, [&](auto t) {
        launch_dense_wide<decltype(t)>(in, out, static_cast<const double2 *>(wide_.p), tb.data(), k, cm, cv, n_,
                                       stream_);
    });
    hip_check(hipGetLastError(), "dense operator launch");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); // h is freed on return
}

void StateVector::apply_pauli_rot(const uint32_t *wires, int nw, bool inverse, double theta, const std::string &word) {
    LQ_ABORT_IF(static_cast<int>(word.size()) != nw, "wires and word have incompatible dimensions.");
    check_wires(n_, wires, nw, nullptr, 0);
    // exp(-i theta/2 P) = V^dagger MultiRZ(theta) V with V the per-wire basis
    // change taking P to a product of Z: H for X, H S^dagger for Y, nothing
    // for Z; every letter other than Z and Y acts as X, as in the reference
    // (GateImplementationsLM.hpp:575-629). An all-Z word is MultiRZ alone.
    auto one = [&](GateId g, uint32_t w, bool inv) {
        enqueue(lower_gate(n_, g, &w, 1, nullptr, nullptr, 0, inv, nullptr));
    };
    for (int j = 0; j < nw; ++j) {
        if (word[j] == 'Z') continue;
        if (word[j] == 'Y') one(GateId::S, wires[j], true);
        one(GateId::Hadamard, wires[j], false);
    }
    enqueue(lower_gate(n_, GateId::MultiRZ, wires, nw, nullptr, nullptr, 0, inverse, &theta));
    for (int j = 0; j < nw; ++j) {
        if (word[j] == 'Z') continue;
        one(GateId::Hadamard, wires[j], false);
        if (word[j] == 'Y') one(GateId::S, wires[j], false);
    }
}

double StateVector::apply_generator(const std::string &name, const uint32_t *wires, int nw, const uint32_t *ctrls,
                                    const uint8_t *cvals, int nc) {
    GateId g;
    LQ_ABORT_IF(!gate_from_name(name, &g), (std::string("Generator does not exist for ") + name).c_str());
    LQ_ABORT_IF(!gate_info(g).has_generator, (std::string("Generator does not exist for ") + name).c_str());
    check_wires(n_, wires, g == GateId::GlobalPhase ? 0 : nw, ctrls, nc);
    double scale = 0.0;
    LOp op = lower_generator(n_, g, wires, nw, ctrls, cvals, nc, &scale);
    if (!(g == GateId::GlobalPhase && nc == 0)) enqueue(std::move(op));
    return scale;
}

void StateVector::apply_oprec(const OpRec &op, bool adj) {
    const bool inv = op.inverse ^ adj;
    const uint32_t *cw = op.ctrls.empty() ? nullptr : op.ctrls.data();
    const uint8_t *cv = op.cvals.empty() ? nullptr : op.cvals.data();
    if (op.named) {
        apply_named(op.name, op.wires.data(), static_cast<int>(op.wires.size()), cw, cv,
                    static_cast<int>(op.ctrls.size()), inv, op.params.data(), static_cast<int>(op.params.size()));
        return;
    }
    LQ_ABORT_IF(!op.ctrls.empty(), (std::string("Controlled operation does not exist for ") + op.name).c_str());
    LQ_ABORT_IF(op.matrix.size() != (static_cast<size_t>(1) << (2 * op.wires.size())),
                (std::string("Operation does not exist for ") + op.name).c_str());
    check_wires(n_, op.wires.data(), static_cast<int>(op.wires.size()), nullptr, 0);
    std::vector<cd> m = op.matrix;
    const size_t d = static_cast<size_t>(1) << op.wires.size();
    if (inv) {
        std::vector<cd> t(d * d);
        for (size_t r = 0; r < d; ++r)
            for (size_t c = 0; c < d; ++c) t[c * d + r] = std::conj(m[r * d + c]);
        m.swap(t);
    }
    std::vector<int> tb;
    for (auto w : op.wires) tb.push_back(phys_bit(n_, w));
    enqueue(lower_matrix(std::move(m), tb, 0, 0, false));
}

double StateVector::norm2() {
    flush();
    const uint32_t nb = static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(1, length() / 256)));
    partials_.ensure(nb * sizeof(double2));
    dispatch(dtype_, [&](auto t) {
        launch_norm2<decltype(t)>(data(), n_, static_cast<double2 *>(partials_.p), nb, stream_);
    });
    std::vector<double2> h(nb);
    read_back(h.data(), partials_.p, nb * sizeof(double2));
    double s = 0.0, c = 0.0;
    for (const auto &v : h) {
        const double t = s + v.x;
        c += (std::fabs(s) >= std::fabs(v.x)) ? (s - t) + v.x : (v.x - t) + s;
        s = t;
    }
    return s + c;
}

cd StateVector::overlap(StateVector &other) {
    LQ_ABORT_IF(other.n_ != n_, "states of different sizes");
    flush();
    other.flush();
    const void *op = other.data(); // before the wait: it may enqueue a deferred reset on other's stream
    other.synchronize();
    const uint32_t nb = static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(1, length() / 256)));
    partials_.ensure(nb * sizeof(double2));
    launch_overlap(data(), dtype_, op, other.dtype_, n_, static_cast<double2 *>(partials_.p), nb, stream_);
    std::vector<double2> h(nb);
    read_back(h.data(), partials_.p, nb * sizeof(double2));
    double sr = 0.0, cr = 0.0, si = 0.0, ci = 0.0;
    for (const auto &v : h) {
        double t = sr + v.x;
        cr += (std::fabs(sr) >= std::fabs(v.x)) ? (sr - t) + v.x : (v.x - t) + sr;
        sr = t;
        t = si + v.y;
        ci += (std::fabs(si) >= std::fabs(v.y)) ? (si - t) + v.y : (v.y - t) + si;
        si = t;
    }
    return cd(sr + cr, si + ci);
}

void StateVector::normalize() {
    const double norm = std::sqrt(norm2());
    const double eps = dtype_ ? std::numeric_limits<double>::epsilon() : std::numeric_limits<float>::epsilon();
    LQ_ABORT_IF(norm < eps * 1e2, "Vector has norm close to zero and cannot be normalized");
    dispatch(dtype_, [&](auto t) { launch_scale<decltype(t)>(data(), n_, 1.0 / norm, 0.0, stream_); });
}

void StateVector::collapse(uint32_t wire, bool branch) {
    LQ_ABORT_IF(static_cast<int>(wire) >= n_, "wire index out of range");
    flush();
    dispatch(dtype_, [&](auto t) {
        launch_collapse<decltype(t)>(data(), n_, phys_bit(n_, wire), branch ? 1 : 0, 1.0, stream_);
    });
    normalize();
}

std::vector<double> StateVector::reduce_rows_to_host(uint32_t rows, uint32_t cols) {
    slots_.ensure(static_cast<size_t>(cols) * sizeof(double));
    launch_reduce_rows(static_cast<const double *>(partials_.p), rows, cols, static_cast<double *>(slots_.p), stream_);
    std::vector<double> h(cols);
    read_back(h.data(), slots_.p, cols * sizeof(double));
    return h;
}

std::vector<double> StateVector::probs(const uint32_t *wires, int nw, bool all) {
    if (!all && nw == 0) return {1.0};
    std::vector<int> bits;
    if (all) {
        for (int b = n_ - 1; b >= 0; --b) bits.push_back(b);
    } else {
        check_wires(n_, wires, nw, nullptr, 0);
        for (int j = 0; j < nw; ++j) bits.push_back(phys_bit(n_, wires[j]));
    }
    const int m = static_cast<int>(bits.size());
    flush();
    if (m > 12) {
        slots_.ensure(sizeof(double) << m);
        dispatch(dtype_, [&](auto t) {
            launch_probs_gather<decltype(t)>(data(), n_, bits.data(), m, static_cast<double *>(slots_.p), stream_);
        });
        std::vector<double> h(static_cast<size_t>(1) << m);
        read_back(h.data(), slots_.p, h.size() * sizeof(double));
        return h;
    }
    const uint32_t nb = static_cast<uint32_t>(std::min<uint64_t>(512, std::max<uint64_t>(1, length() / 1024)));
    partials_.ensure((static_cast<size_t>(nb) * sizeof(double)) << m);
    dispatch(dtype_, [&](auto t) {
        launch_probs_hist<decltype(t)>(data(), n_, bits.data(), m, static_cast<double *>(partials_.p), nb, stream_);
    });
    return reduce_rows_to_host(nb, 1u << m);
}

double StateVector::expval_named(const std::string &name, const uint32_t *wires, int nw) {
    if (name == "Identity") return 1.0;
    const auto &m = named_obs_matrix(name);
    LQ_ABORT_IF(nw != 1, "named observables act on one wire");
    check_wires(n_, wires, 1, nullptr, 0);
    // the pending gates and the dot run as one program
    std::vector<LOp> prog;
    prog.swap(queue_);
    prog.push_back(make_dot(lower_matrix(m, {phys_bit(n_, wires[0])}, 0, 0, true), 0, 0, 0));
    return run(prog, {d_}, 1)[0].real();
}

std::vector<double> StateVector::expval_named_batch(const std::vector<std::string> &names,
                                                    const std::vector<uint32_t> &wires) {
    LQ_ABORT_IF(names.size() != wires.size(), "names and wires differ in length");
    std::vector<LOp> dots;
    std::vector<int> slot_of(names.size(), -1);
    for (size_t i = 0; i < names.size(); ++i) {
        if (names[i] == "Identity") continue;
        const auto &m = named_obs_matrix(names[i]);
        check_wires(n_, &wires[i], 1, nullptr, 0);
        slot_of[i] = static_cast<int>(dots.size());
        dots.push_back(make_dot(lower_matrix(m, {phys_bit(n_, wires[i])}, 0, 0, true), 0, 0, slot_of[i]));
    }
    const int nslots = static_cast<int>(dots.size());
    std::vector<LOp> prog;
    prog.swap(queue_);
    for (auto &d : dots) prog.push_back(std::move(d));
    const auto v = run(prog, {d_}, nslots);
    std::vector<double> out(names.size(), 1.0);
    for (size_t i = 0; i < names.size(); ++i)
        if (slot_of[i] >= 0) out[i] = v[slot_of[i]].real();
    return out;
}

double StateVector::expval_matrix(const void *mat, const uint32_t *wires, int nw) {
    LQ_ABORT_IF(nw < 1, "Number of wires must be larger than 0");
    check_wires(n_, wires, nw, nullptr, 0);
    const size_t d = static_cast<size_t>(1) << nw;
    std::vector<cd> m(d * d);
    if (dtype_) {
        const auto *s = static_cast<const std::complex<double> *>(mat);
        for (size_t i = 0; i < d * d; ++i) m[i] = s[i];
    } else {
        const auto *s = static_cast<const std::complex<float> *>(mat);
        for (size_t i = 0; i < d * d; ++i) m[i] = cd(s[i].real(), s[i].imag());
    }
    std::vector<int> tb(nw);
    for (int j = 0; j < nw; ++j) tb[j] = phys_bit(n_, wires[j]);
    if (nw > kMaxFusedTargets) {
        flush();
        void *tmp = workspace(5);
        apply_wide(std::move(m), tb, 0, 0, data(), tmp);
        return inner(d_, tmp).real();
    }
    std::vector<LOp> prog;
    prog.swap(queue_);
    prog.push_back(make_dot(lower_matrix(std::move(m), tb, 0, 0, true), 0, 0, 0));
    return run(prog, {d_}, 1)[0].real();
}

cd StateVector::inner(const void *bra, const void *ket) {
    return run({identity_dot(1, 0, 0)}, {const_cast<void *>(bra), const_cast<void *>(ket)}, 1)[0];
}

void StateVector::apply_obs_inplace(const Obs &o, void *state, std::vector<LOp> &prog) {
    (void)state;
    switch (o.kind) {
    case Obs::Named: {
        if (o.name == "Identity") return;
        GateId g;
        std::vector<cd> m;
        if (o.name == "PauliX" || o.name == "PauliY" || o.name == "PauliZ" || o.name == "Hadamard") {
            m = named_obs_matrix(o.name);
        } else {
            LQ_ABORT_IF(!gate_from_name(o.name, &g), (std::string("Operation does not exist for ") + o.name).c_str());
            LQ_ABORT_IF(gate_info(g).nparams != 0, "parametric named observables are not supported");
            m = gate_matrix(g, nullptr, false, static_cast<int>(o.wires.size()));
        }
        std::vector<int> tb;
        for (auto w : o.wires) tb.push_back(phys_bit(n_, w));
        prog.push_back(lower_matrix(std::move(m), tb, 0, 0, false));
        return;
    }
    case Obs::Hermitian: {
        std::vector<int> tb;
        for (auto w : o.wires) tb.push_back(phys_bit(n_, w));
        prog.push_back(lower_matrix(o.matrix, tb, 0, 0, false));
        return;
    }
    case Obs::Tensor:
        for (const auto &f : o.terms) apply_obs_inplace(*f, state, prog);
        return;
    case Obs::Ham: LQ_ABORT("a Hamiltonian cannot be applied in place as a product factor");
    case Obs::Sparse: LQ_ABORT("a sparse Hamiltonian cannot be a product factor");
    }
}

void StateVector::apply_obs_to(const Obs &o, const void *in, void *out) {
    for (auto w : o.all_wires()) LQ_ABORT_IF(static_cast<int>(w) >= n_, "observable wire out of range");
    if (o.kind == Obs::Sparse) {
        LQ_ABORT_IF(static_cast<int>(o.wires.size()) != n_, "SparseH wire count does not match state-vector size");
        upload_csr(sparse_view(o));
        const char *base = static_cast<const char *>(sparse_.p);
        dispatch(dtype_, [&](auto t) {
            launch_sparse_rows<decltype(t)>(in, out, reinterpret_cast<const int64_t *>(base),
                                            reinterpret_cast<const int64_t *>(base + csr_indices_off_),
                                            base + csr_values_off_, length(), 1, nullptr,
                                            static_cast<uint32_t>(std::min<uint64_t>(4096, std::max<uint64_t>(1, length() / 256))),
                                            stream_);
        });
        hip_check(hipStreamSynchronize(stream_), "sync");
        return;
    }
    std::vector<PauliWord> words;
    if (obs_to_pauli(o, n_, &words)) {
        std::vector<PauliTerm> terms;
        for (const auto &w : words) {
            cd c = w.coeff;
            for (int i = 0; i < (w.ny & 3); ++i) c *= cd(0, -1);
            terms.push_back({w.x, w.yz, c.real(), c.imag()});
        }
        small_.ensure(terms.size() * sizeof(PauliTerm) + 256);
        upload_small(small_.p, small_.bytes, terms.data(), terms.size() * sizeof(PauliTerm));
        dispatch(dtype_, [&](auto t) {
            launch_pauli_sum<decltype(t)>(in, out, static_cast<const PauliTerm *>(small_.p),
                                          static_cast<uint32_t>(terms.size()), n_, 1, nullptr, 0, stream_);
        });
        hip_check(hipStreamSynchronize(stream_), "sync");
        return;
    }
    if (o.kind == Obs::Ham) {
        void *tmp = workspace(7);
        LQ_ABORT_IF(tmp == in || tmp == out, "workspace clash");
        // a fresh workspace may hold NaN bit patterns: clear it, do not scale it
        hip_check(hipMemsetAsync(out, 0, bytes(), stream_), "hipMemsetAsync");
        for (size_t i = 0; i < o.terms.size(); ++i) {
            LQ_ABORT_IF(o.terms[i]->kind == Obs::Ham, "nested Hamiltonians are not supported");
            apply_obs_to(*o.terms[i], in, tmp);
            dispatch(dtype_, [&](auto t) {
                launch_axpy<decltype(t)>(tmp, out, o.coeffs[i], 0.0, n_, stream_);
            });
        }
        return;
    }
    launch_stream_copy(in, out, bytes(), stream_);
    // factors act in place as one program; a Hermitian factor wider than a
    // tile op runs out of place through a work state between them
    std::vector<const Obs *> factors;
    if (o.kind == Obs::Tensor)
        for (const auto &f : o.terms) factors.push_back(f.get());
    else
        factors.push_back(&o);
    std::vector<LOp> prog;
    for (const Obs *f : factors) {
        if (f->kind == Obs::Hermitian && static_cast<int>(f->wires.size()) > kMaxFusedTargets) {
            run(prog, {out}, 0);
            prog.clear();
            std::vector<int> tb;
            for (auto w : f->wires) tb.push_back(phys_bit(n_, w));
            void *tmp = workspace(8);
            LQ_ABORT_IF(tmp == in || tmp == out, "workspace clash");
            apply_wide(f->matrix, tb, 0, 0, out, tmp);
            launch_stream_copy(tmp, out, bytes(), stream_);
        } else {
            apply_obs_inplace(*f, out, prog);
        }
    }
    run(prog, {out}, 0);
}

double StateVector::expval_obs(const Obs &o) {
    for (auto w : o.all_wires()) LQ_ABORT_IF(static_cast<int>(w) >= n_, "observable wire out of range");
    if (o.kind == Obs::Sparse) {
        LQ_ABORT_IF(static_cast<int>(o.wires.size()) != n_, "SparseH wire count does not match state-vector size");
        double mean, ms;
        sparse_moments(sparse_view(o), &mean, &ms);
        return mean;
    }
    std::vector<PauliWord> words;
    const bool pauli = obs_to_pauli(o, n_, &words);
    const int ktm = std::min(kt_for(1), n_);
    bool fits = pauli && words.size() <= 32;
    for (const auto &w : words) fits = fits && __builtin_popcountll(w.x) <= ktm - 2;
    if (!fits) flush();
    if (pauli) {
        if (fits) {
            // the pending gates and the dots run as one program
            std::vector<LOp> prog;
            prog.swap(queue_);
            for (size_t i = 0; i < words.size(); ++i)
                prog.push_back(make_dot(pauli_dot_op(n_, words[i]), 0, 0, static_cast<int>(i)));
            const auto v = run(prog, {d_}, static_cast<int>(words.size()));
            double e = 0.0;
            for (size_t i = 0; i < words.size(); ++i) e += (words[i].coeff * v[i]).real();
            return e;
        }
        std::vector<PauliTerm> terms;
        for (const auto &w : words) {
            cd c = w.coeff;
            for (int i = 0; i < (w.ny & 3); ++i) c *= cd(0, -1);
            terms.push_back({w.x, w.yz, c.real(), c.imag()});
        }
        small_.ensure(terms.size() * sizeof(PauliTerm) + 256);
        upload_small(small_.p, small_.bytes, terms.data(), terms.size() * sizeof(PauliTerm));
        const uint32_t nb = static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(1, length() / 256)));
        partials_.ensure(nb * sizeof(double2));
        dispatch(dtype_, [&](auto t) {
            launch_pauli_sum<decltype(t)>(data(), nullptr, static_cast<const PauliTerm *>(small_.p),
                                          static_cast<uint32_t>(terms.size()), n_, 0,
                                          static_cast<double2 *>(partials_.p), nb, stream_);
        });
        std::vector<double2> h(nb);
        read_back(h.data(), partials_.p, nb * sizeof(double2));
        double s = 0.0, c = 0.0;
        for (const auto &v : h) {
            const double t = s + v.x;
            c += (std::fabs(s) >= std::fabs(v.x)) ? (s - t) + v.x : (v.x - t) + s;
            s = t;
        }
        return s + c;
    }
    void *tmp = workspace(6);
    apply_obs_to(o, data(), tmp);
    return inner(data(), tmp).real();
}

// ---------------------------------------------------------------------------
// Variances, sparse Hamiltonians, samples
// ---------------------------------------------------------------------------
void StateVector::dense_moments(std::vector<cd> m, const uint32_t *wires, int nw, double *mean, double *mean_square) {
    LQ_ABORT_IF(nw < 0, "Number of wires must not be negative");
    check_wires(n_, wires, nw, nullptr, 0);
    const size_t d = static_cast<size_t>(1) << nw;
    LQ_ABORT_IF(m.size() != d * d, "The size of matrix does not match with the given number of wires");
    if (nw > kMaxFusedTargets) {
        // O psi into a work state: Re<psi|O psi> and <O psi|O psi>
        std::vector<int> tb(static_cast<size_t>(nw));
        for (int j = 0; j < nw; ++j) tb[j] = phys_bit(n_, wires[j]);
        flush();
        void *tmp = workspace(5);
        apply_wide(std::move(m), tb, 0, 0, data(), tmp);
        *mean = inner(d_, tmp).real();
        *mean_square = inner(tmp, tmp).real();
        return;
    }
    // <O psi|O psi> = <psi|O^dagger O|psi>
    std::vector<cd> mm(d * d, 0.0);
    for (size_t r = 0; r < d; ++r)
        for (size_t c = 0; c < d; ++c) {
            cd acc = 0.0;
            for (size_t k = 0; k < d; ++k) acc += std::conj(m[k * d + r]) * m[k * d + c];
            mm[r * d + c] = acc;
        }
    std::vector<int> tb(static_cast<size_t>(nw));
    for (int j = 0; j < nw; ++j) tb[j] = phys_bit(n_, wires[j]);
    // the pending gates and both dots run as one program
    std::vector<LOp> prog;
    prog.swap(queue_);
    prog.push_back(make_dot(lower_matrix(std::move(m), tb, 0, 0, true), 0, 0, 0));
    prog.push_back(make_dot(lower_matrix(std::move(mm), tb, 0, 0, true), 0, 0, 1));
    const auto v = run(prog, {d_}, 2);
    *mean = v[0].real();
    *mean_square = v[1].real();
}

double StateVector::var_named(const std::string &name, const uint32_t *wires, int nw) {
    // the reference applies the operation by name to a copy (Identity is skipped)
    std::vector<cd> m;
    if (name == "Identity") {
        m = {1.0, 0.0, 0.0, 1.0};
        nw = std::min(nw, 1);
    } else if (name == "PauliX" || name == "PauliY" || name == "PauliZ" || name == "Hadamard") {
        LQ_ABORT_IF(nw != 1, "named observables act on one wire");
        m = named_obs_matrix(name);
    } else {
        GateId g;
        LQ_ABORT_IF(!gate_from_name(name, &g), (std::string("Operation does not exist for ") + name).c_str());
        LQ_ABORT_IF(gate_info(g).nparams != 0, "parametric operations have no variance without parameters");
        LQ_ABORT_IF(gate_info(g).nwires > 0 && gate_info(g).nwires != nw, "wrong number of wires for the operation");
        m = gate_matrix(g, nullptr, false, nw);
    }
    double mean, ms;
    if (nw == 0) {
        // Identity on no wires: <psi|psi> - <psi|psi>^2
        const double nn = norm2();
        return nn - nn * nn;
    }
    dense_moments(std::move(m), wires, nw, &mean, &ms);
    return ms - mean * mean;
}

double StateVector::var_matrix(const void *mat, const uint32_t *wires, int nw) {
    LQ_ABORT_IF(nw < 1, "Number of wires must be larger than 0");
    const size_t d = static_cast<size_t>(1) << nw;
    std::vector<cd> m(d * d);
    if (dtype_) {
        const auto *src = static_cast<const std::complex<double> *>(mat);
        for (size_t i = 0; i < d * d; ++i) m[i] = src[i];
    } else {
        const auto *src = static_cast<const std::complex<float> *>(mat);
        for (size_t i = 0; i < d * d; ++i) m[i] = cd(src[i].real(), src[i].imag());
    }
    double mean, ms;
    dense_moments(std::move(m), wires, nw, &mean, &ms);
    return ms - mean * mean;
}

double StateVector::var_obs(const Obs &o) {
    for (auto w : o.all_wires()) LQ_ABORT_IF(static_cast<int>(w) >= n_, "observable wire out of range");
    if (o.kind == Obs::Sparse) {
        LQ_ABORT_IF(static_cast<int>(o.wires.size()) != n_, "SparseH wire count does not match state-vector size");
        const CsrView h = sparse_view(o);
        double mean, ms;
        sparse_moments(h, &mean, &ms);
        return ms - mean * mean;
    }
    flush();
    void *tmp = workspace(6);
    apply_obs_to(o, data(), tmp);
    // <O psi|O psi> and Re<O psi|psi> in one read pass
    const auto v = run({identity_dot(1, 1, 0), identity_dot(0, 1, 1)}, {d_, tmp}, 2);
    const double mean = v[1].real();
    return v[0].real() - mean * mean;
}

CsrView StateVector::sparse_view(const Obs &o) {
    const size_t nnz = o.sp_values.size();
    sparse_host_.resize(nnz * elem());
    if (dtype_) {
        std::memcpy(sparse_host_.data(), o.sp_values.data(), nnz * sizeof(cd));
    } else {
        auto *dst = reinterpret_cast<std::complex<float> *>(sparse_host_.data());
        for (size_t i = 0; i < nnz; ++i)
            dst[i] = std::complex<float>(static_cast<float>(o.sp_values[i].real()), static_cast<float>(o.sp_values[i].imag()));
    }
    return CsrView{o.sp_indptr.data(), o.sp_indptr.size() - 1, o.sp_indices.data(), sparse_host_.data(), nnz};
}

void StateVector::sparse_moments(const CsrView &h, double *mean, double *mean_square) {
    upload_csr(h);
    const uint32_t nb = static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(1, length() / 256)));
    partials_.ensure(2 * nb * sizeof(double2));
    const char *base = static_cast<const char *>(sparse_.p);
    const auto *indptr = reinterpret_cast<const int64_t *>(base);
    const auto *indices = reinterpret_cast<const int64_t *>(base + csr_indices_off_);
    const void *values = base + csr_values_off_;
    dispatch(dtype_, [&](auto t) {
        launch_sparse_rows<decltype(t)>(data(), nullptr, indptr, indices, values, h.nrows, 0,
                                        static_cast<double2 *>(partials_.p), nb, stream_);
    });
    std::vector<double2> hp(2 * static_cast<size_t>(nb));
    read_back(hp.data(), partials_.p, hp.size() * sizeof(double2));
    double s0 = 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
    auto add = [](double &sum, double &comp, double x) {
        const double t = sum + x;
        comp += (std::fabs(sum) >= std::fabs(x)) ? (sum - t) + x : (x - t) + sum;
        sum = t;
    };
    for (uint32_t b = 0; b < nb; ++b) {
        add(s0, c0, hp[2 * b].x);
        add(s1, c1, hp[2 * b + 1].x);
    }
    *mean = s0 + c0;
    *mean_square = s1 + c1;
}

void StateVector::upload_csr(const CsrView &h) {
    LQ_ABORT_IF(h.nrows != length(), "Statevector and Hamiltonian have incompatible sizes.");
    // every index the kernel follows is checked here
    LQ_ABORT_IF(h.indptr[0] != 0, "sparse matrix offsets must start at 0");
    for (uint64_t i = 0; i < h.nrows; ++i)
        LQ_ABORT_IF(h.indptr[i + 1] < h.indptr[i], "sparse matrix offsets must not decrease");
    LQ_ABORT_IF(static_cast<uint64_t>(h.indptr[h.nrows]) != h.nnz, "sparse matrix offsets do not match the number of values");
    const int64_t len = static_cast<int64_t>(length());
    for (uint64_t k = 0; k < h.nnz; ++k)
        LQ_ABORT_IF(h.indices[k] < 0 || h.indices[k] >= len, "sparse matrix column index out of range");
    flush();
    auto align = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    csr_indices_off_ = align((h.nrows + 1) * sizeof(int64_t));
    csr_values_off_ = align(csr_indices_off_ + h.nnz * sizeof(int64_t));
    const size_t total = csr_values_off_ + h.nnz * elem() + 16;
    sparse_.ensure(total);
    char *base = static_cast<char *>(sparse_.p);
    hip_check(hipMemcpyAsync(base, h.indptr, (h.nrows + 1) * sizeof(int64_t), hipMemcpyHostToDevice, stream_), "memcpy");
    if (h.nnz) {
        hip_check(hipMemcpyAsync(base + csr_indices_off_, h.indices, h.nnz * sizeof(int64_t), hipMemcpyHostToDevice,
                                 stream_),
                  "memcpy");
        hip_check(hipMemcpyAsync(base + csr_values_off_, h.values, h.nnz * elem(), hipMemcpyHostToDevice, stream_),
                  "memcpy");
    }
}

double StateVector::expval_sparse(const CsrView &h) {
    double mean, ms;
    sparse_moments(h, &mean, &ms);
    return mean;
}

double StateVector::var_sparse(const CsrView &h) {
    double mean, ms;
    sparse_moments(h, &mean, &ms);
    return ms - mean * mean;
}

namespace {
// Alias table (Vose's method with LIFO work lists) over probabilities p:
// entry i keeps i with probability first, else returns its partner
// (lq/measurements/MeasurementKernels.hpp:347-380).
template <typename T> std::vector<std::pair<double, size_t>> alias_table(const std::vector<T> &p) {
    const size_t N = p.size(), none = std::numeric_limits<size_t>::max();
    std::vector<std::pair<double, size_t>> b(N, {0.0, none});
    std::vector<size_t> under, over;
    for (size_t i = 0; i < N; ++i) {
        b[i].first = static_cast<T>(N) * p[i];
        (b[i].first < 1.0 ? under : over).push_back(i);
    }
    while (!under.empty() && !over.empty()) {
        const size_t i = over.back();
        over.pop_back();
        const size_t j = under.back();
        under.pop_back();
        b[j].second = i;
        b[i].first += b[j].first - 1.0;
        (b[i].first < 1.0 ? under : over).push_back(i);
    }
    return b;
}
} // namespace

void StateVector::generate_samples(const uint32_t *wires, int nw, bool all, uint64_t shots, bool seeded, uint64_t seed,
                                   uint64_t *out) {
    const std::vector<double> p = probs(wires, nw, all);
    const int k = all ? n_ : nw;
    std::mt19937 gen;
    if (seeded) {
        gen.seed(static_cast<std::mt19937::result_type>(seed));
    } else {
        std::random_device rd;
        gen.seed(rd());
    }
    auto draw = [&](auto tag) {
        using T = decltype(tag);
        const std::vector<T> pt(p.begin(), p.end());
        const auto table = alias_table(pt);
        const size_t N = pt.size(), none = std::numeric_limits<size_t>::max();
        std::uniform_real_distribution<T> u(0.0, 1.0);
        for (uint64_t s = 0; s < shots; ++s) {
            size_t idx = static_cast<size_t>(u(gen) * static_cast<T>(N));
            if (u(gen) >= table[idx].first && table[idx].second != none) idx = table[idx].second;
            for (int j = 0; j < k; ++j) out[s * k + (k - 1 - j)] = (idx >> j) & 1u;
        }
    };
    if (dtype_) draw(double{});
    else draw(float{});
}

// ---------------------------------------------------------------------------
// Adjoint Jacobian (lq/algorithms/AdjointJacobianLQubit.hpp:225-491)
// ---------------------------------------------------------------------------
// Lowered op of an OpsList entry, as applied forward (inverse flag honoured)
// or, with `dagger`, as its adjoint U^dagger
// (core/algorithms/AdjointJacobianBase.hpp:108-127).
LOp lower_record(int n, const OpRec &op, bool dagger) {
    const uint32_t *cw = op.ctrls.empty() ? nullptr : op.ctrls.data();
    const uint8_t *cv = op.cvals.empty() ? nullptr : op.cvals.data();
    const bool inv = op.inverse != dagger;
    if (op.named)
        return lower_gate(n, op.gid, op.wires.data(), static_cast<int>(op.wires.size()), cw, cv,
                          static_cast<int>(op.ctrls.size()), inv, op.params.data());
    // the reference resolves a controlled op by name only
    // (lq/StateVectorLQubit.hpp:457-482, DynamicDispatcher.hpp:385-388)
    LQ_ABORT_IF(!op.ctrls.empty(), (std::string("Controlled operation does not exist for ") + op.name).c_str());
    std::vector<cd> mm = op.matrix;
    const size_t d = static_cast<size_t>(1) << op.wires.size();
    LQ_ABORT_IF(mm.size() != d * d, (std::string("Operation does not exist for ") + op.name).c_str());
    if (inv) {
        std::vector<cd> t(d * d);
        for (size_t r = 0; r < d; ++r)
            for (size_t c = 0; c < d; ++c) t[c * d + r] = std::conj(mm[r * d + c]);
        mm.swap(t);
    }
    std::vector<int> tb;
    for (auto w : op.wires) tb.push_back(phys_bit(n, w));
    return lower_matrix(std::move(mm), tb, 0, 0, false);
}

void apply_forward(StateVector &sv, const OpsList &ops, void *state) {
    std::vector<LOp> fwd;
    for (const auto &op : ops.ops) {
        if (op.named && op.gid == GateId::Identity) continue;
        fwd.push_back(lower_record(sv.n(), op, false));
    }
    sv.run(fwd, {state}, 0);
}

// The backward sweep (lq/algorithms/AdjointJacobianLQubit.hpp:225-319 and
// lq/algorithms/VectorJacobianProduct.hpp:84-157) as one program on the
// states {lambda, b_1 .. b_m}: for each trainable op, from the last, the dots
// <b_j | G lambda> into slot slot0 + t*m + j (t = position in the sorted
// trainable list), then U^dagger on every state. Ops before the first
// trainable parameter change no result and are left out. Returns the
// generator scales, negated for inverse ops.
std::vector<double> backward_sweep(int n, const OpsList &ops, const std::vector<uint64_t> &tp, size_t m, int slot0,
                                   std::vector<LOp> &prog) {
    const size_t ntp = tp.size();
    const uint32_t all = (1u << (m + 1)) - 1u;
    std::vector<double> scales(ntp, 0.0);
    size_t tp_pos = ntp; // tp[tp_pos-1] is the next trainable index going backwards
    int64_t cur = static_cast<int64_t>(ops.num_par_ops) - 1;
    int64_t tnum = static_cast<int64_t>(ntp) - 1;
    size_t last_dot = prog.size();
    for (int64_t k = static_cast<int64_t>(ops.ops.size()) - 1; k >= 0; --k) {
        const OpRec &op = ops.ops[static_cast<size_t>(k)];
        LQ_ABORT_IF(op.params.size() > 1, "The operation is not supported using the adjoint differentiation method");
        if (op.name == "StatePrep" || op.name == "BasisState") continue;
        if (tp_pos == 0) break;
        if (!op.params.empty()) {
            if (cur == static_cast<int64_t>(tp[tp_pos - 1])) {
                LQ_ABORT_IF(!op.named || !gate_info(op.gid).has_generator,
                            (std::string("The operation is not supported using the adjoint differentiation method: ") +
                             op.name).c_str());
                double s = 0.0;
                const uint32_t *cw = op.ctrls.empty() ? nullptr : op.ctrls.data();
                const uint8_t *cv = op.cvals.empty() ? nullptr : op.cvals.data();
                LOp g = lower_generator(n, op.gid, op.wires.data(), static_cast<int>(op.wires.size()), cw, cv,
                                        static_cast<int>(op.ctrls.size()), &s);
                scales[static_cast<size_t>(tnum)] = s * (op.inverse ? -1.0 : 1.0);
                for (size_t j = 0; j < m; ++j)
                    prog.push_back(make_dot(g, 0, 1 + static_cast<uint32_t>(j),
                                            slot0 + static_cast<int>(tnum * static_cast<int64_t>(m) + j)));
                last_dot = prog.size();
                --tnum;
                --tp_pos;
            }
            --cur;
        }
        if (op.named && op.gid == GateId::Identity) continue;
        LOp u = lower_record(n, op, true);
        u.smask = all;
        prog.push_back(std::move(u));
    }
    // ops after the last dot change no result
    prog.resize(last_dot);
    return scales;
}


std::atomic<uint64_t> g_sweeps_grouped{0}, g_sweeps_single{0};

void adjoint_sweep_counts(bool reset, uint64_t *grouped, uint64_t *single) {
    if (grouped) *grouped = reset ? g_sweeps_grouped.exchange(0) : g_sweeps_grouped.load();
    if (single) *single = reset ? g_sweeps_single.exchange(0) : g_sweeps_single.load();
}

void adjoint_jacobian(StateVector &sv, const OpsList &ops, const std::vector<std::shared_ptr<Obs>> &obs,
                      const std::vector<uint64_t> &tp_in, bool apply_operations, double *jac, double *expvals) {
    std::vector<uint64_t> tp = tp_in;
    std::sort(tp.begin(), tp.end());
    const size_t ntp = tp.size(), nobs = obs.size();
    if (ntp == 0) return;
    LQ_ABORT_IF(nobs == 0, "no observables given");
    sv.flush();
    // lambda, kept unchanged across observables as workspace 0
    void *lam0 = sv.workspace(0);
    hip_check(hipMemcpyAsync(lam0, sv.data(), sv.bytes(), hipMemcpyDeviceToDevice, sv.stream()), "memcpy");
    if (apply_operations) apply_forward(sv, ops, lam0);
    // Observables share a sweep in groups of up to three (lambda and their
    // H|lambda> in one program of up to four states, so the inverse gates
    // on lambda run once per group) when that program's specialised kernels
    // are ready; otherwise, and for the observables left over, one H|lambda>
    // per sweep re-runs it from a saved lambda. Each program is built and
    // planned on the host while the forward pass runs.
    struct Sweep {
        size_t m = 0;
        int exp_slots = 0, nslots = 0;
        std::vector<double> scales;
        StateVector::Planned planned;
    };
    auto make_sweep = [&](size_t m) {
        Sweep w;
        w.m = m;
        std::vector<LOp> prog;
        if (expvals)
            for (size_t j = 0; j < m; ++j) prog.push_back(identity_dot(0, 1 + static_cast<uint32_t>(j), w.exp_slots++));
        w.scales = backward_sweep(sv.n(), ops, tp, m, w.exp_slots, prog);
        w.nslots = w.exp_slots + static_cast<int>(ntp * m);
        w.planned = sv.plan(prog, 1 + static_cast<int>(m));
        return w;
    };
    Sweep single = make_sweep(1);
    Sweep group;
    const int grouping = tuning().adjoint_group.load();
    if (nobs > 1 && grouping > 0) {
        group = make_sweep(std::min<size_t>(3, nobs));
        // state traffic of the sweeps (passes x states each pass streams):
        // a group against one sweep per observable
        const double g_cost = static_cast<double>(group.planned.passes.size()) * static_cast<double>(1 + group.m);
        const double s_cost = static_cast<double>(single.planned.passes.size()) * 2.0 * static_cast<double>(group.m);
        const bool ready = group.planned.all_specialised();
        const bool take = ready && (grouping == 2 || g_cost < s_cost);
        if (tuning().adjoint_debug)
            std::fprintf(stderr, "[lq adjoint] observables %zu: group of %zu in %zu passes (specialised %d), single %zu "
                                 "passes: %s\n",
                         nobs, group.m, group.planned.passes.size(), ready ? 1 : 0, single.planned.passes.size(),
                         take ? "groups" : "one per sweep");
        if (!take) group = Sweep{};
    }
    std::vector<double> out(nobs * ntp, 0.0);
    for (size_t o0 = 0; o0 < nobs;) {
        const Sweep &w = (group.m > 0 && nobs - o0 >= group.m) ? group : single;
        const size_t m = w.m;
        // the sweep runs its inverse gates on lambda in place: groups with
        // more to follow work on a copy, the last one on lambda itself
        void *lam = lam0;
        if (o0 + m < nobs) {
            lam = sv.workspace(1);
            hip_check(hipMemcpyAsync(lam, lam0, sv.bytes(), hipMemcpyDeviceToDevice, sv.stream()), "memcpy");
        }
        std::vector<void *> states{lam};
        for (size_t j = 0; j < m; ++j) {
            void *h = sv.workspace(2 + static_cast<int>(j));
            sv.apply_obs_to(*obs[o0 + j], lam, h);
            states.push_back(h);
        }
        const auto v = sv.execute(w.planned, states, w.nslots);
        (m > 1 ? g_sweeps_grouped : g_sweeps_single).fetch_add(1, std::memory_order_relaxed);
        for (size_t j = 0; j < m; ++j) {
            if (expvals) expvals[o0 + j] = v[j].real();
            for (size_t t = 0; t < ntp; ++t)
                out[(o0 + j) * ntp + t] = -2.0 * w.scales[t] * v[w.exp_slots + t * m + j].imag();
        }
        o0 += m;
    }
    std::copy(out.begin(), out.end(), jac);
}

void vector_jacobian_product(StateVector &sv, const OpsList &ops, const void *dy_host, uint64_t len,
                             const std::vector<uint64_t> &tp_in, bool apply_operations, cd *out) {
    LQ_ABORT_IF(len != sv.length(), "Assertion failed: dy.size() == jd.getSizeStateVec()");
    std::vector<uint64_t> tp = tp_in;
    std::sort(tp.begin(), tp.end());
    const size_t ntp = tp.size();
    for (size_t t = 0; t < ntp; ++t) out[t] = cd(0.0, 0.0);
    if (ntp == 0) return;
    sv.flush();
    void *lam = sv.workspace(0);
    void *mu = sv.workspace(2);
    hip_check(hipMemcpyAsync(lam, sv.data(), sv.bytes(), hipMemcpyDeviceToDevice, sv.stream()), "memcpy");
    hip_check(hipMemcpyAsync(mu, dy_host, sv.bytes(), hipMemcpyHostToDevice, sv.stream()), "memcpy(dy)");
    if (apply_operations) apply_forward(sv, ops, lam);
    // planned on the host while the forward pass runs
    std::vector<LOp> prog;
    const std::vector<double> scales = backward_sweep(sv.n(), ops, tp, 1, 0, prog);
    const auto v = sv.execute(sv.plan(prog, 2), {lam, mu}, static_cast<int>(ntp));
    // w_t = i s <G mu | lambda> = i s <mu | G lambda>
    for (size_t t = 0; t < ntp; ++t) out[t] = cd(0.0, scales[t]) * v[t];
}

} // namespace lqamd
