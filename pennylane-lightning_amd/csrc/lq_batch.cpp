// Host side of the packed small-request path: program building (the same
// lowering and backward sweep as the engine's adjoint) and launches.
#include "lq_batch.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace lqamd {

namespace {

bool to_bop(const LOp &o, BatchProgram *p, std::string *why) {
    BOp b{};
    b.flags = o.is_dot() ? BOP_F_DOT : 0;
    b.smask = static_cast<uint8_t>(o.smask & 3u);
    b.cmask = static_cast<uint32_t>(o.cmask);
    b.cval = static_cast<uint32_t>(o.cval & o.cmask);
    if (o.is_dot()) {
        if (o.sa != 0 || o.sb != 1) {
            *why = "dot between unsupported states";
            return false;
        }
        b.arg2 = static_cast<uint32_t>(o.slot);
    }
    if (o.flags & TF_ZERO_UNCTRL) {
        *why = "generator applied to a state";
        return false;
    }
    b.arg = static_cast<uint32_t>(p->coef.size());
    switch (o.kind) {
    case TOP_DENSE1:
    case TOP_DENSE2:
    case TOP_DENSEK:
        if (o.k == 1) {
            b.kind = BOP_DENSE1;
            b.t0 = static_cast<uint8_t>(o.bits[0]);
        } else if (o.k == 2) {
            b.kind = BOP_DENSE2;
            b.t0 = static_cast<uint8_t>(o.bits[0]);
            b.t1 = static_cast<uint8_t>(o.bits[1]);
        } else {
            *why = "dense op on more than 2 qubits";
            return false;
        }
        break;
    case TOP_DIAG:
        if (o.k > 2) {
            *why = "diagonal op on more than 2 qubits";
            return false;
        }
        b.kind = BOP_DIAG;
        b.k = static_cast<uint8_t>(o.k);
        b.t0 = static_cast<uint8_t>(o.k > 0 ? o.bits[0] : 0);
        b.t1 = static_cast<uint8_t>(o.k > 1 ? o.bits[1] : 0);
        break;
    case TOP_PARITY:
        b.kind = BOP_PARITY;
        if (o.is_dot()) {
            *why = "parity dot";
            return false;
        }
        b.arg2 = static_cast<uint32_t>(o.pmask);
        break;
    default:
        *why = "op kind not in the packed kernel";
        return false;
    }
    for (const cd &c : o.coef) p->coef.push_back(make_double2(c.real(), c.imag()));
    p->ops.push_back(b);
    return true;
}

} // namespace

bool build_batch_program(int n, int dtype, const OpsList &ops, const Obs &ham, const std::vector<uint64_t> &tp_in,
                         BatchProgram *out, std::string *why) {
    if (n < 1 || n > batch_max_qubits(dtype)) {
        *why = "state too large for the packed kernel";
        return false;
    }
    std::vector<PauliWord> words;
    if (!obs_to_pauli(ham, n, &words)) {
        *why = "observable is not a Pauli sum";
        return false;
    }
    BatchProgram p;
    p.n = n;
    // forward from |0..0> on lambda
    std::vector<LOp> fwd;
    for (const auto &op : ops.ops) {
        if (op.named && op.gid == GateId::Identity) continue;
        if (op.name == "StatePrep" || op.name == "BasisState") {
            *why = "state preparation in the ops";
            return false;
        }
        LOp l = lower_record(n, op, false);
        l.smask = 1;
        if (!try_merge_1q(fwd, l)) fwd.push_back(std::move(l));
    }
    for (const auto &l : fwd)
        if (!to_bop(l, &p, why)) return false;
    // mu = H lambda
    BOp h{};
    h.kind = BOP_HAM;
    h.arg = 0;
    h.arg2 = static_cast<uint32_t>(words.size());
    for (const auto &w : words) {
        cd c = w.coeff;
        for (int i = 0; i < (w.ny & 3); ++i) c *= cd(0, -1);
        p.terms.push_back({w.x, w.yz, c.real(), c.imag()});
    }
    p.ops.push_back(h);
    // energy <mu|lambda> into slot 0
    BOp e{};
    e.kind = BOP_DOTID;
    e.flags = BOP_F_DOT;
    e.arg2 = 0;
    p.ops.push_back(e);
    // backward sweep: dots into slots 1 + t
    std::vector<uint64_t> tp = tp_in;
    std::sort(tp.begin(), tp.end());
    std::vector<LOp> bwd;
    if (!tp.empty()) p.scales = backward_sweep(n, ops, tp, 1, 1, bwd);
    for (const auto &l : bwd)
        if (!to_bop(l, &p, why)) return false;
    p.nslots = 1 + static_cast<uint32_t>(tp.size());
    *out = std::move(p);
    return true;
}

namespace {
// Initial sizes of a runner's buffers: 64 MiB of programs is ~2000 requests
// (a request's program is ~30-40 KiB), more than a server backlog of its
// default 4096-request cap reaches in practice; results are 16 bytes a slot.
constexpr size_t kRunnerInBytes = 64u << 20;
constexpr size_t kRunnerOutBytes = 1u << 20;
constexpr size_t kSectionAlign = 256; // each of the 5 sections starts 256-byte aligned
} // namespace

BatchRunner::BatchRunner() {
    // a hardware queue of its own: a CU-masked stream (every CU enabled) gets
    // a new queue instead of sharing one with other streams once the
    // process's queue pool is full, where it would wait behind their kernels
    int dev = 0, cus = 0;
    const char *oq = std::getenv("LQ_BATCH_OWN_QUEUE");
    const char *ku = std::getenv("LQ_BATCH_KERNEL_UPLOAD");
    kernel_upload_ = !ku || std::atoi(ku) != 0;
    if ((!oq || std::atoi(oq) != 0) && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) {
        std::vector<uint32_t> mask((static_cast<size_t>(cus) + 31) / 32, 0u);
        for (int c = 0; c < cus; ++c) mask[static_cast<size_t>(c) / 32] |= 1u << (c % 32);
        if (hipExtStreamCreateWithCUMask(&stream_, static_cast<uint32_t>(mask.size()), mask.data()) != hipSuccess)
            stream_ = nullptr;
    }
    if (!stream_) hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "hipEventCreate");
    for (hipEvent_t *e : {&t0_, &t_in_, &t_k_, &t_out_}) hip_check(hipEventCreate(e), "hipEventCreate");
    // programs and results move through kernels that read or write host
    // memory on this runner's queue: a copy-engine transfer waited in the
    // copy engine's queue behind other streams' transfers, themselves
    // waiting for those streams' kernels (uploads of ~1 MB and downloads of
    // a few KB took up to 1.3 s under load). Fine-grained (coherent) host
    // memory: the GPU does not cache it, so a launch reads what the host just
    // wrote and the host reads what the kernel wrote once the launch is done.
    if (kernel_upload_) hin_.flags = hout_.flags = hipHostMallocCoherent;
    hin_.ensure(kRunnerInBytes);
    dbuf_.ensure(kRunnerInBytes);
    hout_.ensure(kRunnerOutBytes);
}

size_t BatchRunner::program_bytes(const BatchProgram &p) {
    return sizeof(BReq) + p.ops.size() * sizeof(BOp) + p.coef.size() * sizeof(double2) +
           p.terms.size() * sizeof(PauliTerm) + p.nslots * sizeof(double2);
}

size_t BatchRunner::capacity() const {
    const size_t slack = 5 * kSectionAlign;
    const size_t in = std::min(hin_.bytes, dbuf_.bytes);
    return in > slack ? in - slack : 0;
}

BatchRunner::~BatchRunner() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    if (done_) (void)hipEventDestroy(done_);
    for (hipEvent_t e : {t0_, t_in_, t_k_, t_out_})
        if (e) (void)hipEventDestroy(e);
}

void BatchRunner::segments(double out3[3]) const {
    for (int i = 0; i < 3; ++i) out3[i] = seg_[i];
}

void BatchRunner::launch(const std::vector<const BatchProgram *> &progs, int dtype) {
    progs_ = progs;
    const size_t nreq = progs.size();
    if (nreq == 0) return;
    auto align = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    // LDS-resident requests first, then the ones working in global scratch
    const int lds_q = batch_lds_qubits(dtype);
    std::vector<size_t> order;
    for (size_t i = 0; i < nreq; ++i)
        if (progs[i]->n <= lds_q) order.push_back(i);
    const size_t n_lds = order.size();
    for (size_t i = 0; i < nreq; ++i)
        if (progs[i]->n > lds_q) order.push_back(i);
    size_t nops = 0, ncoef = 0, nterms = 0, nslots = 0, nscratch = 0;
    int max_lds_n = 1, max_n = 1;
    for (const auto *p : progs) {
        nops += p->ops.size();
        ncoef += p->coef.size();
        nterms += p->terms.size();
        nslots += p->nslots;
        max_n = std::max(max_n, p->n);
        if (p->n <= lds_q) max_lds_n = std::max(max_lds_n, p->n);
        else nscratch += 2ull << p->n;
    }
    LQ_ABORT_IF(max_n > batch_max_qubits(dtype), "request too large for the packed kernel");
    const size_t reqs_at = 0;
    const size_t ops_at = align(reqs_at + nreq * sizeof(BReq));
    const size_t coef_at = align(ops_at + nops * sizeof(BOp));
    const size_t terms_at = align(coef_at + ncoef * sizeof(double2));
    const size_t out_at = align(terms_at + nterms * sizeof(PauliTerm));
    const size_t total_in = out_at;
    const size_t total = out_at + nslots * sizeof(double2);
    // the previous launch on this runner has finished (finish() was called).
    // Growing a buffer frees the old one, and hipFree / hipHostFree wait for
    // the whole device, i.e. for every engine request in flight: the server
    // keeps its launches within capacity(), so these only grow for a single
    // program larger than the initial sizes.
    hin_.ensure(total_in);
    hout_.ensure(nslots * sizeof(double2));
    dbuf_.ensure(total);
    if (nscratch) scratch_.ensure(nscratch * (dtype ? 16 : 8));
    char *h = static_cast<char *>(hin_.p);
    BReq *rq = reinterpret_cast<BReq *>(h + reqs_at);
    BOp *op = reinterpret_cast<BOp *>(h + ops_at);
    double2 *cf = reinterpret_cast<double2 *>(h + coef_at);
    PauliTerm *tm = reinterpret_cast<PauliTerm *>(h + terms_at);
    out_off_.assign(nreq, 0);
    size_t o_ops = 0, o_coef = 0, o_terms = 0, o_out = 0, o_state = 0;
    for (size_t k = 0; k < nreq; ++k) {
        const size_t i = order[k];
        const BatchProgram &p = *progs[i];
        rq[k].n = static_cast<uint32_t>(p.n);
        rq[k].op_begin = static_cast<uint32_t>(o_ops);
        rq[k].op_end = static_cast<uint32_t>(o_ops + p.ops.size());
        rq[k].pad = 0;
        rq[k].out_off = o_out;
        rq[k].state_off = 0;
        if (p.n > lds_q) {
            rq[k].state_off = o_state;
            o_state += 2ull << p.n;
        }
        out_off_[i] = o_out;
        for (size_t j = 0; j < p.ops.size(); ++j) {
            BOp b = p.ops[j];
            if (b.kind == BOP_HAM) b.arg += static_cast<uint32_t>(o_terms);
            else if (b.kind != BOP_DOTID) b.arg += static_cast<uint32_t>(o_coef);
            op[o_ops + j] = b;
        }
        std::memcpy(cf + o_coef, p.coef.data(), p.coef.size() * sizeof(double2));
        std::memcpy(tm + o_terms, p.terms.data(), p.terms.size() * sizeof(PauliTerm));
        o_ops += p.ops.size();
        o_coef += p.coef.size();
        o_terms += p.terms.size();
        o_out += p.nslots;
    }
    out_slots_ = nslots;
    char *d = static_cast<char *>(dbuf_.p);
    hip_check(hipEventRecord(t0_, stream_), "hipEventRecord");
    void *hin_dev = nullptr;
    if (kernel_upload_ && hipHostGetDevicePointer(&hin_dev, hin_.p, 0) != hipSuccess) hin_dev = nullptr;
    if (hin_dev) {
        launch_copy16(hin_dev, d, total_in, stream_);
        hip_check(hipGetLastError(), "batch upload launch");
    } else {
        hip_check(hipMemcpyAsync(d, h, total_in, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync(batch)");
    }
    hip_check(hipEventRecord(t_in_, stream_), "hipEventRecord");
    const BReq *dr = reinterpret_cast<const BReq *>(d + reqs_at);
    const BOp *dop = reinterpret_cast<const BOp *>(d + ops_at);
    const double2 *dcf = reinterpret_cast<const double2 *>(d + coef_at);
    const PauliTerm *dtm = reinterpret_cast<const PauliTerm *>(d + terms_at);
    double2 *dout = reinterpret_cast<double2 *>(d + out_at);
    if (n_lds)
        launch_batch(dr, dop, dcf, dtm, dout, nullptr, static_cast<uint32_t>(n_lds), max_lds_n, true, dtype, stream_);
    if (nreq > n_lds)
        launch_batch(dr + n_lds, dop, dcf, dtm, dout, scratch_.p, static_cast<uint32_t>(nreq - n_lds), max_n, false,
                     dtype, stream_);
    hip_check(hipGetLastError(), "batch launch");
    hip_check(hipEventRecord(t_k_, stream_), "hipEventRecord");
    void *hout_dev = nullptr;
    if (hin_dev && hipHostGetDevicePointer(&hout_dev, hout_.p, 0) != hipSuccess) hout_dev = nullptr;
    if (hout_dev) {
        launch_copy16(d + out_at, hout_dev, nslots * sizeof(double2), stream_);
        hip_check(hipGetLastError(), "batch download launch");
    } else {
        hip_check(hipMemcpyAsync(hout_.p, d + out_at, nslots * sizeof(double2), hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync(batch results)");
    }
    hip_check(hipEventRecord(t_out_, stream_), "hipEventRecord");
    hip_check(hipEventRecord(done_, stream_), "hipEventRecord");
}

bool BatchRunner::done() {
    if (progs_.empty()) return true;
    const hipError_t e = hipEventQuery(done_);
    if (e == hipSuccess) return true;
    if (e == hipErrorNotReady) return false;
    hip_check(e, "hipEventQuery");
    return false;
}

void BatchRunner::finish(std::vector<double> *energies, std::vector<std::vector<double>> *grads) {
    energies->clear();
    grads->clear();
    if (progs_.empty()) return;
    hip_check(hipEventSynchronize(done_), "hipEventSynchronize");
    float ms[3] = {0.f, 0.f, 0.f};
    if (hipEventElapsedTime(&ms[0], t0_, t_in_) == hipSuccess && hipEventElapsedTime(&ms[1], t_in_, t_k_) == hipSuccess &&
        hipEventElapsedTime(&ms[2], t_k_, t_out_) == hipSuccess)
        for (int i = 0; i < 3; ++i) seg_[i] = ms[i];
    const double2 *r = static_cast<const double2 *>(hout_.p);
    for (size_t i = 0; i < progs_.size(); ++i) {
        const BatchProgram &p = *progs_[i];
        const double2 *s = r + out_off_[i];
        energies->push_back(s[0].x);
        std::vector<double> g(p.scales.size());
        for (size_t t = 0; t < g.size(); ++t) g[t] = -2.0 * p.scales[t] * s[1 + t].y;
        grads->push_back(std::move(g));
    }
    progs_.clear();
}

} // namespace lqamd
