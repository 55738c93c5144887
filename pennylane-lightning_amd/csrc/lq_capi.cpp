// extern "C" boundary (include/lq_amd.h) over the engine.
#include "lq_amd.h"

#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <memory>
#include <string>
#include <vector>

#include "lq_engine.hpp"
#include "lq_jit.hpp"
#include "lq_serve.hpp"

using namespace lqamd;

struct lq_sv_s {
    // shared with the DLPack exports of the state, which keep its device
    // memory alive after lq_sv_destroy until the consumer releases them
    std::shared_ptr<StateVector> sv;
};
struct lq_obs_s {
    std::shared_ptr<Obs> obs;
};
struct lq_rec_s {
    std::shared_ptr<StateVector> sv; // the graph writes this state
    std::unique_ptr<StateVector::Recording> rec;
};
struct lq_ops_s {
    OpsList list;
};

namespace {
thread_local std::string g_last_error;

template <typename F> int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return 1;
    } catch (...) {
        g_last_error = "unknown error";
        return 1;
    }
}

StateVector &S(lq_sv_t h) {
    LQ_ABORT_IF(h == nullptr || !h->sv, "null state vector handle");
    return *h->sv;
}

// DLPack tensor exchange, ABI of dlpack.h v0.8 (DLManagedTensor; the form
// torch.utils.dlpack.from_dlpack consumes): a 1-D compact tensor of 2^n
// complex amplitudes on a ROCm device.
struct DlDevice {
    int32_t device_type;
    int32_t device_id;
};
struct DlDataType {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
};
struct DlTensor {
    void *data;
    DlDevice device;
    int32_t ndim;
    DlDataType dtype;
    int64_t *shape;
    int64_t *strides;
    uint64_t byte_offset;
};
struct DlManagedTensor {
    DlTensor dl_tensor;
    void *manager_ctx;
    void (*deleter)(DlManagedTensor *);
};
constexpr int32_t kDlRocm = 10;
constexpr uint8_t kDlComplex = 5;

struct DlExport {
    DlManagedTensor managed;
    int64_t shape[1];
    std::shared_ptr<StateVector> keep;
};

void dl_export_delete(DlManagedTensor *t) { delete static_cast<DlExport *>(t->manager_ctx); }
} // namespace

extern "C" {

const char *lq_last_error(void) { return g_last_error.c_str(); }

const char *lq_backend_info(void) { return "lightning.qubit hot path on MI355X (gfx950, HIP tile programs)"; }

int lq_device_count(int *out) {
    return guard([&] { hip_check(hipGetDeviceCount(out), "hipGetDeviceCount"); });
}

int lq_sv_create(uint32_t num_qubits, int dtype, lq_sv_t *out) {
    return guard([&] {
        auto h = std::make_unique<lq_sv_s>();
        h->sv = std::make_shared<StateVector>(static_cast<int>(num_qubits), dtype);
        *out = h.release();
    });
}

int lq_sv_destroy(lq_sv_t sv) {
    return guard([&] { delete sv; });
}

int lq_sv_num_qubits(lq_sv_t sv, uint32_t *out) {
    return guard([&] { *out = static_cast<uint32_t>(S(sv).n()); });
}

int lq_sv_dtype(lq_sv_t sv, int *out) {
    return guard([&] { *out = S(sv).dtype(); });
}

int lq_sv_reset(lq_sv_t sv) {
    return guard([&] { S(sv).reset(); });
}

int lq_sv_set_basis_state(lq_sv_t sv, const uint32_t *bits, const uint32_t *wires, uint32_t nw) {
    return guard([&] { S(sv).set_basis_state(bits, wires, static_cast<int>(nw)); });
}

int lq_sv_set_state_vector(lq_sv_t sv, const void *state, const uint32_t *wires, uint32_t nw) {
    return guard([&] { S(sv).set_state_vector(state, wires, static_cast<int>(nw)); });
}

int lq_sv_update_data(lq_sv_t sv, const void *host, uint64_t len) {
    return guard([&] { S(sv).update_data(host, len); });
}

int lq_sv_get_data(lq_sv_t sv, void *host, uint64_t len) {
    return guard([&] { S(sv).get_data(host, len); });
}

int lq_sv_copy(lq_sv_t dst, lq_sv_t src) {
    return guard([&] { S(dst).copy_from(S(src)); });
}

int lq_sv_synchronize(lq_sv_t sv) {
    return guard([&] { S(sv).synchronize(); });
}

int lq_sv_device_ptr(lq_sv_t sv, void **out) {
    return guard([&] {
        // data() first: it may enqueue the write of a deferred reset
        *out = S(sv).data();
        S(sv).synchronize();
    });
}

int lq_sv_record(lq_sv_t sv, lq_ops_t ops, lq_rec_t *out) {
    return guard([&] {
        LQ_ABORT_IF(ops == nullptr || out == nullptr, "null argument");
        StateVector &s = S(sv);
        auto h = std::make_unique<lq_rec_s>();
        h->sv = sv->sv;
        h->rec = s.record(ops->list);
        *out = h.release();
    });
}

int lq_rec_replay(lq_rec_t rec, const double *params, uint64_t nparams) {
    return guard([&] {
        LQ_ABORT_IF(rec == nullptr || !rec->rec, "null recording handle");
        rec->sv->replay(*rec->rec, params, static_cast<size_t>(nparams));
    });
}

int lq_rec_stats(lq_rec_t rec, uint64_t *captures, uint64_t *replays) {
    return guard([&] {
        LQ_ABORT_IF(rec == nullptr || !rec->rec, "null recording handle");
        if (captures) *captures = rec->rec->captures;
        if (replays) *replays = rec->rec->replays;
    });
}

int lq_rec_destroy(lq_rec_t rec) {
    return guard([&] {
        if (rec == nullptr) return;
        // the graph may still run on the state's stream
        if (rec->sv) (void)hipStreamSynchronize(rec->sv->stream());
        delete rec;
    });
}

int lq_sv_update_data_device(lq_sv_t sv, const void *src, uint64_t len) {
    return guard([&] { S(sv).update_data_device(src, len); });
}

int lq_sv_get_data_device(lq_sv_t sv, void *dst, uint64_t len) {
    return guard([&] { S(sv).get_data_device(dst, len); });
}

int lq_sv_dlpack(lq_sv_t sv, void **out) {
    return guard([&] {
        LQ_ABORT_IF(out == nullptr, "null argument");
        StateVector &s = S(sv);
        void *p = s.data(); // executes a deferred reset's write first
        s.synchronize();
        hipPointerAttribute_t attr{};
        hip_check(hipPointerGetAttributes(&attr, p), "hipPointerGetAttributes");
        auto *e = new DlExport{};
        e->keep = sv->sv;
        e->shape[0] = static_cast<int64_t>(s.length());
        e->managed.dl_tensor.data = p;
        e->managed.dl_tensor.device = DlDevice{kDlRocm, attr.device};
        e->managed.dl_tensor.ndim = 1;
        e->managed.dl_tensor.dtype = DlDataType{kDlComplex, static_cast<uint8_t>(s.elem() * 8), 1};
        e->managed.dl_tensor.shape = e->shape;
        e->managed.dl_tensor.strides = nullptr;
        e->managed.dl_tensor.byte_offset = 0;
        e->managed.manager_ctx = e;
        e->managed.deleter = dl_export_delete;
        *out = &e->managed;
    });
}

int lq_dlpack_release(void *managed) {
    return guard([&] {
        auto *t = static_cast<DlManagedTensor *>(managed);
        if (t != nullptr && t->deleter != nullptr) t->deleter(t);
    });
}

int lq_sv_apply_op(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, const uint32_t *ctrls,
                   const uint8_t *ctrl_vals, uint32_t nc, int inverse, const double *params, uint32_t np) {
    return guard([&] {
        S(sv).apply_named(name, wires, static_cast<int>(nw), ctrls, ctrl_vals, static_cast<int>(nc), inverse != 0,
                          params, static_cast<int>(np));
    });
}

int lq_sv_apply_operations(lq_sv_t sv, uint32_t nops, const char *names, const uint32_t *wire_off,
                           const uint32_t *wires, const uint8_t *inverses, const uint32_t *param_off,
                           const double *params) {
    return guard([&] {
        LQ_ABORT_IF(nops && (!names || !wire_off || !inverses), "null argument");
        StateVector &s = S(sv);
        const char *nm = names;
        std::string name;
        for (uint32_t i = 0; i < nops; ++i) {
            const size_t len = std::strlen(nm);
            name.assign(nm, len);
            nm += len + 1;
            LQ_ABORT_IF(wire_off[i + 1] < wire_off[i], "invalid wire offsets");
            const uint32_t np = param_off ? param_off[i + 1] - param_off[i] : 0;
            LQ_ABORT_IF(np && !params, "null parameters");
            s.apply_named(name, wires + wire_off[i], static_cast<int>(wire_off[i + 1] - wire_off[i]), nullptr, nullptr,
                          0, inverses[i] != 0, np ? params + param_off[i] : nullptr, static_cast<int>(np));
        }
    });
}

int lq_sv_apply_matrix(lq_sv_t sv, const void *matrix, const uint32_t *wires, uint32_t nw, const uint32_t *ctrls,
                       const uint8_t *ctrl_vals, uint32_t nc, int inverse) {
    return guard([&] {
        S(sv).apply_matrix(matrix, wires, static_cast<int>(nw), ctrls, ctrl_vals, static_cast<int>(nc), inverse != 0);
    });
}

int lq_sv_apply_sparse_matrix(lq_sv_t sv, const int64_t *indptr, uint64_t indptr_len, const int64_t *indices,
                              const void *values, uint64_t nnz, const uint32_t *wires, uint32_t nw,
                              const uint32_t *ctrls, const uint8_t *ctrl_vals, uint32_t nc, int inverse) {
    return guard([&] {
        LQ_ABORT_IF(nw < 1, "Number of wires must be larger than 0");
        LQ_ABORT_IF(inverse != 0, "Inverse not implemented for sparse ops.");
        LQ_ABORT_IF(nw > 10, "sparse matrices are supported on at most 10 wires");
        const uint64_t d = 1ull << nw;
        LQ_ABORT_IF(indptr_len != d + 1, "The size of matrix does not match with the given number of wires");
        LQ_ABORT_IF(indptr[0] != 0, "sparse matrix offsets must start at 0");
        for (uint64_t i = 0; i < d; ++i)
            LQ_ABORT_IF(indptr[i + 1] < indptr[i], "sparse matrix offsets must not decrease");
        LQ_ABORT_IF(static_cast<uint64_t>(indptr[d]) != nnz, "sparse matrix offsets do not match the number of values");
        // the row-major dense form of the operator, in the state's precision
        StateVector &st = S(sv);
        const size_t el = st.elem();
        std::vector<char> dense(static_cast<size_t>(d * d) * el, 0);
        for (uint64_t r = 0; r < d; ++r)
            for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
                LQ_ABORT_IF(indices[k] < 0 || static_cast<uint64_t>(indices[k]) >= d,
                            "sparse matrix column index out of range");
                const size_t at = static_cast<size_t>(r * d + static_cast<uint64_t>(indices[k]));
                if (el == 16) {
                    auto *m = reinterpret_cast<std::complex<double> *>(dense.data());
                    m[at] += static_cast<const std::complex<double> *>(values)[k];
                } else {
                    auto *m = reinterpret_cast<std::complex<float> *>(dense.data());
                    m[at] += static_cast<const std::complex<float> *>(values)[k];
                }
            }
        st.apply_matrix(dense.data(), wires, static_cast<int>(nw), ctrls, ctrl_vals, static_cast<int>(nc), false);
    });
}

int lq_sv_apply_pauli_rot(lq_sv_t sv, const uint32_t *wires, uint32_t nw, int inverse, double theta,
                          const char *word) {
    return guard([&] { S(sv).apply_pauli_rot(wires, static_cast<int>(nw), inverse != 0, theta, word); });
}

int lq_sv_apply_generator(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, const uint32_t *ctrls,
                          const uint8_t *ctrl_vals, uint32_t nc, int adj, double *scale) {
    (void)adj; // the generator does not depend on adj (GateImplementationsLM.hpp:2181-2952)
    return guard([&] {
        *scale = S(sv).apply_generator(name, wires, static_cast<int>(nw), ctrls, ctrl_vals, static_cast<int>(nc));
    });
}

int lq_sv_normalize(lq_sv_t sv) {
    return guard([&] { S(sv).normalize(); });
}

int lq_sv_norm2(lq_sv_t sv, double *out) {
    return guard([&] {
        LQ_ABORT_IF(!out, "null argument");
        *out = S(sv).norm2();
    });
}

int lq_sv_collapse(lq_sv_t sv, uint32_t wire, int branch) {
    return guard([&] { S(sv).collapse(wire, branch != 0); });
}

int lq_probs(lq_sv_t sv, const uint32_t *wires, uint32_t nw, int all, double *out) {
    return guard([&] {
        const auto p = S(sv).probs(wires, static_cast<int>(nw), all != 0);
        std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
}

int lq_expval_named(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, double *out) {
    return guard([&] { *out = S(sv).expval_named(name, wires, static_cast<int>(nw)); });
}

int lq_expval_named_batch(lq_sv_t sv, uint32_t count, const char *const *names, const uint32_t *wires,
                          double *out) {
    return guard([&] {
        std::vector<std::string> nm(names, names + count);
        std::vector<uint32_t> w(wires, wires + count);
        const auto v = S(sv).expval_named_batch(nm, w);
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}

int lq_expval_matrix(lq_sv_t sv, const void *matrix, const uint32_t *wires, uint32_t nw, double *out) {
    return guard([&] { *out = S(sv).expval_matrix(matrix, wires, static_cast<int>(nw)); });
}

int lq_expval_obs(lq_sv_t sv, lq_obs_t obs, double *out) {
    return guard([&] {
        LQ_ABORT_IF(!obs, "null observable");
        *out = S(sv).expval_obs(*obs->obs);
    });
}

int lq_var_named(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, double *out) {
    return guard([&] { *out = S(sv).var_named(name, wires, static_cast<int>(nw)); });
}

int lq_var_matrix(lq_sv_t sv, const void *matrix, const uint32_t *wires, uint32_t nw, double *out) {
    return guard([&] { *out = S(sv).var_matrix(matrix, wires, static_cast<int>(nw)); });
}

int lq_var_obs(lq_sv_t sv, lq_obs_t obs, double *out) {
    return guard([&] {
        LQ_ABORT_IF(!obs, "null observable");
        *out = S(sv).var_obs(*obs->obs);
    });
}

namespace {
CsrView csr_args(const int64_t *indptr, uint64_t indptr_len, const int64_t *indices, const void *values, uint64_t nnz) {
    LQ_ABORT_IF(indptr_len < 1 || !indptr || (nnz && (!indices || !values)), "invalid sparse matrix");
    return CsrView{indptr, indptr_len - 1, indices, values, nnz};
}
} // namespace

int lq_expval_sparse(lq_sv_t sv, const int64_t *indptr, uint64_t indptr_len, const int64_t *indices,
                     const void *values, uint64_t nnz, double *out) {
    return guard([&] { *out = S(sv).expval_sparse(csr_args(indptr, indptr_len, indices, values, nnz)); });
}

int lq_var_sparse(lq_sv_t sv, const int64_t *indptr, uint64_t indptr_len, const int64_t *indices,
                  const void *values, uint64_t nnz, double *out) {
    return guard([&] { *out = S(sv).var_sparse(csr_args(indptr, indptr_len, indices, values, nnz)); });
}

int lq_generate_samples(lq_sv_t sv, const uint32_t *wires, uint32_t nw, int all, uint64_t shots, int seeded,
                        uint64_t seed, uint64_t *out) {
    return guard([&] {
        S(sv).generate_samples(wires, static_cast<int>(nw), all != 0, shots, seeded != 0, seed, out);
    });
}

int lq_obs_sparse(const void *values_c128, const int64_t *indices, uint64_t nnz, const int64_t *offsets,
                  uint64_t offsets_len, const uint32_t *wires, uint32_t nw, lq_obs_t *out) {
    return guard([&] {
        LQ_ABORT_IF(offsets_len < 1 || !offsets || (nnz && (!indices || !values_c128)), "invalid sparse matrix");
        auto o = std::make_shared<Obs>();
        o->kind = Obs::Sparse;
        o->wires.assign(wires, wires + nw);
        const auto *v = static_cast<const cd *>(values_c128);
        o->sp_values.assign(v, v + nnz);
        o->sp_indices.assign(indices, indices + nnz);
        o->sp_indptr.assign(offsets, offsets + offsets_len);
        *out = new lq_obs_s{o};
    });
}

int lq_obs_named(const char *name, const uint32_t *wires, uint32_t nw, lq_obs_t *out) {
    return guard([&] {
        GateId g;
        const std::string nm(name);
        LQ_ABORT_IF(!gate_from_name(nm, &g), (std::string("Operation does not exist for ") + nm).c_str());
        LQ_ABORT_IF(gate_info(g).nwires > 0 && static_cast<int>(nw) != gate_info(g).nwires,
                    "wrong number of wires for the named observable");
        auto o = std::make_shared<Obs>();
        o->kind = Obs::Named;
        o->name = nm;
        o->wires.assign(wires, wires + nw);
        *out = new lq_obs_s{o};
    });
}

int lq_obs_hermitian(const void *matrix_c128, const uint32_t *wires, uint32_t nw, lq_obs_t *out) {
    return guard([&] {
        LQ_ABORT_IF(nw < 1 || nw > 30, "Hermitian observables are supported on 1 to 30 wires");
        auto o = std::make_shared<Obs>();
        o->kind = Obs::Hermitian;
        o->wires.assign(wires, wires + nw);
        const size_t d = static_cast<size_t>(1) << nw;
        const auto *m = static_cast<const cd *>(matrix_c128);
        o->matrix.assign(m, m + d * d);
        *out = new lq_obs_s{o};
    });
}

int lq_obs_tensor(const lq_obs_t *factors, uint32_t count, lq_obs_t *out) {
    return guard([&] {
        auto o = std::make_shared<Obs>();
        o->kind = Obs::Tensor;
        std::vector<uint32_t> seen;
        for (uint32_t i = 0; i < count; ++i) {
            LQ_ABORT_IF(!factors[i], "null observable");
            for (auto w : factors[i]->obs->all_wires()) {
                for (auto s : seen) LQ_ABORT_IF(s == w, "All wires in observables must be disjoint.");
                seen.push_back(w);
            }
            o->terms.push_back(factors[i]->obs);
        }
        LQ_ABORT_IF(count == 1 && factors[0]->obs->kind == Obs::Tensor,
                    "A new TensorProdObsBase observable cannot be created from a single TensorProdObsBase.");
        *out = new lq_obs_s{o};
    });
}

int lq_obs_hamiltonian(const double *coeffs, const lq_obs_t *terms, uint32_t count, lq_obs_t *out) {
    return guard([&] {
        auto o = std::make_shared<Obs>();
        o->kind = Obs::Ham;
        for (uint32_t i = 0; i < count; ++i) {
            LQ_ABORT_IF(!terms[i], "null observable");
            o->terms.push_back(terms[i]->obs);
            o->coeffs.push_back(coeffs[i]);
        }
        *out = new lq_obs_s{o};
    });
}

int lq_obs_destroy(lq_obs_t obs) {
    return guard([&] { delete obs; });
}

int lq_ops_create(uint32_t nops, const char *const *names, const uint32_t *param_off, const double *params,
                  const uint32_t *wire_off, const uint32_t *wires, const uint8_t *inverses, const uint32_t *mat_off,
                  const double *matrices, const uint32_t *ctrl_off, const uint32_t *ctrls, const uint8_t *ctrl_vals,
                  lq_ops_t *out) {
    return guard([&] {
        auto h = std::make_unique<lq_ops_s>();
        for (uint32_t i = 0; i < nops; ++i) {
            OpRec r;
            r.name = names[i];
            r.named = gate_from_name(r.name, &r.gid);
            r.params.assign(params + param_off[i], params + param_off[i + 1]);
            r.wires.assign(wires + wire_off[i], wires + wire_off[i + 1]);
            r.inverse = inverses[i] != 0;
            for (uint32_t j = mat_off[i]; j < mat_off[i + 1]; ++j) r.matrix.emplace_back(matrices[2 * j], matrices[2 * j + 1]);
            r.ctrls.assign(ctrls + ctrl_off[i], ctrls + ctrl_off[i + 1]);
            r.cvals.assign(ctrl_vals + ctrl_off[i], ctrl_vals + ctrl_off[i + 1]);
            if (!r.params.empty()) ++h->list.num_par_ops;
            h->list.ops.push_back(std::move(r));
        }
        *out = h.release();
    });
}

int lq_ops_destroy(lq_ops_t ops) {
    return guard([&] { delete ops; });
}

int lq_ops_num_par_ops(lq_ops_t ops, uint64_t *out) {
    return guard([&] { *out = ops->list.num_par_ops; });
}

int lq_sv_apply_ops(lq_sv_t sv, lq_ops_t ops, int adjoint) {
    return guard([&] {
        LQ_ABORT_IF(!ops, "null ops");
        if (adjoint) {
            for (auto it = ops->list.ops.rbegin(); it != ops->list.ops.rend(); ++it) S(sv).apply_oprec(*it, true);
        } else {
            for (const auto &op : ops->list.ops) S(sv).apply_oprec(op, false);
        }
    });
}

int lq_adjoint_jacobian(lq_sv_t sv, lq_ops_t ops, const lq_obs_t *obs, uint32_t nobs, const uint64_t *trainable,
                        uint32_t ntp, int apply_operations, double *jac, double *expvals) {
    return guard([&] {
        LQ_ABORT_IF(!ops, "null ops");
        std::vector<std::shared_ptr<Obs>> o;
        for (uint32_t i = 0; i < nobs; ++i) {
            LQ_ABORT_IF(!obs[i], "null observable");
            o.push_back(obs[i]->obs);
        }
        std::vector<uint64_t> tp(trainable, trainable + ntp);
        adjoint_jacobian(S(sv), ops->list, o, tp, apply_operations != 0, jac, expvals);
    });
}

int lq_vector_jacobian_product(lq_sv_t sv, lq_ops_t ops, const void *dy, uint64_t len, const uint64_t *trainable,
                               uint32_t ntp, int apply_operations, double *out) {
    return guard([&] {
        LQ_ABORT_IF(!ops, "null ops");
        LQ_ABORT_IF(!dy, "null cotangent");
        std::vector<uint64_t> tp(trainable, trainable + ntp);
        std::vector<cd> w(ntp);
        vector_jacobian_product(S(sv), ops->list, dy, len, tp, apply_operations != 0, w.data());
        for (uint32_t t = 0; t < ntp; ++t) {
            out[2 * t] = w[t].real();
            out[2 * t + 1] = w[t].imag();
        }
    });
}

int lq_sv_profile(lq_sv_t sv, int enable) {
    return guard([&] { S(sv).profile(enable != 0); });
}

int lq_sv_profile_read(lq_sv_t sv, int reset, uint64_t *passes, double *ms, double *bytes, double *max_ms) {
    return guard([&] {
        const auto st = S(sv).profile_read(reset != 0);
        if (passes) *passes = st.passes;
        if (ms) *ms = st.ms;
        if (bytes) *bytes = st.bytes;
        if (max_ms) *max_ms = st.max_ms;
    });
}

int lq_sv_profile_passes(lq_sv_t sv, double *out, uint64_t cap, uint64_t *count) {
    return guard([&] {
        const auto recs = S(sv).profile_passes();
        if (count) *count = recs.size();
        LQ_ABORT_IF(cap && !out, "null argument");
        for (size_t i = 0; i < recs.size() && i < cap; ++i) {
            const auto &r = recs[i];
            double *o = out + 7 * i;
            o[0] = r.ms;
            o[1] = r.bytes;
            o[2] = r.ops;
            o[3] = r.stages;
            o[4] = r.transposes;
            o[5] = r.specialised;
            o[6] = static_cast<double>(r.tile_bits);
        }
    });
}

int lq_sv_stream(lq_sv_t sv, void **out) {
    return guard([&] { *out = static_cast<void *>(S(sv).stream()); });
}

int lq_sv_overlap(lq_sv_t a, lq_sv_t b, double *re, double *im) {
    return guard([&] {
        const cd v = S(a).overlap(S(b));
        *re = v.real();
        *im = v.imag();
    });
}

int lq_sv_release_workspace(lq_sv_t sv) {
    return guard([&] { S(sv).release_workspace(); });
}

struct lq_server_s {
    std::unique_ptr<lqamd::Server> srv;
};

int lq_server_create(const lq_server_config *cfg, lq_server_t *out) {
    return guard([&] {
        LQ_ABORT_IF(!cfg || !out, "null argument");
        LQ_ABORT_IF(cfg->dtype != LQ_C64 && cfg->dtype != LQ_C128, "invalid dtype");
        lqamd::ServerConfig c;
        c.dtype = cfg->dtype;
        c.batched = cfg->batched;
        if (cfg->hbm_budget) c.hbm_budget = cfg->hbm_budget;
        c.small_max_qubits = cfg->small_max_qubits;
        if (cfg->max_batch > 0) c.max_batch = cfg->max_batch;
        if (cfg->batch_streams > 0) c.batch_streams = cfg->batch_streams;
        if (cfg->large_workers > 0) c.large_workers = cfg->large_workers;
        if (cfg->prep_threads > 0) c.prep_threads = cfg->prep_threads;
        if (cfg->lane_workers != 0) c.lane_workers = std::max(0, cfg->lane_workers);
        if (cfg->lane_max_qubits > 0) c.lane_max_qubits = cfg->lane_max_qubits;
        auto *h = new lq_server_s;
        h->srv = std::make_unique<lqamd::Server>(c);
        *out = h;
    });
}

int lq_server_destroy(lq_server_t srv) {
    return guard([&] { delete srv; });
}

int lq_server_submit(lq_server_t srv, uint32_t num_qubits, lq_ops_t ops, lq_obs_t ham, const uint64_t *trainable,
                     uint32_t ntp, uint64_t *id) {
    return guard([&] {
        LQ_ABORT_IF(!srv || !ops || !ham || !id, "null argument");
        std::vector<uint64_t> tp(trainable, trainable + ntp);
        *id = srv->srv->submit(static_cast<int>(num_qubits), ops->list, ham->obs, tp);
    });
}

int lq_server_wait(lq_server_t srv, uint64_t id, double timeout_s, int *done, double *energy, double *grad,
                   uint32_t ntp, double *latency_ms, int *path) {
    return guard([&] {
        LQ_ABORT_IF(!srv || !done, "null argument");
        lqamd::ServeResult r;
        *done = srv->srv->wait(id, timeout_s, &r) ? 1 : 0;
        if (!*done) return;
        LQ_ABORT_IF(!r.error.empty(), r.error.c_str());
        if (energy) *energy = r.energy;
        if (grad) {
            LQ_ABORT_IF(r.grad.size() != ntp, "gradient size mismatch");
            std::copy(r.grad.begin(), r.grad.end(), grad);
        }
        if (latency_ms) *latency_ms = r.latency_ms;
        if (path) *path = r.path;
    });
}

int lq_server_replay(lq_server_t srv, uint32_t count, const uint32_t *n, const lq_ops_t *ops, const lq_obs_t *ham,
                     const uint64_t *tp, const uint64_t *tp_off, const double *arrival_s, double *energies,
                     double *grads, double *latency_ms, int *path, double *wall_s) {
    return guard([&] {
        LQ_ABORT_IF(!srv, "null server");
        std::vector<uint64_t> ids(count, 0);
        std::atomic<uint32_t> submitted{0};
        std::string err;
        std::mutex err_mu;
        using Clock = std::chrono::steady_clock;
        const auto start = Clock::now();
        Clock::time_point last_done = start;
        std::thread collector([&] {
          try {
            for (uint32_t i = 0; i < count; ++i) {
                while (submitted.load(std::memory_order_acquire) <= i) {
                    {
                        std::lock_guard<std::mutex> lk(err_mu);
                        if (!err.empty()) return;
                    }
                    std::this_thread::sleep_for(std::chrono::microseconds(50));
                }
                lqamd::ServeResult r;
                while (!srv->srv->wait(ids[i], 1.0, &r)) {
                }
                if (!r.error.empty()) {
                    std::lock_guard<std::mutex> lk(err_mu);
                    if (err.empty()) err = r.error;
                }
                energies[i] = r.energy;
                std::copy(r.grad.begin(), r.grad.end(), grads + tp_off[i]);
                latency_ms[i] = r.latency_ms;
                path[i] = r.path;
                last_done = Clock::now();
            }
          } catch (const std::exception &ex) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (err.empty()) err = ex.what();
          }
        });
        try {
            for (uint32_t i = 0; i < count; ++i) {
                std::this_thread::sleep_until(start + std::chrono::duration_cast<Clock::duration>(
                                                          std::chrono::duration<double>(arrival_s[i])));
                std::vector<uint64_t> t(tp + tp_off[i], tp + tp_off[i + 1]);
                ids[i] = srv->srv->submit(static_cast<int>(n[i]), ops[i]->list, ham[i]->obs, t);
                submitted.store(i + 1, std::memory_order_release);
            }
        } catch (const std::exception &ex) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (err.empty()) err = ex.what();
        }
        collector.join();
        LQ_ABORT_IF(!err.empty(), err.c_str());
        *wall_s = std::chrono::duration<double>(last_done - (start + std::chrono::duration_cast<Clock::duration>(
                                                                         std::chrono::duration<double>(arrival_s[0]))))
                      .count();
    });
}

int lq_server_stats(lq_server_t srv, uint64_t *out6) {
    return guard([&] {
        const auto s = srv->srv->stats();
        out6[0] = s.submitted;
        out6[1] = s.completed;
        out6[2] = s.packed;
        out6[3] = s.engine;
        out6[4] = s.launches;
        out6[5] = s.hbm_in_use;
    });
}

int lq_server_pool_stats(lq_server_t srv, uint64_t *made, uint64_t *freed) {
    return guard([&] {
        const auto s = srv->srv->stats();
        if (made) *made = s.states_made;
        if (freed) *freed = s.states_freed;
    });
}

int lq_server_timing(lq_server_t srv, double *out6, int reset) {
    return guard([&] {
        LQ_ABORT_IF(!srv || !out6, "null argument");
        const auto s = srv->srv->stats();
        out6[0] = s.ready_wait_max_ms;
        out6[1] = s.launch_call_max_ms;
        out6[2] = s.launch_max_ms;
        for (int j = 0; j < 3; ++j) out6[3 + j] = s.slow_launch_seg_ms[j];
        if (reset) srv->srv->reset_timing();
    });
}

int lq_jit_wait(void) {
    return guard([&] { jit_wait(); });
}

int lq_jit_shutdown(void) {
    return guard([&] { jit_shutdown(); });
}

int lq_jit_stats(int *ready, int *failed, int *pending) {
    return guard([&] { jit_stats(ready, failed, pending); });
}

int lq_sv_pass_counts(lq_sv_t sv, int reset, uint64_t *specialised, uint64_t *generic) {
    return guard([&] { S(sv).pass_counts(reset != 0, specialised, generic); });
}

int lq_set_split_fallback(int mode, int *previous) {
    return guard([&] {
        LQ_ABORT_IF(mode < 0 || mode > 2, "split fallback mode must be 0, 1 or 2");
        const int old = tuning().split_fallback.exchange(mode);
        if (previous) *previous = old;
    });
}

int lq_set_engine_option(const char *name, int value, int *previous) {
    return guard([&] {
        LQ_ABORT_IF(name == nullptr, "null argument");
        const std::string k(name);
        EngineTuning &t = tuning();
        std::atomic<int> *opt = nullptr;
        if (k == "split_fallback") {
            LQ_ABORT_IF(value < 0 || value > 2, "split fallback mode must be 0, 1 or 2");
            opt = &t.split_fallback;
        } else if (k == "mfma_dense5") {
            opt = &t.mfma_dense5;
        } else if (k == "product_fold") {
            opt = &t.product_fold;
        } else if (k == "plan_search") {
            opt = &t.plan_search;
        } else if (k == "adjoint_group") {
            LQ_ABORT_IF(value < 0 || value > 2, "adjoint_group must be 0, 1 or 2");
            opt = &t.adjoint_group;
        } else {
            LQ_ABORT((std::string("unknown engine option ") + k).c_str());
        }
        if (k != "split_fallback" && k != "adjoint_group") value = value ? 1 : 0;
        const int old = opt->exchange(value);
        if (previous) *previous = old;
    });
}

int lq_adjoint_sweep_counts(int reset, uint64_t *grouped, uint64_t *single) {
    return guard([&] { adjoint_sweep_counts(reset != 0, grouped, single); });
}

int lq_probs_noalloc_init(void) {
    return guard([&] { probs_noalloc_init(); });
}

int lq_probs_noalloc(const void *state_dev, uint32_t num_qubits, int dtype, const uint32_t *wires, uint32_t nw,
                     double *out_dev, void *stream) {
    return guard([&] {
        probs_noalloc(state_dev, static_cast<int>(num_qubits), dtype, wires, static_cast<int>(nw), out_dev,
                      static_cast<hipStream_t>(stream));
    });
}

int lq_set_mfma_dense5(int on, int *previous) {
    return guard([&] {
        const int old = tuning().mfma_dense5.exchange(on ? 1 : 0);
        if (previous) *previous = old;
    });
}

} // extern "C"
