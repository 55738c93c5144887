// Test infrastructure only: a C-ABI driver over the reference lightning.qubit
// C++ classes, compiled from the reference sources in place (see
// oracle/Makefile) into oracle/_ref/liblq_ref.so. It is used by tests/ to pin
// the numpy oracle against the real reference, by tests/golden/make_golden.py
// to write golden fixtures, and by bench.py's cpu_baseline leg. It is never
// part of the product path.
//
// Reference classes driven here:
//   StateVectorLQubitManaged<T>   lq/StateVectorLQubitManaged.hpp:52-188
//   applyOperation / applyMatrix  lq/StateVectorLQubit.hpp:378-710
//   applyGenerator                lq/StateVectorLQubit.hpp:527-560
//   Measurements<SV>              lq/measurements/MeasurementsLQubit.hpp:72-768
//                                 (expval, var, sparse expval/var, generate_samples)
//   NamedObs / TensorProdObs / HermitianObs / Hamiltonian
//                                 lq/observables/ObservablesLQubit.hpp:52-366
//   OpsData / JacobianData        core/algorithms/JacobianData.hpp:39-393
//   AdjointJacobian               lq/algorithms/AdjointJacobianLQubit.hpp:347-491
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <sstream>
#include <span>
#include <string>
#include <vector>

#include "AdjointJacobianLQubit.hpp"
#include "JacobianData.hpp"
#include "MeasurementsLQubit.hpp"
#include "ObservablesLQubit.hpp"
#include "StateVectorLQubitManaged.hpp"
#include "VectorJacobianProduct.hpp"

using namespace Pennylane::LightningQubit;
using namespace Pennylane::LightningQubit::Measures;
using namespace Pennylane::LightningQubit::Observables;
using namespace Pennylane::LightningQubit::Algorithms;
using Pennylane::Algorithms::JacobianData;
using Pennylane::Algorithms::OpsData;

namespace {
thread_local std::string g_err;

template <class T> using SV = StateVectorLQubitManaged<T>;

template <class T>
std::vector<std::size_t> to_wires(const uint32_t *w, uint32_t n) {
    return std::vector<std::size_t>(w, w + n);
}
std::vector<bool> to_bools(const uint8_t *v, uint32_t n) {
    std::vector<bool> out(n);
    for (uint32_t i = 0; i < n; i++) out[i] = v[i] != 0;
    return out;
}

// Observable text spec, whitespace separated tokens:
//   obs  := "N" name wire                       NamedObs
//         | "T" k (name wire){k}                TensorProdObs of NamedObs
//         | "M" k wire{k} (re im){4^k}          HermitianObs
//         | "H" m (coeff obs){m}                Hamiltonian
template <class T>
std::shared_ptr<Observable<SV<T>>> parse_obs(std::istringstream &in) {
    using C = std::complex<T>;
    std::string tag;
    in >> tag;
    if (tag == "N") {
        std::string name;
        std::size_t w;
        in >> name >> w;
        return std::make_shared<NamedObs<SV<T>>>(name, std::vector<std::size_t>{w});
    }
    if (tag == "T") {
        std::size_t k;
        in >> k;
        std::vector<std::shared_ptr<Observable<SV<T>>>> f;
        for (std::size_t i = 0; i < k; i++) {
            std::string name;
            std::size_t w;
            in >> name >> w;
            f.push_back(std::make_shared<NamedObs<SV<T>>>(name, std::vector<std::size_t>{w}));
        }
        return TensorProdObs<SV<T>>::create(f);
    }
    if (tag == "M") {
        std::size_t k;
        in >> k;
        std::vector<std::size_t> wires(k);
        for (auto &w : wires) in >> w;
        std::vector<C> mat(std::size_t{1} << (2 * k));
        for (auto &m : mat) {
            double re, im;
            in >> re >> im;
            m = C(static_cast<T>(re), static_cast<T>(im));
        }
        return std::make_shared<HermitianObs<SV<T>>>(mat, wires);
    }
    if (tag == "H") {
        std::size_t m;
        in >> m;
        std::vector<T> coeffs(m);
        std::vector<std::shared_ptr<Observable<SV<T>>>> terms;
        for (std::size_t i = 0; i < m; i++) {
            double c;
            in >> c;
            coeffs[i] = static_cast<T>(c);
            terms.push_back(parse_obs<T>(in));
        }
        return std::make_shared<Hamiltonian<SV<T>>>(coeffs, terms);
    }
    throw std::runtime_error("bad observable spec tag: " + tag);
}

// Flat ops list, mirroring create_ops_list (core/bindings/Bindings.hpp:772-799).
template <class T>
OpsData<SV<T>> make_ops(uint32_t nops, const char *const *names,
                        const uint32_t *param_off, const double *params,
                        const uint32_t *wire_off, const uint32_t *wires,
                        const uint8_t *inverses, const uint32_t *mat_off,
                        const double *mats, const uint32_t *ctrl_off,
                        const uint32_t *ctrls, const uint8_t *ctrl_vals) {
    using C = std::complex<T>;
    std::vector<std::string> n;
    std::vector<std::vector<T>> p;
    std::vector<std::vector<std::size_t>> w, cw;
    std::vector<bool> inv;
    std::vector<std::vector<C>> m;
    std::vector<std::vector<bool>> cv;
    for (uint32_t i = 0; i < nops; i++) {
        n.emplace_back(names[i]);
        std::vector<T> pi;
        for (uint32_t j = param_off[i]; j < param_off[i + 1]; j++) pi.push_back(static_cast<T>(params[j]));
        p.push_back(pi);
        w.emplace_back(wires + wire_off[i], wires + wire_off[i + 1]);
        inv.push_back(inverses[i] != 0);
        std::vector<C> mi;
        for (uint32_t j = mat_off[i]; j < mat_off[i + 1]; j++)
            mi.emplace_back(static_cast<T>(mats[2 * j]), static_cast<T>(mats[2 * j + 1]));
        m.push_back(mi);
        cw.emplace_back(ctrls + ctrl_off[i], ctrls + ctrl_off[i + 1]);
        std::vector<bool> cvi;
        for (uint32_t j = ctrl_off[i]; j < ctrl_off[i + 1]; j++) cvi.push_back(ctrl_vals[j] != 0);
        cv.push_back(cvi);
    }
    return OpsData<SV<T>>(n, p, w, inv, m, cw, cv);
}

#define GUARD(body)                                                            \
    try {                                                                      \
        body;                                                                  \
        return 0;                                                              \
    } catch (const std::exception &e) {                                        \
        g_err = e.what();                                                      \
        return 1;                                                              \
    }

template <class T> SV<T> &sv(void *h) { return *static_cast<SV<T> *>(h); }
} // namespace

#define DISPATCH(prec, call)                                                   \
    if (prec == 1) {                                                           \
        using T = double;                                                      \
        call;                                                                  \
    } else {                                                                   \
        using T = float;                                                       \
        call;                                                                  \
    }

extern "C" {

const char *ref_last_error(void) { return g_err.c_str(); }

void *ref_sv_new(uint32_t nq, int prec) {
    try {
        if (prec == 1) return new SV<double>(nq);
        return new SV<float>(nq);
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_sv_free(void *h, int prec) {
    if (prec == 1) delete static_cast<SV<double> *>(h);
    else delete static_cast<SV<float> *>(h);
}

int ref_sv_set(void *h, int prec, const void *data, uint64_t len) {
    GUARD(DISPATCH(prec, sv<T>(h).updateData(static_cast<const std::complex<T> *>(data), len)))
}

int ref_sv_get(void *h, int prec, void *out, uint64_t len) {
    GUARD(DISPATCH(prec, {
        auto &s = sv<T>(h);
        if (len != s.getLength()) throw std::runtime_error("length mismatch");
        std::memcpy(out, s.getData(), len * sizeof(std::complex<T>));
    }))
}

int ref_apply(void *h, int prec, const char *name, const uint32_t *wires,
              uint32_t nw, const uint32_t *ctrls, const uint8_t *cvals,
              uint32_t nc, int inverse, const double *params, uint32_t np) {
    GUARD(DISPATCH(prec, {
        std::vector<T> p(params, params + np);
        if (nc == 0)
            sv<T>(h).applyOperation(name, to_wires<T>(wires, nw), inverse != 0, p);
        else
            sv<T>(h).applyOperation(name, to_wires<T>(ctrls, nc), to_bools(cvals, nc),
                                    to_wires<T>(wires, nw), inverse != 0, p);
    }))
}

int ref_apply_matrix(void *h, int prec, const void *mat, const uint32_t *wires,
                     uint32_t nw, const uint32_t *ctrls, const uint8_t *cvals,
                     uint32_t nc, int inverse) {
    GUARD(DISPATCH(prec, {
        const auto *m = static_cast<const std::complex<T> *>(mat);
        std::vector<std::complex<T>> mv(m, m + (std::size_t{1} << (2 * nw)));
        if (nc == 0)
            sv<T>(h).applyMatrix(mv, to_wires<T>(wires, nw), inverse != 0);
        else
            sv<T>(h).applyControlledMatrix(mv, to_wires<T>(ctrls, nc), to_bools(cvals, nc),
                                           to_wires<T>(wires, nw), inverse != 0);
    }))
}

int ref_apply_pauli_rot(void *h, int prec, const uint32_t *wires, uint32_t nw,
                        int inverse, double theta, const char *word) {
    GUARD(DISPATCH(prec, sv<T>(h).applyPauliRot(to_wires<T>(wires, nw), inverse != 0,
                                                 std::vector<T>{static_cast<T>(theta)}, word)))
}

int ref_apply_generator(void *h, int prec, const char *name,
                        const uint32_t *wires, uint32_t nw,
                        const uint32_t *ctrls, const uint8_t *cvals,
                        uint32_t nc, int adj, double *scale) {
    GUARD(DISPATCH(prec, {
        T s;
        if (nc == 0)
            s = sv<T>(h).applyGenerator(name, to_wires<T>(wires, nw), adj != 0);
        else
            s = sv<T>(h).applyGenerator(name, to_wires<T>(ctrls, nc), to_bools(cvals, nc),
                                        to_wires<T>(wires, nw), adj != 0);
        *scale = static_cast<double>(s);
    }))
}

int ref_normalize(void *h, int prec) { GUARD(DISPATCH(prec, sv<T>(h).normalize())) }

int ref_collapse(void *h, int prec, uint32_t wire, int branch) {
    GUARD(DISPATCH(prec, sv<T>(h).collapse(wire, branch != 0)))
}

int ref_set_basis_state(void *h, int prec, const uint32_t *bits,
                        const uint32_t *wires, uint32_t nw) {
    GUARD(DISPATCH(prec, sv<T>(h).setBasisState(std::vector<std::size_t>(bits, bits + nw),
                                                 to_wires<T>(wires, nw))))
}

int ref_set_state_vector(void *h, int prec, const void *state,
                         const uint32_t *wires, uint32_t nw) {
    GUARD(DISPATCH(prec, sv<T>(h).setStateVector(static_cast<const std::complex<T> *>(state),
                                                  to_wires<T>(wires, nw))))
}

int ref_probs(void *h, int prec, const uint32_t *wires, uint32_t nw,
              int all, double *out) {
    GUARD(DISPATCH(prec, {
        Measurements<SV<T>> m(sv<T>(h));
        std::vector<T> p = all ? m.probs() : m.probs(to_wires<T>(wires, nw));
        for (std::size_t i = 0; i < p.size(); i++) out[i] = p[i];
    }))
}

int ref_expval_named(void *h, int prec, const char *name,
                     const uint32_t *wires, uint32_t nw, double *out) {
    GUARD(DISPATCH(prec, {
        Measurements<SV<T>> m(sv<T>(h));
        *out = m.expval(std::string(name), to_wires<T>(wires, nw));
    }))
}

int ref_expval_matrix(void *h, int prec, const void *mat,
                      const uint32_t *wires, uint32_t nw, double *out) {
    GUARD(DISPATCH(prec, {
        const auto *mp = static_cast<const std::complex<T> *>(mat);
        std::vector<std::complex<T>> mv(mp, mp + (std::size_t{1} << (2 * nw)));
        Measurements<SV<T>> m(sv<T>(h));
        *out = m.expval(mv, to_wires<T>(wires, nw));
    }))
}

int ref_expval_obs(void *h, int prec, const char *spec, double *out) {
    GUARD(DISPATCH(prec, {
        std::istringstream in(spec);
        auto ob = parse_obs<T>(in);
        Measurements<SV<T>> m(sv<T>(h));
        *out = m.expval(*ob);
    }))
}

// applyControlledSparseMatrix (applySparseMatrix without controls); values complex128
int ref_apply_sparse(void *h, int prec, const int64_t *indptr, uint64_t indptr_len, const int64_t *indices,
                     const void *values_c128, uint64_t nnz, const uint32_t *wires, uint32_t nw,
                     const uint32_t *ctrls, const uint8_t *cvals, uint32_t nc) {
    GUARD(DISPATCH(prec, {
        std::vector<std::size_t> rp(indptr, indptr + indptr_len);
        std::vector<std::size_t> ci(indices, indices + nnz);
        const auto *v = static_cast<const std::complex<double> *>(values_c128);
        std::vector<std::complex<T>> vals(nnz);
        for (uint64_t k = 0; k < nnz; k++) vals[k] = std::complex<T>(static_cast<T>(v[k].real()), static_cast<T>(v[k].imag()));
        if (nc)
            sv<T>(h).applyControlledSparseMatrix(rp.data(), ci.data(), vals.data(), to_wires<T>(ctrls, nc),
                                                 to_bools(cvals, nc), to_wires<T>(wires, nw), false);
        else
            sv<T>(h).applySparseMatrix(rp.data(), ci.data(), vals.data(), to_wires<T>(wires, nw), false);
    }))
}

int ref_var_named(void *h, int prec, const char *name, const uint32_t *wires, uint32_t nw, double *out) {
    GUARD(DISPATCH(prec, {
        Measurements<SV<T>> m(sv<T>(h));
        *out = m.var(std::string(name), to_wires<T>(wires, nw));
    }))
}

int ref_var_obs(void *h, int prec, const char *spec, double *out) {
    GUARD(DISPATCH(prec, {
        std::istringstream in(spec);
        auto ob = parse_obs<T>(in);
        Measurements<SV<T>> m(sv<T>(h));
        *out = m.var(*ob);
    }))
}

// CSR operator over the state index; values given as complex128
int ref_sparse(void *h, int prec, int variance, const int64_t *indptr, uint64_t indptr_len,
               const int64_t *indices, const void *values_c128, uint64_t nnz, double *out) {
    GUARD(DISPATCH(prec, {
        std::vector<std::size_t> rp(indptr, indptr + indptr_len);
        std::vector<std::size_t> ci(indices, indices + nnz);
        const auto *v = static_cast<const std::complex<double> *>(values_c128);
        std::vector<std::complex<T>> vals(nnz);
        for (uint64_t k = 0; k < nnz; k++) vals[k] = std::complex<T>(static_cast<T>(v[k].real()), static_cast<T>(v[k].imag()));
        Measurements<SV<T>> m(sv<T>(h));
        *out = variance ? m.var(rp.data(), rp.size(), ci.data(), vals.data(), vals.size())
                        : m.expval(rp.data(), rp.size(), ci.data(), vals.data(), vals.size());
    }))
}

// generate_samples(wires, shots) (all wires when all != 0) after setSeed(seed)
int ref_generate_samples(void *h, int prec, const uint32_t *wires, uint32_t nw, int all, uint64_t shots,
                         uint64_t seed, uint64_t *out) {
    GUARD(DISPATCH(prec, {
        Measurements<SV<T>> m(sv<T>(h));
        m.setSeed(static_cast<std::size_t>(seed));
        std::vector<std::size_t> s = all ? m.generate_samples(static_cast<std::size_t>(shots))
                                         : m.generate_samples(to_wires<T>(wires, nw), static_cast<std::size_t>(shots));
        for (std::size_t i = 0; i < s.size(); i++) out[i] = s[i];
    }))
}

// Adjoint Jacobian on the state held by `h` (the final state of the ops list;
// apply_operations=0 as LightningAdjointJacobian uses it). Output [obs][param].
int ref_adjoint(void *h, int prec, uint32_t nops, const char *const *names,
                const uint32_t *param_off, const double *params,
                const uint32_t *wire_off, const uint32_t *wires,
                const uint8_t *inverses, const uint32_t *mat_off,
                const double *mats, const uint32_t *ctrl_off,
                const uint32_t *ctrls, const uint8_t *ctrl_vals,
                uint32_t nobs, const char *const *obs_specs,
                const uint64_t *tp, uint32_t ntp, double *jac) {
    GUARD(DISPATCH(prec, {
        auto ops = make_ops<T>(nops, names, param_off, params, wire_off, wires, inverses,
                               mat_off, mats, ctrl_off, ctrls, ctrl_vals);
        std::vector<std::shared_ptr<Observable<SV<T>>>> obs;
        for (uint32_t i = 0; i < nobs; i++) {
            std::istringstream in(obs_specs[i]);
            obs.push_back(parse_obs<T>(in));
        }
        std::vector<std::size_t> tpv(tp, tp + ntp);
        auto &s = sv<T>(h);
        JacobianData<SV<T>> jd(ops.getTotalNumParams(), s.getLength(), s.getData(), obs, ops, tpv);
        std::vector<T> j(static_cast<std::size_t>(nobs) * ntp);
        AdjointJacobian<SV<T>> adj;
        adj.adjointJacobian(std::span<T>{j}, jd, s, false);
        for (std::size_t i = 0; i < j.size(); i++) jac[i] = j[i];
    }))
}

// VectorJacobianProduct on the handle's (final) state; dy: len complex values
// of the handle's precision; out: ntp complex values as (re, im) doubles.
int ref_vjp(void *h, int prec, uint32_t nops, const char *const *names,
            const uint32_t *param_off, const double *params,
            const uint32_t *wire_off, const uint32_t *wires,
            const uint8_t *inverses, const uint32_t *mat_off,
            const double *mats, const uint32_t *ctrl_off,
            const uint32_t *ctrls, const uint8_t *ctrl_vals,
            const void *dy, uint64_t len, const uint64_t *tp, uint32_t ntp, double *out) {
    GUARD(DISPATCH(prec, {
        auto ops = make_ops<T>(nops, names, param_off, params, wire_off, wires, inverses,
                               mat_off, mats, ctrl_off, ctrls, ctrl_vals);
        std::vector<std::size_t> tpv(tp, tp + ntp);
        auto &s = sv<T>(h);
        JacobianData<SV<T>> jd(ops.getTotalNumParams(), s.getLength(), s.getData(), {}, ops, tpv);
        std::vector<std::complex<T>> w(ntp);
        Pennylane::LightningQubit::Algorithms::VectorJacobianProduct<SV<T>> vjp;
        vjp(std::span{w}, jd, std::span<const std::complex<T>>{static_cast<const std::complex<T> *>(dy), len}, false);
        for (std::size_t i = 0; i < w.size(); i++) {
            out[2 * i] = w[i].real();
            out[2 * i + 1] = w[i].imag();
        }
    }))
}

} // extern "C"
