// Gate tables, matrices and generators (host, double precision).
//
// Matrices follow gates/Gates.hpp:29-1330 (row-major, wires[0] = local MSB);
// `inverse` is the conjugate transpose, which equals the reference's per-gate
// rule (negated angles; Rot(-omega,-theta,-phi); conjugated S, T, SX).
// Generators follow the LM kernels' action and scale
// (lq/gates/cpu_kernels/GateImplementationsLM.hpp:2181-2952).
#include "lq_gates.hpp"

#include <cmath>
#include <cstring>
#include <unordered_map>

namespace lqamd {

namespace {
// name, wires, params, controllable, generator  (gates/Constant.hpp:60-431)
const GateInfo kGates[] = {
    {"Identity", 1, 0, false, false},
    {"PauliX", 1, 0, true, false},
    {"PauliY", 1, 0, true, false},
    {"PauliZ", 1, 0, true, false},
    {"Hadamard", 1, 0, true, false},
    {"S", 1, 0, true, false},
    {"SX", 1, 0, true, false},
    {"T", 1, 0, true, false},
    {"PhaseShift", 1, 1, true, true},
    {"RX", 1, 1, true, true},
    {"RY", 1, 1, true, true},
    {"RZ", 1, 1, true, true},
    {"Rot", 1, 3, true, false},
    {"CNOT", 2, 0, false, false},
    {"CY", 2, 0, false, false},
    {"CZ", 2, 0, false, false},
    {"SWAP", 2, 0, true, false},
    {"IsingXX", 2, 1, true, true},
    {"IsingXY", 2, 1, true, true},
    {"IsingYY", 2, 1, true, true},
    {"IsingZZ", 2, 1, true, true},
    {"ControlledPhaseShift", 2, 1, false, true},
    {"CRX", 2, 1, false, true},
    {"CRY", 2, 1, false, true},
    {"CRZ", 2, 1, false, true},
    {"CRot", 2, 3, false, false},
    {"SingleExcitation", 2, 1, true, true},
    {"SingleExcitationMinus", 2, 1, true, true},
    {"SingleExcitationPlus", 2, 1, true, true},
    {"Toffoli", 3, 0, false, false},
    {"CSWAP", 3, 0, false, false},
    {"DoubleExcitation", 4, 1, true, true},
    {"DoubleExcitationMinus", 4, 1, true, true},
    {"DoubleExcitationPlus", 4, 1, true, true},
    {"MultiRZ", -1, 1, true, true},
    {"GlobalPhase", 1, 1, true, true},
    {"PSWAP", 2, 1, true, true},
    {"PCPhase", -1, 2, true, false},
};
static_assert(sizeof(kGates) / sizeof(kGates[0]) == static_cast<size_t>(GateId::Count), "gate table size");

const cd I1{0.0, 1.0};

std::vector<cd> eye(int d) {
    std::vector<cd> m(static_cast<size_t>(d) * d, 0.0);
    for (int i = 0; i < d; ++i) m[static_cast<size_t>(i) * d + i] = 1.0;
    return m;
}

// |0><0| (x) I + |1><1| (x) U: the control is the local MSB
std::vector<cd> ctrl_block(const std::vector<cd> &u) {
    const int d = static_cast<int>(std::lround(std::sqrt(static_cast<double>(u.size()))));
    auto m = eye(2 * d);
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) m[static_cast<size_t>(d + r) * (2 * d) + d + c] = u[static_cast<size_t>(r) * d + c];
    return m;
}

std::vector<cd> rot(double phi, double theta, double omega) {
    const double c = std::cos(theta / 2), s = std::sin(theta / 2);
    const double p = phi + omega, m = phi - omega;
    return {cd(std::cos(p / 2) * c, -std::sin(p / 2) * c), cd(-std::cos(m / 2) * s, -std::sin(m / 2) * s),
            cd(std::cos(m / 2) * s, -std::sin(m / 2) * s), cd(std::cos(p / 2) * c, std::sin(p / 2) * c)};
}

std::vector<cd> kron(const std::vector<cd> &a, int da, const std::vector<cd> &b, int db) {
    const int d = da * db;
    std::vector<cd> m(static_cast<size_t>(d) * d);
    for (int r1 = 0; r1 < da; ++r1)
        for (int c1 = 0; c1 < da; ++c1)
            for (int r2 = 0; r2 < db; ++r2)
                for (int c2 = 0; c2 < db; ++c2)
                    m[static_cast<size_t>(r1 * db + r2) * d + c1 * db + c2] =
                        a[static_cast<size_t>(r1) * da + c1] * b[static_cast<size_t>(r2) * db + c2];
    return m;
}

const std::vector<cd> kX{0, 1, 1, 0};
const std::vector<cd> kY{0, cd(0, -1), cd(0, 1), 0};
const std::vector<cd> kZ{1, 0, 0, -1};
const std::vector<cd> kP1{0, 0, 0, 1};

std::vector<cd> dagger(const std::vector<cd> &m) {
    const int d = static_cast<int>(std::lround(std::sqrt(static_cast<double>(m.size()))));
    std::vector<cd> r(m.size());
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) r[static_cast<size_t>(j) * d + i] = std::conj(m[static_cast<size_t>(i) * d + j]);
    return r;
}
} // namespace

const GateInfo &gate_info(GateId g) { return kGates[static_cast<int>(g)]; }
int gate_count() { return static_cast<int>(GateId::Count); }

bool gate_from_name(const std::string &name, GateId *out) {
    static const std::unordered_map<std::string, GateId> table = [] {
        std::unordered_map<std::string, GateId> t;
        for (int i = 0; i < static_cast<int>(GateId::Count); ++i) t.emplace(kGates[i].name, static_cast<GateId>(i));
        return t;
    }();
    auto it = table.find(name);
    if (it == table.end()) return false;
    *out = it->second;
    return true;
}

std::vector<cd> gate_matrix(GateId g, const double *p, bool inverse, int nwires) {
    std::vector<cd> u;
    switch (g) {
    case GateId::Identity: u = eye(2); break;
    case GateId::PauliX: u = kX; break;
    case GateId::PauliY: u = kY; break;
    case GateId::PauliZ: u = kZ; break;
    case GateId::Hadamard: {
        const double r = 1.0 / std::sqrt(2.0);
        u = {r, r, r, -r};
        break;
    }
    case GateId::S: u = {1, 0, 0, I1}; break;
    case GateId::T: u = {1, 0, 0, std::exp(I1 * (M_PI / 4))}; break;
    case GateId::SX: u = {cd(0.5, 0.5), cd(0.5, -0.5), cd(0.5, -0.5), cd(0.5, 0.5)}; break;
    case GateId::PhaseShift: u = {1, 0, 0, std::exp(I1 * p[0])}; break;
    case GateId::RX: {
        const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
        u = {c, cd(0, -s), cd(0, -s), c};
        break;
    }
    case GateId::RY: {
        const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
        u = {c, -s, s, c};
        break;
    }
    case GateId::RZ: u = {std::exp(-I1 * (p[0] / 2)), 0, 0, std::exp(I1 * (p[0] / 2))}; break;
    case GateId::Rot: u = rot(p[0], p[1], p[2]); break;
    case GateId::GlobalPhase: {
        const cd e = std::exp(-I1 * p[0]);
        u = {e, 0, 0, e};
        break;
    }
    case GateId::CNOT: u = ctrl_block(kX); break;
    case GateId::CY: u = ctrl_block(kY); break;
    case GateId::CZ: u = ctrl_block(kZ); break;
    case GateId::SWAP: u = {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1}; break;
    case GateId::ControlledPhaseShift: {
        u = eye(4);
        u[15] = std::exp(I1 * p[0]);
        break;
    }
    case GateId::CRX: u = ctrl_block(gate_matrix(GateId::RX, p, false, 1)); break;
    case GateId::CRY: u = ctrl_block(gate_matrix(GateId::RY, p, false, 1)); break;
    case GateId::CRZ: u = ctrl_block(gate_matrix(GateId::RZ, p, false, 1)); break;
    case GateId::CRot: u = ctrl_block(rot(p[0], p[1], p[2])); break;
    case GateId::IsingXX:
    case GateId::IsingYY: {
        const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
        const auto pp = g == GateId::IsingXX ? kron(kX, 2, kX, 2) : kron(kY, 2, kY, 2);
        u = eye(4);
        for (size_t i = 0; i < 16; ++i) u[i] = u[i] * c - I1 * s * pp[i];
        break;
    }
    case GateId::IsingZZ: {
        const cd e = std::exp(-I1 * (p[0] / 2));
        u = {e, 0, 0, 0, 0, std::conj(e), 0, 0, 0, 0, std::conj(e), 0, 0, 0, 0, e};
        break;
    }
    case GateId::IsingXY: {
        const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
        u = eye(4);
        u[5] = u[10] = c;
        u[6] = u[9] = cd(0, s);
        break;
    }
    case GateId::SingleExcitation:
    case GateId::SingleExcitationMinus:
    case GateId::SingleExcitationPlus: {
        const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
        const cd e = g == GateId::SingleExcitation        ? cd(1.0)
                     : g == GateId::SingleExcitationMinus ? std::exp(-I1 * (p[0] / 2))
                                                          : std::exp(I1 * (p[0] / 2));
        u = std::vector<cd>(16, 0.0);
        u[0] = u[15] = e;
        u[5] = u[10] = c;
        u[6] = -s;
        u[9] = s;
        break;
    }
    case GateId::PSWAP: {
        const cd e = std::exp(I1 * p[0]);
        u = std::vector<cd>(16, 0.0);
        u[0] = u[15] = 1;
        u[6] = u[9] = e;
        break;
    }
    case GateId::Toffoli: u = ctrl_block(ctrl_block(kX)); break;
    case GateId::CSWAP: u = ctrl_block(gate_matrix(GateId::SWAP, p, false, 2)); break;
    case GateId::DoubleExcitation:
    case GateId::DoubleExcitationMinus:
    case GateId::DoubleExcitationPlus: {
        const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
        const cd e = g == GateId::DoubleExcitation        ? cd(1.0)
                     : g == GateId::DoubleExcitationMinus ? std::exp(-I1 * (p[0] / 2))
                                                          : std::exp(I1 * (p[0] / 2));
        u = eye(16);
        for (auto &x : u) x *= e;
        u[3 * 16 + 3] = u[12 * 16 + 12] = c;
        u[3 * 16 + 12] = -s;
        u[12 * 16 + 3] = s;
        break;
    }
    case GateId::MultiRZ: {
        const int d = 1 << nwires;
        u = std::vector<cd>(static_cast<size_t>(d) * d, 0.0);
        for (int i = 0; i < d; ++i)
            u[static_cast<size_t>(i) * d + i] =
                (__builtin_popcount(i) & 1) ? std::exp(I1 * (p[0] / 2)) : std::exp(-I1 * (p[0] / 2));
        break;
    }
    case GateId::PCPhase: {
        const int d = 1 << nwires;
        const long dim = std::lround(p[1]);
        u = std::vector<cd>(static_cast<size_t>(d) * d, 0.0);
        for (int i = 0; i < d; ++i) u[static_cast<size_t>(i) * d + i] = i < dim ? std::exp(I1 * p[0]) : std::exp(-I1 * p[0]);
        break;
    }
    case GateId::Count: LQ_ABORT("invalid gate");
    }
    return inverse ? dagger(u) : u;
}

Generator gate_generator(GateId g, int nwires) {
    switch (g) {
    case GateId::RX: return {kX, -0.5};
    case GateId::RY: return {kY, -0.5};
    case GateId::RZ: return {kZ, -0.5};
    case GateId::PhaseShift: return {kP1, 1.0};
    case GateId::GlobalPhase: return {eye(2), -1.0};
    case GateId::CRX: return {kron(kP1, 2, kX, 2), -0.5};
    case GateId::CRY: return {kron(kP1, 2, kY, 2), -0.5};
    case GateId::CRZ: return {kron(kP1, 2, kZ, 2), -0.5};
    case GateId::ControlledPhaseShift: return {kron(kP1, 2, kP1, 2), 1.0};
    case GateId::IsingXX: return {kron(kX, 2, kX, 2), -0.5};
    case GateId::IsingYY: return {kron(kY, 2, kY, 2), -0.5};
    case GateId::IsingZZ: return {kron(kZ, 2, kZ, 2), -0.5};
    case GateId::IsingXY: {
        std::vector<cd> m(16, 0.0);
        m[6] = m[9] = 1.0;
        return {m, 0.5};
    }
    case GateId::SingleExcitation:
    case GateId::SingleExcitationMinus:
    case GateId::SingleExcitationPlus: {
        const double d = g == GateId::SingleExcitation ? 0.0 : g == GateId::SingleExcitationMinus ? 1.0 : -1.0;
        std::vector<cd> m(16, 0.0);
        m[0] = m[15] = d;
        m[6] = cd(0, -1);
        m[9] = cd(0, 1);
        return {m, -0.5};
    }
    case GateId::PSWAP: {
        // the reference's PSWAP generator: swap block, scale +1
        // (GateImplementationsLM.hpp:2666-2689)
        std::vector<cd> m(16, 0.0);
        m[6] = m[9] = 1.0;
        return {m, 1.0};
    }
    case GateId::DoubleExcitation: {
        std::vector<cd> m(256, 0.0);
        m[3 * 16 + 12] = cd(0, -1);
        m[12 * 16 + 3] = cd(0, 1);
        return {m, -0.5};
    }
    case GateId::DoubleExcitationMinus:
    case GateId::DoubleExcitationPlus: {
        const bool minus = g == GateId::DoubleExcitationMinus;
        auto m = eye(16);
        m[3 * 16 + 3] = m[12 * 16 + 12] = 0.0;
        m[3 * 16 + 12] = minus ? cd(0, -1) : cd(0, 1);
        m[12 * 16 + 3] = minus ? cd(0, 1) : cd(0, -1);
        return {m, minus ? -0.5 : 0.5};
    }
    case GateId::MultiRZ: {
        const int d = 1 << nwires;
        std::vector<cd> m(static_cast<size_t>(d) * d, 0.0);
        for (int i = 0; i < d; ++i) m[static_cast<size_t>(i) * d + i] = (__builtin_popcount(i) & 1) ? -1.0 : 1.0;
        return {m, -0.5};
    }
    default: break;
    }
    LQ_ABORT(("The gate " + std::string(gate_info(g).name) + " has no generator").c_str());
}

} // namespace lqamd
