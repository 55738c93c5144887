// Named gates: the name/arity tables of the reference and host-side matrices
// and generators in double precision. Device kernels never evaluate
// trigonometry; every coefficient comes from here.
#pragma once

#include <string>
#include <vector>

#include "lq_common.hpp"

namespace lqamd {

enum class GateId : int {
    Identity, PauliX, PauliY, PauliZ, Hadamard, S, SX, T, PhaseShift, RX, RY, RZ, Rot,
    CNOT, CY, CZ, SWAP, IsingXX, IsingXY, IsingYY, IsingZZ, ControlledPhaseShift,
    CRX, CRY, CRZ, CRot, SingleExcitation, SingleExcitationMinus, SingleExcitationPlus,
    Toffoli, CSWAP, DoubleExcitation, DoubleExcitationMinus, DoubleExcitationPlus,
    MultiRZ, GlobalPhase, PSWAP, PCPhase, Count
};

struct GateInfo {
    const char *name;
    int nwires;       // -1: any number (MultiRZ, PCPhase)
    int nparams;
    bool controllable; // accepted by the controlled form
    bool has_generator;
};

const GateInfo &gate_info(GateId g);
// Name lookup (gates/Constant.hpp:60-99); false if unknown.
bool gate_from_name(const std::string &name, GateId *out);
int gate_count();

// Row-major 2^k x 2^k matrix, wires[0] the local MSB; inverse gives U^dagger.
std::vector<cd> gate_matrix(GateId g, const double *params, bool inverse, int nwires);

// Generator G and scale s with U(theta) = exp(i s theta G)
// (lq/gates/cpu_kernels/PauliGenerator.hpp:31-60, GateImplementationsLM.hpp:2181-2952).
struct Generator {
    std::vector<cd> g; // row-major, 2^k x 2^k
    double scale;
};
Generator gate_generator(GateId g, int nwires);

} // namespace lqamd
