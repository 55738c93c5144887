// Lowering of gates/generators/observables to physical-bit ops, and the pass
// planner that groups them into LDS tile programs.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lq_common.hpp"
#include "lq_gates.hpp"
#include "lq_tileop.hpp"

namespace lqamd {

// An op on physical bits (bit b = wire n-1-b), before tiling.
struct LOp {
    uint32_t kind = TOP_DIAG;
    uint32_t flags = 0;
    int k = 0;
    int bits[kMaxTileOpTargets] = {}; // targets; bits[0] is the local MSB
    uint64_t cmask = 0, cval = 0;     // controls
    uint64_t xmask = 0, yzmask = 0;   // PAULI
    int ny = 0;                       // PAULI
    uint64_t pmask = 0;               // PARITY
    uint64_t pc_dim = 0;              // PCPHASE
    std::vector<cd> coef;
    uint32_t smask = 1;               // gate ops: states it applies to
    uint32_t sa = 0, sb = 0;          // dots: ket / bra
    int32_t slot = -1;                // dots: caller's slot id

    bool is_dot() const { return (flags & TF_DOT) != 0; }
    uint64_t target_mask() const;
    // bits that must be tile bits (non-diagonal action)
    uint64_t nmask() const;
    // bits acted on diagonally (controls, diagonal targets, Z parts)
    uint64_t dmask() const;
    double cost() const;
};

inline int phys_bit(int n, uint32_t wire) { return n - 1 - static_cast<int>(wire); }

// Dense matrix (row-major 2^k x 2^k, k = tbits.size(), tbits[0] local MSB) on
// physical bits with controls. Peels target bits on which the matrix is
// block diagonal with an identity block (unitary mode) or a zero block
// (generator mode) into controls, drops bits it acts on trivially, and emits
// DIAG when the rest is diagonal.
LOp lower_matrix(std::vector<cd> mat, std::vector<int> tbits, uint64_t cmask, uint64_t cval, bool generator_mode);

// Named gate (lq/StateVectorLQubit.hpp:378-423 semantics)
LOp lower_gate(int n, GateId g, const uint32_t *wires, int nw, const uint32_t *ctrls, const uint8_t *cvals, int nc,
               bool inverse, const double *params);
// Generator as an op on the state (non-control subspace zeroed) or as a dot;
// returns the scale through *scale.
LOp lower_generator(int n, GateId g, const uint32_t *wires, int nw, const uint32_t *ctrls, const uint8_t *cvals,
                    int nc, double *scale);

// Peephole: fold `op` into the last queued op when both act on exactly the
// same single target bit with the same controls and nothing in between
// touches that bit. Returns true if merged.
bool try_merge_1q(std::vector<LOp> &queue, const LOp &op);

// Write every uncontrolled 2x2 unitary gate U as D(alpha) R(beta) D(delta):
// a diagonal phase, a rotation with real coefficients, a diagonal phase
// (the Z-Y-Z Euler form, global phase folded into the last factor). Register
// passes then apply the rotation with real arithmetic and apply the phases of
// a stage's register bits together as one table (TOP_RPHASE). Gates whose
// factors do not reproduce U to 1e-13 are kept whole.
void split_1q_rotations(std::vector<LOp> &prog);

// The state before `prog` is the product state with amplitude
// prod_q f[2q + bit q of the index] (f holds one pair per physical bit).
// Fold into f every uncontrolled one-bit gate of state 0 that comes before
// any other op touching its bit (such a gate commutes with everything before
// it), and global phases; remove them from `prog`. Dots read the whole state:
// nothing after the first one is folded. Returns the number of ops folded.
size_t absorb_into_product(std::vector<LOp> &prog, std::vector<cd> &f);

struct PassPlan {
    std::vector<int> tbits;        // tile bits, ascending
    std::vector<TileOp> ops;
    std::vector<cd> coef;
    uint32_t load_mask = 0, store_mask = 0;
    std::vector<int32_t> slot_ids; // caller slot id of each local dot slot
    // register-resident form (tile_reg_kernel): ops grouped into stages
    bool reg = false;
    std::vector<TileStage> stages;
    uint32_t end_transpose = 0;
    // where each coefficient comes from, so a plan can be reused for a
    // program of the same structure with other coefficients: entry i of coef
    // is ops[coef_src[i].op].coef[coef_src[i].idx], or (op == kNoSrc) an
    // entry of a phase table, rebuilt from the entries it multiplies
    static constexpr uint32_t kNoSrc = 0xffffffffu;
    struct CoefSrc {
        uint32_t op, idx;
    };
    std::vector<CoefSrc> coef_src;
    struct PhaseTable {
        uint64_t off;                                 // first entry in coef
        int rb;                                       // 2^rb entries
        std::vector<std::pair<uint64_t, int>> phases; // (coef index of d0, register index)
    };
    std::vector<PhaseTable> tables;
    // rotations kept as [[1, -t], [t, 1]] (coef offsets) and the phase table
    // that carries the product of their cosines (see PlanConfig::scale_rotations)
    std::vector<uint64_t> scaled;
    int64_t scale_table = -1;
    double scale_growth_limit = 1e300;
    // eligible rotations left in the full form because of the growth bound
    std::vector<uint64_t> scale_dropped;
};


struct PlanConfig {
    int n = 0;
    int ns = 1;          // states resident per pass
    int kt_max = 12;     // max tile bits
    int low_bits = 4;    // contiguous low bits always in the tile
    double budget = 24;  // max op cost per pass
    // > 0: the first pass's budget (a pass that forms a deferred state
    // writes the state without reading it, so it has half the memory time to
    // hide its arithmetic behind)
    double first_budget = -1;
    int reg_rb = 0;      // > 0: emit register stages for tiles of 8 + reg_rb bits
    int max_dots = 1 << 30; // dot slots per pass
    // In one-state register passes without dots, store each real rotation
    // [[c, -s], [s, c]] as [[1, -t], [t, 1]] with t = s/c and multiply the
    // product of the cosines into one phase table of the pass: two FMAs per
    // amplitude instead of a multiply and an FMA per component. Rotations with
    // |c| below 1e-100 keep the full form. The unscaled intermediate
    // amplitudes grow by 1/c per rotation: a pass whose product of 1/c would
    // exceed max_scale_growth keeps every rotation in the full form.
    bool scale_rotations = false;
    double max_scale_growth = 1e300;
    // choose each pass's tile bits by lookahead (the bits that let the most
    // work run) instead of taking the first ready op's bits
    bool lookahead = false;
};

std::vector<PassPlan> plan_passes(const std::vector<LOp> &ops, const PlanConfig &cfg);

// Structural key of a program (everything planning depends on except the
// coefficient values, plus the value tests planning makes: one-bit and
// global diagonals equal to +-1 become sign flips).
std::string plan_structure_key(const std::vector<LOp> &ops, const PlanConfig &cfg);
// Refill the coefficients of `passes`, planned for a program of the same
// structure, from `ops`.
// Returns false when the new coefficients do not fit the plan's value-dependent
// choices (a scaled rotation whose cosine is now 0): plan afresh then.
bool refresh_plan_coefs(std::vector<PassPlan> &passes, const std::vector<LOp> &ops);

TileGeom make_geom(int n, const std::vector<int> &tbits);

} // namespace lqamd
