// Tile-program format shared by the host planner and the HIP tile kernel.
//
// A pass loads, for every state it touches, a tile of 2^kt amplitudes whose
// indices vary over kt chosen "tile bits" (always the lowest `c` bits, for
// coalesced 2^c-amplitude segments, plus any higher bits the pass needs);
// the remaining n-kt "outside" bits are fixed per tile (one tile per
// workgroup). Ops then run in LDS in program order. Tile-local index l maps to
// global index base(tile) | spread(l).
//
// Every op's targets that act non-diagonally must be tile bits (lpos >= 0).
// Diagonal action (controls, diagonal targets, Z parts of Pauli strings) may
// sit on any bit: its outside part is resolved once per tile from base.
#pragma once

#include <cstdint>

namespace lqamd {

constexpr int kMaxTileOpTargets = 12; // dense k <= 10, diag table k <= 12
constexpr int kMaxStates = 4;         // states resident per pass
constexpr int kBlock = 256;           // threads per tile workgroup
constexpr int kMaxOutsideBits = 64;

enum TileOpKind : uint32_t {
    TOP_DENSE1 = 1,  // 2x2 on one tile bit
    TOP_DENSE2 = 2,  // 4x4 on two tile bits (lpos[0] is the local MSB)
    TOP_DENSEK = 3,  // 2^k x 2^k, 3 <= k <= 10, on tile bits
    TOP_DIAG = 4,    // table of 2^k phases indexed by the target bits (any bits)
    TOP_PARITY = 5,  // coef[parity(index & mask)] (MultiRZ, Pauli-Z products)
    TOP_PCPHASE = 6, // local index < dim ? coef[0] : coef[1]  (PCPhase)
    TOP_PAULI = 7,   // Pauli string: (-i)^ny (-1)^popcnt(i & yzmask) a[i ^ xmask]
    TOP_SIGN = 8,    // negate where (index & gcmask) == gcval  (Z, CZ, multi-controlled Z)
    TOP_CZSET = 9,   // negate where sum_i popcnt(index & (index >> sd[i]) & sm[i]) is odd
    TOP_RPHASE = 10, // register stages: amplitude at register index e times coef[e] (2^RB table)
    TOP_TPHASE = 11, // register stages: every amplitude times prod_j coef[2j + bit gpos[j] of the
                     // index], over k bits fixed per thread (one-bit phases off the register bits)
};

// Register-resident tile programs (tile_reg_kernel): a pass is a list of
// stages; in a REG stage every thread holds 2^RB amplitudes per state whose
// tile-local indices differ in the stage's register bits; non-diagonal
// targets of the stage's ops are register bits. LDS stages run ops on the
// LDS copy of the tile (ops the register path does not cover).
enum TileStageKind : uint32_t { STAGE_REG = 1, STAGE_LDS = 2 };
enum TileStageEntry : uint32_t { ENTRY_NONE = 0, ENTRY_TRANSPOSE = 1, ENTRY_FROM_LDS = 2 };
constexpr int kMaxRegBits = 4;    // 16 amplitudes per state per thread
constexpr int kMaxPassDots = 32;  // dot slots per register-kernel pass
// dot slots per specialised pass of a program on three or four states (the
// adjoint sweep of several observables: each trainable op dots G lambda with
// every observable's state, so such passes fill their slots 2-3x faster)
constexpr int kMaxJitPassDots = 96;

struct TileStage {
    uint32_t kind;
    uint32_t op_begin, op_end;
    uint32_t entry;                 // how the stage's mapping is reached
    int32_t rbits[kMaxRegBits];     // tile-local bit of register index bit i
    int32_t hbits[8];               // tile-local bit of thread index bit j
};

enum TileOpFlags : uint32_t {
    TF_DOT = 1u,         // read-only: accumulate sum conj(state[sb]) * (G state[sa]) into slot
    TF_ZERO_UNCTRL = 2u, // zero the amplitudes outside the control subspace (generators)
    TF_REAL = 4u,        // dense op with real coefficients (imaginary parts zero)
    TF_SCALED = 8u,      // real rotation stored as [[1, -t], [t, 1]]; its cosine is applied once per pass
};

struct alignas(16) TileOp {
    uint32_t kind;
    uint32_t flags;
    uint32_t k;       // number of targets
    uint32_t smask;   // gate ops: bitmask of states the op is applied to
    uint32_t sa, sb;  // dot ops: ket (G applied) and bra state indices
    uint32_t slot;    // dot ops: output slot
    uint32_t lcmask;  // control bits inside the tile (tile-local positions)
    uint32_t lcval;
    uint32_t lxmask;  // PAULI: flipped bits (tile-local); PARITY: tile-local mask
    uint32_t lzmask;  // PAULI: Y|Z bits inside the tile (tile-local)
    uint32_t ny;      // PAULI: number of Y factors
    uint64_t ocmask;  // control bits outside the tile (global positions)
    uint64_t ocval;
    uint64_t omask;   // PARITY/PAULI: outside bits contributing to the sign/parity
    uint64_t coef_off; // offset (complex entries) into the pass coefficient buffer
    uint64_t pc_dim;   // PCPHASE: threshold on the local index
    uint64_t gcmask;   // controls on global bits (register stages)
    uint64_t gcval;
    uint64_t gpmask;   // PARITY: global mask (register stages)
    uint64_t sm[4];    // CZSET: bit a of sm[i] pairs bits a and a + sd[i]
    uint32_t sd[4];
    uint32_t nz;       // CZSET: number of (sd, sm) terms
    uint32_t pad2;
    int32_t lpos[kMaxTileOpTargets]; // tile-local position of target j (-1: outside)
    int32_t gpos[kMaxTileOpTargets]; // global bit of target j (target 0 = local MSB)
};

struct TileGeom {
    int32_t n;                // qubits
    int32_t kt;               // tile bits
    int32_t nout;             // outside bits = n - kt
    int32_t c;                // leading contiguous low bits (tb[j] = j for j < c)
    int32_t tb[16];           // tile bits, ascending
    int32_t ob[kMaxOutsideBits]; // outside bits, ascending
};

} // namespace lqamd
