// Lowering and pass planning (host).
//
// Gate semantics follow the reference: controls select the subspace
// ctrl == value (GateImplementationsLM.hpp:407-440); generators zero the
// non-control subspace (:2181-2228); MultiRZ is a parity phase (:1988-2010);
// PCPhase is a threshold phase on the local index (:2113-2162).
//
// Planning: ops form a DAG in which two ops commute when every bit they share
// is acted on diagonally by both (controls, diagonal targets), or when both
// are read-only dots. A pass greedily takes ready ops whose non-diagonal bits
// fit in its tile (at most kt_max bits, the low bits always included), up to
// a compute budget; diagonal ops fit any pass.
#include "lq_plan.hpp"

#include <algorithm>
#include <cmath>
#include <queue>
#include <set>

namespace lqamd {

namespace {
inline int popc(uint64_t x) { return __builtin_popcountll(x); }

bool all_zero(const std::vector<cd> &m) {
    for (const auto &x : m)
        if (x != cd(0.0)) return false;
    return true;
}
bool is_identity(const std::vector<cd> &m, int d) {
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c)
            if (m[static_cast<size_t>(r) * d + c] != (r == c ? cd(1.0) : cd(0.0))) return false;
    return true;
}
bool is_diagonal(const std::vector<cd> &m, int d) {
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c)
            if (r != c && m[static_cast<size_t>(r) * d + c] != cd(0.0)) return false;
    return true;
}
} // namespace

uint64_t LOp::target_mask() const {
    uint64_t m = 0;
    for (int j = 0; j < k; ++j) m |= 1ull << bits[j];
    return m;
}

uint64_t LOp::nmask() const {
    switch (kind) {
    case TOP_DENSE1:
    case TOP_DENSE2:
    case TOP_DENSEK: return target_mask();
    case TOP_PAULI: return xmask;
    default: return 0;
    }
}

uint64_t LOp::dmask() const {
    uint64_t m = cmask;
    switch (kind) {
    case TOP_DIAG:
    case TOP_PCPHASE: m |= target_mask(); break;
    case TOP_PARITY: m |= pmask; break;
    case TOP_PAULI: m |= yzmask & ~xmask; break;
    default: break;
    }
    return m;
}

double LOp::cost() const {
    double c = 0.5;
    switch (kind) {
    case TOP_DENSE1: c = (flags & TF_REAL) ? 0.5 : 1.0; break; // real rotations take half the arithmetic
    case TOP_DENSE2: c = 2.5; break;
    case TOP_DENSEK: c = static_cast<double>(1 << k) / 2.0; break;
    case TOP_DIAG: c = (k == 1 && cmask == 0) ? 0.15 : 0.5; break; // one-bit phases share a table per stage
    case TOP_PARITY: c = 0.4; break;
    case TOP_PCPHASE: c = 0.5; break;
    case TOP_PAULI: c = 0.6; break;
    default: break;
    }
    if (is_dot()) return c + 0.5;
    return c * std::max(1, popc(smask));
}

LOp lower_matrix(std::vector<cd> mat, std::vector<int> tbits, uint64_t cmask, uint64_t cval, bool gen) {
    int k = static_cast<int>(tbits.size());
    LQ_ABORT_IF(mat.size() != (static_cast<size_t>(1) << (2 * k)), "The size of matrix does not match with the given number of wires");
    bool changed = true;
    while (changed && k > 0) {
        changed = false;
        for (int j = 0; j < k && !changed; ++j) {
            const int d = 1 << k, h = d / 2, lb = k - 1 - j;
            bool offzero = true;
            for (int r = 0; r < d && offzero; ++r)
                for (int c = 0; c < d; ++c)
                    if ((((r >> lb) ^ (c >> lb)) & 1) && mat[static_cast<size_t>(r) * d + c] != cd(0.0)) {
                        offzero = false;
                        break;
                    }
            if (!offzero) continue;
            std::vector<cd> A(static_cast<size_t>(h) * h), B(static_cast<size_t>(h) * h);
            auto ins = [lb](int x, int v) { return ((x >> lb) << (lb + 1)) | (v << lb) | (x & ((1 << lb) - 1)); };
            for (int r = 0; r < h; ++r)
                for (int c = 0; c < h; ++c) {
                    A[static_cast<size_t>(r) * h + c] = mat[static_cast<size_t>(ins(r, 0)) * d + ins(c, 0)];
                    B[static_cast<size_t>(r) * h + c] = mat[static_cast<size_t>(ins(r, 1)) * d + ins(c, 1)];
                }
            int ctrl = -1;
            std::vector<cd> *keep = nullptr;
            if (A == B) {
                keep = &A;
            } else if (!gen && is_identity(A, h)) {
                keep = &B;
                ctrl = 1;
            } else if (!gen && is_identity(B, h)) {
                keep = &A;
                ctrl = 0;
            } else if (gen && all_zero(A)) {
                keep = &B;
                ctrl = 1;
            } else if (gen && all_zero(B)) {
                keep = &A;
                ctrl = 0;
            }
            if (!keep) continue;
            if (ctrl >= 0) {
                cmask |= 1ull << tbits[j];
                if (ctrl) cval |= 1ull << tbits[j];
            }
            mat = *keep;
            tbits.erase(tbits.begin() + j);
            --k;
            changed = true;
        }
    }
    LOp op;
    op.cmask = cmask;
    op.cval = cval;
    op.k = k;
    for (int j = 0; j < k; ++j) op.bits[j] = tbits[j];
    const int d = 1 << k;
    if (k == 0 || is_diagonal(mat, d)) {
        LQ_ABORT_IF(k > kMaxTileOpTargets, "diagonal operator on too many wires");
        op.kind = TOP_DIAG;
        op.coef.resize(d);
        for (int i = 0; i < d; ++i) op.coef[i] = mat[static_cast<size_t>(i) * d + i];
        return op;
    }
    LQ_ABORT_IF(k > 10, "dense operators are supported on at most 10 wires");
    op.kind = k == 1 ? TOP_DENSE1 : k == 2 ? TOP_DENSE2 : TOP_DENSEK;
    op.coef = std::move(mat);
    return op;
}

namespace {
void controls_to_mask(int n, const uint32_t *ctrls, const uint8_t *cvals, int nc, uint64_t *cm, uint64_t *cv) {
    *cm = 0;
    *cv = 0;
    for (int i = 0; i < nc; ++i) {
        const int b = phys_bit(n, ctrls[i]);
        *cm |= 1ull << b;
        if (cvals == nullptr || cvals[i]) *cv |= 1ull << b;
    }
}
} // namespace

LOp lower_gate(int n, GateId g, const uint32_t *wires, int nw, const uint32_t *ctrls, const uint8_t *cvals, int nc,
               bool inverse, const double *params) {
    uint64_t cm, cv;
    controls_to_mask(n, ctrls, cvals, nc, &cm, &cv);
    std::vector<int> tb(nw);
    for (int j = 0; j < nw; ++j) tb[j] = phys_bit(n, wires[j]);
    const cd I1(0.0, 1.0);
    if (g == GateId::MultiRZ) {
        LOp op;
        op.kind = TOP_PARITY;
        op.cmask = cm;
        op.cval = cv;
        for (int b : tb) op.pmask |= 1ull << b;
        const cd first = std::exp(-I1 * (params[0] / 2)), second = std::exp(I1 * (params[0] / 2));
        op.coef = {inverse ? std::conj(first) : first, inverse ? std::conj(second) : second};
        return op;
    }
    if (g == GateId::PCPhase) {
        LQ_ABORT_IF(nw > kMaxTileOpTargets, "PCPhase is supported on at most 12 wires");
        const double dim = std::round(params[1]);
        LQ_ABORT_IF(dim < 0 || dim > std::ldexp(1.0, n),
                    "The dimension of the PCPhase gate must be a positive integer and less than or equal to statevector size.");
        LOp op;
        op.kind = TOP_PCPHASE;
        op.cmask = cm;
        op.cval = cv;
        op.k = nw;
        for (int j = 0; j < nw; ++j) op.bits[j] = tb[j];
        op.pc_dim = static_cast<uint64_t>(dim);
        const double ph = inverse ? -params[0] : params[0];
        op.coef = {std::exp(I1 * ph), std::exp(-I1 * ph)};
        return op;
    }
    if (g == GateId::GlobalPhase) {
        LOp op;
        op.kind = TOP_DIAG;
        op.k = 0;
        op.cmask = cm;
        op.cval = cv;
        op.coef = {std::exp(cd(0.0, inverse ? params[0] : -params[0]))};
        return op;
    }
    return lower_matrix(gate_matrix(g, params, inverse, nw), tb, cm, cv, false);
}

LOp lower_generator(int n, GateId g, const uint32_t *wires, int nw, const uint32_t *ctrls, const uint8_t *cvals,
                    int nc, double *scale) {
    uint64_t cm, cv;
    controls_to_mask(n, ctrls, cvals, nc, &cm, &cv);
    std::vector<int> tb(nw);
    for (int j = 0; j < nw; ++j) tb[j] = phys_bit(n, wires[j]);
    LOp op;
    if (g == GateId::MultiRZ) {
        op.kind = TOP_PARITY;
        op.cmask = cm;
        op.cval = cv;
        for (int b : tb) op.pmask |= 1ull << b;
        op.coef = {1.0, -1.0};
        *scale = -0.5;
    } else if (g == GateId::GlobalPhase) {
        op.kind = TOP_DIAG;
        op.k = 0;
        op.cmask = cm;
        op.cval = cv;
        op.coef = {1.0};
        *scale = -1.0;
    } else {
        Generator gen = gate_generator(g, nw);
        op = lower_matrix(gen.g, tb, cm, cv, true);
        *scale = gen.scale;
    }
    op.flags |= TF_ZERO_UNCTRL;
    return op;
}

namespace {
uint64_t touch_mask(const LOp &o) { return o.target_mask() | o.cmask | o.pmask | o.xmask | o.yzmask; }

bool single_bit_1q(const LOp &o) {
    return !o.is_dot() && !(o.flags & TF_ZERO_UNCTRL) && o.k == 1 && (o.kind == TOP_DENSE1 || o.kind == TOP_DIAG);
}

std::vector<cd> as_2x2(const LOp &o) {
    if (o.kind == TOP_DENSE1) return o.coef;
    return {o.coef[0], 0.0, 0.0, o.coef[1]};
}
} // namespace

size_t absorb_into_product(std::vector<LOp> &prog, std::vector<cd> &f) {
    const int nq = static_cast<int>(f.size() / 2);
    LQ_ABORT_IF(nq < 1 || nq > 64, "product state of unsupported size");
    const uint64_t all = nq == 64 ? ~0ull : ((1ull << nq) - 1ull);
    uint64_t blocked = 0;
    size_t folded = 0, i = 0;
    std::vector<LOp> rest;
    rest.reserve(prog.size());
    for (; i < prog.size() && blocked != all; ++i) {
        LOp &o = prog[i];
        const bool own = !o.is_dot() && !(o.flags & TF_ZERO_UNCTRL) && o.cmask == 0 && o.smask == 1u;
        if (own && o.kind == TOP_DIAG && o.k == 0 && o.coef.size() == 1) {
            f[0] *= o.coef[0];
            f[1] *= o.coef[0];
            ++folded;
            continue;
        }
        if (own && single_bit_1q(o) && o.bits[0] < nq && !((blocked >> o.bits[0]) & 1ull)) {
            const size_t q = static_cast<size_t>(o.bits[0]);
            const std::vector<cd> m = as_2x2(o);
            const cd a = f[2 * q], b = f[2 * q + 1];
            f[2 * q] = m[0] * a + m[1] * b;
            f[2 * q + 1] = m[2] * a + m[3] * b;
            ++folded;
            continue;
        }
        blocked |= o.is_dot() ? all : touch_mask(o);
        rest.push_back(std::move(o));
    }
    for (; i < prog.size(); ++i) rest.push_back(std::move(prog[i]));
    prog.swap(rest);
    return folded;
}

void split_1q_rotations(std::vector<LOp> &prog) {
    std::vector<LOp> out;
    out.reserve(prog.size() * 3);
    for (auto &o : prog) {
        const bool candidate = (o.kind == TOP_DENSE1 || o.kind == TOP_DENSEK) && o.k == 1 && o.cmask == 0 &&
                               !o.is_dot() && !(o.flags & TF_ZERO_UNCTRL) && o.coef.size() == 4;
        if (!candidate) {
            out.push_back(std::move(o));
            continue;
        }
        const cd u00 = o.coef[0], u01 = o.coef[1], u10 = o.coef[2], u11 = o.coef[3];
        const double gamma = std::arg(u00 * u11 - u01 * u10) / 2.0;
        const cd eg = std::polar(1.0, -gamma);
        const cd a = eg * u00, b = eg * u10;
        const double c = std::abs(a), sn = std::abs(b);
        const double pa = c > 1e-300 ? std::arg(a) : 0.0, pb = sn > 1e-300 ? std::arg(b) : 0.0;
        const double alpha = pb - pa, delta = -pa - pb;
        const cd d1_0 = std::polar(1.0, -delta / 2), d1_1 = std::polar(1.0, delta / 2);
        const cd d2_0 = std::polar(1.0, gamma - alpha / 2), d2_1 = std::polar(1.0, gamma + alpha / 2);
        // U' = D2 R D1 with R = [[c, -s], [s, c]]
        const cd r00 = d2_0 * c * d1_0, r01 = d2_0 * (-sn) * d1_1, r10 = d2_1 * sn * d1_0, r11 = d2_1 * c * d1_1;
        const double err = std::max(std::max(std::abs(r00 - u00), std::abs(r01 - u01)),
                                    std::max(std::abs(r10 - u10), std::abs(r11 - u11)));
        if (!(err <= 1e-13)) {
            out.push_back(std::move(o));
            continue;
        }
        LOp p1;
        p1.kind = TOP_DIAG;
        p1.k = 1;
        p1.bits[0] = o.bits[0];
        p1.coef = {d1_0, d1_1};
        p1.smask = o.smask;
        LOp rot = o;
        rot.kind = TOP_DENSE1;
        rot.flags |= TF_REAL;
        rot.coef = {cd(c), cd(-sn), cd(sn), cd(c)};
        LOp p2 = p1;
        p2.coef = {d2_0, d2_1};
        out.push_back(std::move(p1));
        out.push_back(std::move(rot));
        out.push_back(std::move(p2));
    }
    // Move each one-bit phase later past the ops it commutes with (anything
    // acting diagonally on its bit or on other states) and fold it into the
    // next one-bit phase on the same bit, so that consecutive rotations on a
    // wire share one phase between them.
    auto is_phase = [](const LOp &o) {
        return o.kind == TOP_DIAG && o.k == 1 && o.cmask == 0 && !o.is_dot() && o.coef.size() == 2;
    };
    std::vector<char> gone(out.size(), 0);
    for (size_t i = 0; i < out.size(); ++i) {
        if (!is_phase(out[i])) continue;
        const uint64_t bit = 1ull << out[i].bits[0];
        const size_t lim = std::min(out.size(), i + 4096);
        for (size_t j = i + 1; j < lim; ++j) {
            const LOp &o = out[j];
            if (gone[j]) continue;
            if (o.is_dot()) break;
            if (!(o.smask & out[i].smask)) continue;
            if (is_phase(o) && o.bits[0] == out[i].bits[0] && o.smask == out[i].smask) {
                out[j].coef[0] *= out[i].coef[0];
                out[j].coef[1] *= out[i].coef[1];
                gone[i] = 1;
                break;
            }
            if ((o.nmask() & bit) || o.smask != out[i].smask) break;
        }
    }
    prog.clear();
    for (size_t i = 0; i < out.size(); ++i)
        if (!gone[i]) prog.push_back(std::move(out[i]));
}

bool try_merge_1q(std::vector<LOp> &queue, const LOp &op) {
    if (!single_bit_1q(op)) return false;
    const uint64_t mine = touch_mask(op);
    const int lim = std::max(0, static_cast<int>(queue.size()) - 256);
    for (int i = static_cast<int>(queue.size()) - 1; i >= lim; --i) {
        LOp &q = queue[i];
        if (!(touch_mask(q) & mine)) continue;
        if (!single_bit_1q(q) || q.bits[0] != op.bits[0] || q.cmask != op.cmask || q.cval != op.cval ||
            q.smask != op.smask)
            return false;
        const auto a = as_2x2(op), b = as_2x2(q);
        std::vector<cd> m = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2],
                             a[2] * b[1] + a[3] * b[3]};
        if (m[1] == cd(0.0) && m[2] == cd(0.0)) {
            q.kind = TOP_DIAG;
            q.coef = {m[0], m[3]};
        } else {
            q.kind = TOP_DENSE1;
            q.coef = m;
        }
        return true;
    }
    return false;
}

namespace {
// Ops of one stage of a register-kernel pass.
struct StagePlan {
    bool reg = false;
    uint32_t R = 0;       // tile-local register bits (register stages)
    std::vector<int> ops; // indices into the program
};

// Two ops must keep their order when a bit they share is acted on
// non-diagonally by one of them; read-only dots commute with each other.
bool conflicts(const LOp &a, const LOp &b) {
    if (a.is_dot() && b.is_dot()) return false;
    const uint64_t na = a.nmask(), da = a.dmask() & ~na, nb = b.nmask(), db = b.dmask() & ~nb;
    return ((na & (nb | db)) | (da & nb)) != 0;
}

bool reg_capable(const LOp &o) {
    if (o.flags & TF_ZERO_UNCTRL) return false;
    switch (o.kind) {
    case TOP_DENSE1:
    case TOP_DENSE2:
    case TOP_PARITY:
    case TOP_PCPHASE: return true;
    case TOP_DIAG: return o.k <= 2;
    default: return false;
    }
}

// Split a pass's ops into register stages (each with at most rb register
// bits among its non-diagonal targets) and LDS stages, respecting the order
// constraints between ops. Register bits of a stage are padded with the
// previous stage's bits so that unchanged mappings need no transpose.
template <typename LocalF>
std::vector<StagePlan> plan_stages(const std::vector<LOp> &ops, const std::vector<int> &sel, LocalF local, int rb,
                                   int kt) {
    const int m = static_cast<int>(sel.size());
    std::vector<std::vector<int>> succ(m);
    std::vector<int> indeg(m, 0);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < j; ++i)
            if (conflicts(ops[sel[i]], ops[sel[j]])) {
                succ[i].push_back(j);
                ++indeg[j];
            }
    std::vector<uint32_t> nl(m);
    std::vector<char> rc(m);
    for (int i = 0; i < m; ++i) {
        nl[i] = local(ops[sel[i]].nmask());
        rc[i] = reg_capable(ops[sel[i]]) ? 1 : 0;
    }
    std::set<int> ready;
    for (int i = 0; i < m; ++i)
        if (indeg[i] == 0) ready.insert(i);
    int left = m;
    auto take = [&](int i) {
        ready.erase(i);
        --left;
        for (int s : succ[i])
            if (--indeg[s] == 0) ready.insert(s);
    };
    uint32_t prevR = ((1u << kt) - 1u) & ~0xFFu; // the load mapping's register bits
    std::vector<StagePlan> out;
    while (left > 0) {
        StagePlan sp;
        uint32_t U = 0;
        while (true) {
            int pick = -1;
            for (int i : ready)
                if (rc[i] && (nl[i] & ~U) == 0) {
                    pick = i;
                    break;
                }
            if (pick < 0)
                for (int i : ready)
                    if (rc[i] && popc(U | nl[i]) <= rb) {
                        pick = i;
                        break;
                    }
            if (pick < 0) break;
            U |= nl[pick];
            sp.ops.push_back(sel[pick]);
            take(pick);
        }
        if (!sp.ops.empty()) {
            sp.reg = true;
            uint32_t R = U;
            for (int b = 0; b < kt && popc(R) < rb; ++b)
                if ((prevR >> b) & 1u) R |= 1u << b;
            for (int b = kt - 1; b >= 0 && popc(R) < rb; --b) R |= 1u << b;
            sp.R = R;
            prevR = R;
        } else {
            while (true) {
                int pick = -1;
                for (int i : ready)
                    if (!rc[i]) {
                        pick = i;
                        break;
                    }
                if (pick < 0) break;
                sp.ops.push_back(sel[pick]);
                take(pick);
            }
            LQ_ABORT_IF(sp.ops.empty(), "stage planning made no progress");
        }
        out.push_back(std::move(sp));
    }
    return out;
}

// Targets as register-index bits, controls and diagonal bits as global masks;
// a diagonal +-1 table becomes a sign flip.
void to_register_form(TileOp &t, const LOp &o, const TileStage &st, const int *pos, int rb) {
    auto ridx = [&](int lb) -> int {
        for (int i = 0; i < rb; ++i)
            if (st.rbits[i] == lb) return i;
        LQ_ABORT("register bit not found");
    };
    t.gcmask = o.cmask;
    t.gcval = o.cval & o.cmask;
    if (o.kind == TOP_DENSE1) {
        t.lpos[0] = ridx(pos[o.bits[0]]);
    } else if (o.kind == TOP_DENSE2) {
        t.lpos[0] = ridx(pos[o.bits[0]]);
        t.lpos[1] = ridx(pos[o.bits[1]]);
    } else if (o.kind == TOP_PARITY) {
        t.gpmask = o.pmask;
    } else if (o.kind == TOP_DIAG && !o.is_dot()) {
        const cd one(1.0), mone(-1.0);
        if (o.k == 1 && o.coef[0] == one && o.coef[1] == mone) {
            t.kind = TOP_SIGN;
            t.gcmask |= 1ull << o.bits[0];
            t.gcval |= 1ull << o.bits[0];
        } else if (o.k == 1 && o.coef[0] == mone && o.coef[1] == one) {
            t.kind = TOP_SIGN;
            t.gcmask |= 1ull << o.bits[0];
        } else if (o.k == 0 && o.coef[0] == mone) {
            t.kind = TOP_SIGN;
        }
    }
}

bool diagonal_kind(uint32_t k) {
    return k == TOP_DIAG || k == TOP_SIGN || k == TOP_CZSET || k == TOP_PARITY || k == TOP_PCPHASE ||
           k == TOP_RPHASE || k == TOP_TPHASE;
}

// Order of a register stage's ops by earliest level: an op's level is one
// more than that of the latest earlier op it does not commute with. Ops
// commute when they act on different states, or when neither acts
// non-diagonally on a bit the other touches. Dots and ops without a known
// bit set keep their place relative to everything.
std::vector<int> stage_levels_order(const std::vector<TileOp> &ops) {
    const int M = static_cast<int>(ops.size());
    const uint64_t ALL = ~0ull;
    std::vector<uint64_t> nd(static_cast<size_t>(M)), dg(static_cast<size_t>(M));
    for (int x = 0; x < M; ++x) {
        const TileOp &t = ops[x];
        uint64_t a = 0, d = t.gcmask;
        if (t.flags & TF_DOT) {
            a = ALL;
        } else if (t.kind == TOP_DENSE1) {
            a = 1ull << t.gpos[0];
        } else if (t.kind == TOP_DENSE2) {
            a = (1ull << t.gpos[0]) | (1ull << t.gpos[1]);
        } else if (t.kind == TOP_DIAG || t.kind == TOP_PCPHASE || t.kind == TOP_TPHASE) {
            for (uint32_t j = 0; j < t.k; ++j) d |= 1ull << t.gpos[j];
        } else if (t.kind == TOP_PARITY) {
            d |= t.gpmask;
        } else if (t.kind == TOP_SIGN) {
        } else {
            a = ALL;
        }
        nd[x] = a;
        dg[x] = d;
    }
    auto commute = [&](int x, int y) {
        if (nd[x] == ALL || nd[y] == ALL) return false;
        if (!(ops[x].smask & ops[y].smask)) return true;
        return !(nd[x] & (nd[y] | dg[y])) && !(nd[y] & (nd[x] | dg[x]));
    };
    std::vector<int> level(static_cast<size_t>(M), 0);
    for (int y = 0; y < M; ++y)
        for (int x = 0; x < y; ++x)
            if (!commute(x, y)) level[y] = std::max(level[y], level[x] + 1);
    std::vector<int> perm(static_cast<size_t>(M));
    for (int x = 0; x < M; ++x) perm[x] = x;
    std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return level[x] < level[y]; });
    return perm;
}

// In a register stage the non-diagonal ops act on register bits only, so a
// one-bit phase on any other bit (a thread or outside bit: fixed per thread)
// commutes with every op of the stage but a dot or an op of unknown action.
// Between two such blockers, the phases of one state set become one
// TOP_TPHASE op at the first one's place: each thread forms the product of
// their factors once and multiplies its amplitudes once, instead of once per
// phase. Coefficients are copied (pairs d0, d1 per phase) with their sources.
void group_thread_phases(PassPlan &p, std::vector<TileOp> &ops, const int *ridx) {
    auto is_tphase = [&](const TileOp &t) {
        return t.kind == TOP_DIAG && t.k == 1 && t.gcmask == 0 && !(t.flags & TF_DOT) && t.gpos[0] >= 0 &&
               t.gpos[0] < 64 && ridx[t.gpos[0]] < 0;
    };
    auto blocker = [&](const TileOp &t) {
        if (t.flags & TF_DOT) return true;
        return !diagonal_kind(t.kind) && t.kind != TOP_DENSE1 && t.kind != TOP_DENSE2;
    };
    int count = 0;
    for (const TileOp &t : ops) count += is_tphase(t);
    if (count < 2) return;
    std::vector<TileOp> out;
    std::vector<std::pair<uint32_t, size_t>> open; // (state mask, index into out)
    for (const TileOp &t : ops) {
        if (blocker(t)) open.clear();
        if (!is_tphase(t)) {
            out.push_back(t);
            continue;
        }
        size_t at = out.size();
        bool found = false;
        for (const auto &g : open)
            if (g.first == t.smask && out[g.second].k < static_cast<uint32_t>(kMaxTileOpTargets)) {
                at = g.second;
                found = true;
            }
        if (!found) {
            TileOp n{};
            n.kind = TOP_TPHASE;
            n.smask = t.smask;
            n.k = 0;
            for (int j = 0; j < kMaxTileOpTargets; ++j) n.lpos[j] = n.gpos[j] = -1;
            n.coef_off = p.coef.size();
            // room for a full op: coefficients of one op are contiguous
            p.coef.insert(p.coef.end(), 2 * kMaxTileOpTargets, cd(1.0));
            p.coef_src.insert(p.coef_src.end(), 2 * kMaxTileOpTargets, PassPlan::CoefSrc{PassPlan::kNoSrc, 0});
            out.push_back(n);
            open.emplace_back(t.smask, at);
        }
        TileOp &g = out[at];
        const uint64_t dst = g.coef_off + 2 * static_cast<uint64_t>(g.k);
        for (int c = 0; c < 2; ++c) {
            p.coef[dst + c] = p.coef[t.coef_off + c];
            p.coef_src[dst + c] = p.coef_src[t.coef_off + c];
        }
        g.gpos[g.k] = t.gpos[0];
        ++g.k;
    }
    ops.swap(out);
}

// In each register stage, one-bit phases on register bits commute with
// diagonal ops and with dense ops on other register bits, so each can sit
// anywhere between the ops that block it (a dense op on its bit, a dot, any
// other non-diagonal op). Choose the fewest points that every phase's range
// contains (interval stabbing, greedy by right end) and replace the phases at
// each point by one table over the stage's register index.
void group_register_phases(PassPlan &p, int rb) {
    std::vector<TileOp> out;
    const int E = 1 << rb;
    for (auto &st : p.stages) {
        const uint32_t begin = static_cast<uint32_t>(out.size());
        std::vector<TileOp> ops(p.ops.begin() + st.op_begin, p.ops.begin() + st.op_end);
        if (st.kind == STAGE_REG) {
            int ridx[64];
            for (int b = 0; b < 64; ++b) ridx[b] = -1;
            for (int i = 0; i < rb; ++i) ridx[p.tbits[st.rbits[i]]] = i;
            group_thread_phases(p, ops, ridx);
            auto is_phase = [&](const TileOp &t) {
                return t.kind == TOP_DIAG && t.k == 1 && t.gcmask == 0 && !(t.flags & TF_DOT) && t.gpos[0] >= 0 &&
                       t.gpos[0] < 64 && ridx[t.gpos[0]] >= 0;
            };
            auto blocks = [&](const TileOp &l, int r, uint32_t sm) {
                if (l.flags & TF_DOT) return true;
                if (diagonal_kind(l.kind)) return false;
                if (!(l.smask & sm)) return false;
                if (l.kind == TOP_DENSE1) return l.lpos[0] == r;
                if (l.kind == TOP_DENSE2) return l.lpos[0] == r || l.lpos[1] == r;
                return true;
            };
            std::vector<TileOp> rest;      // non-phase ops; gap g sits before rest[g]
            struct Ph { int lo, hi, r; uint32_t sm; cd d0, d1; int at; uint64_t src; };
            std::vector<Ph> ph;
            for (const TileOp &t : ops) {
                if (is_phase(t)) {
                    ph.push_back({0, 0, ridx[t.gpos[0]], t.smask, p.coef[t.coef_off], p.coef[t.coef_off + 1],
                                  static_cast<int>(rest.size()), t.coef_off});
                } else {
                    rest.push_back(t);
                }
            }
            const int M = static_cast<int>(rest.size());
            // each phase's blockers (indices into rest) before and after it
            std::vector<int> prev_b(ph.size(), -1), next_b(ph.size(), M);
            for (size_t j = 0; j < ph.size(); ++j) {
                for (int x = ph[j].at - 1; x >= 0; --x)
                    if (blocks(rest[x], ph[j].r, ph[j].sm)) { prev_b[j] = x; break; }
                for (int x = ph[j].at; x < M; ++x)
                    if (blocks(rest[x], ph[j].r, ph[j].sm)) { next_b[j] = x; break; }
            }
            // Reorder the other ops by their earliest level (ops that commute
            // may pass each other), so that ops on different bits at the same
            // depth sit together and the phases between them share one point.
            std::vector<int> perm = stage_levels_order(rest);
            std::vector<int> newpos(static_cast<size_t>(M));
            for (int x = 0; x < M; ++x) newpos[perm[x]] = x;
            {
                std::vector<TileOp> re(static_cast<size_t>(M));
                for (int x = 0; x < M; ++x) re[x] = rest[perm[x]];
                rest.swap(re);
            }
            for (size_t j = 0; j < ph.size(); ++j) {
                ph[j].lo = prev_b[j] < 0 ? 0 : newpos[prev_b[j]] + 1;
                ph[j].hi = next_b[j] >= M ? M : newpos[next_b[j]];
            }
            std::vector<int> order(ph.size());
            for (size_t i = 0; i < ph.size(); ++i) order[i] = static_cast<int>(i);
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ph[x].hi < ph[y].hi; });
            std::vector<int> point(ph.size(), -1);
            for (int i : order) {
                if (point[i] >= 0) continue;
                const int g = ph[i].hi;
                for (size_t j = 0; j < ph.size(); ++j)
                    if (point[j] < 0 && ph[j].sm == ph[i].sm && ph[j].lo <= g && g <= ph[j].hi) point[j] = g;
            }
            std::vector<TileOp> merged;
            for (int g = 0; g <= M; ++g) {
                std::vector<uint32_t> sms;
                for (size_t j = 0; j < ph.size(); ++j)
                    if (point[j] == g && std::find(sms.begin(), sms.end(), ph[j].sm) == sms.end()) sms.push_back(ph[j].sm);
                for (uint32_t sm : sms) {
                    std::vector<cd> table(static_cast<size_t>(E), cd(1.0));
                    TileOp t{};
                    PassPlan::PhaseTable rec;
                    rec.off = p.coef.size();
                    rec.rb = rb;
                    for (size_t j = 0; j < ph.size(); ++j) {
                        if (point[j] != g || ph[j].sm != sm) continue;
                        for (int e = 0; e < E; ++e) table[e] *= ((e >> ph[j].r) & 1) ? ph[j].d1 : ph[j].d0;
                        rec.phases.emplace_back(ph[j].src, ph[j].r);
                    }
                    t.kind = TOP_RPHASE;
                    t.k = static_cast<uint32_t>(rb);
                    t.smask = sm;
                    for (int j = 0; j < kMaxTileOpTargets; ++j) t.lpos[j] = t.gpos[j] = -1;
                    t.coef_off = p.coef.size();
                    p.coef.insert(p.coef.end(), table.begin(), table.end());
                    p.coef_src.insert(p.coef_src.end(), table.size(), PassPlan::CoefSrc{PassPlan::kNoSrc, 0});
                    p.tables.push_back(std::move(rec));
                    merged.push_back(t);
                }
                if (g < M) merged.push_back(rest[g]);
            }
            ops.swap(merged);
        }
        out.insert(out.end(), ops.begin(), ops.end());
        st.op_begin = begin;
        st.op_end = static_cast<uint32_t>(out.size());
    }
    p.ops.swap(out);
}

// Merge runs of CZ sign flips into popcount-parity ops and set how each
// stage's register mapping is reached.
void finish_reg_pass(PassPlan &p, int kt) {
    group_register_phases(p, kt - 8);
    std::vector<TileOp> merged;
    for (auto &st : p.stages) {
        const uint32_t begin = static_cast<uint32_t>(merged.size());
        TileOp cur{};
        bool open = false;
        auto flush = [&]() {
            if (open) merged.push_back(cur);
            open = false;
        };
        for (uint32_t i = st.op_begin; i < st.op_end; ++i) {
            const TileOp &t = p.ops[i];
            const bool cz = st.kind == STAGE_REG && t.kind == TOP_SIGN && !(t.flags & TF_DOT) &&
                            popc(t.gcmask) == 2 && t.gcval == t.gcmask;
            if (!cz) {
                flush();
                merged.push_back(t);
                continue;
            }
            const int lo = __builtin_ctzll(t.gcmask), hi = 63 - __builtin_clzll(t.gcmask);
            const uint32_t d = static_cast<uint32_t>(hi - lo);
            int slot = -1;
            if (open && cur.smask == t.smask) {
                for (uint32_t z = 0; z < cur.nz; ++z)
                    if (cur.sd[z] == d) slot = static_cast<int>(z);
                if (slot < 0 && cur.nz < 4) slot = static_cast<int>(cur.nz++);
            }
            if (slot < 0) {
                flush();
                cur = t;
                cur.kind = TOP_CZSET;
                cur.gcmask = cur.gcval = 0;
                cur.nz = 1;
                cur.sd[0] = d;
                cur.sm[0] = 0;
                for (int z = 1; z < 4; ++z) {
                    cur.sd[z] = 0;
                    cur.sm[z] = 0;
                }
                open = true;
                slot = 0;
            }
            cur.sd[slot] = d;
            cur.sm[slot] |= 1ull << lo;
        }
        flush();
        st.op_begin = begin;
        st.op_end = static_cast<uint32_t>(merged.size());
    }
    p.ops.swap(merged);

    const uint32_t m0 = ((1u << kt) - 1u) & ~0xFFu;
    uint32_t prevR = m0;
    bool prev_lds = false;
    for (auto &st : p.stages) {
        if (st.kind != STAGE_REG) {
            prev_lds = true;
            continue;
        }
        uint32_t R = 0;
        for (int i = 0; i < kt - 8; ++i) R |= 1u << st.rbits[i];
        st.entry = prev_lds ? ENTRY_FROM_LDS : (R == prevR ? ENTRY_NONE : ENTRY_TRANSPOSE);
        prevR = R;
        prev_lds = false;
    }
    p.end_transpose = (!prev_lds && prevR != m0) ? 1u : 0u;
}
} // namespace

TileGeom make_geom(int n, const std::vector<int> &tbits) {
    TileGeom g{};
    g.n = n;
    g.kt = static_cast<int>(tbits.size());
    g.nout = n - g.kt;
    LQ_ABORT_IF(g.kt > 16 || g.nout > kMaxOutsideBits, "tile geometry out of range");
    uint64_t tm = 0;
    for (int j = 0; j < g.kt; ++j) {
        g.tb[j] = tbits[j];
        tm |= 1ull << tbits[j];
    }
    g.c = 0;
    while (g.c < g.kt && g.tb[g.c] == g.c) ++g.c;
    int o = 0;
    for (int b = 0; b < n; ++b)
        if (!((tm >> b) & 1ull)) g.ob[o++] = b;
    return g;
}

namespace {
constexpr double kMinCos = 1e-100;

// Convert the recorded (or, at plan time, the eligible) real rotations of a
// pass to [[1, -t], [t, 1]] and multiply the product of their cosines into
// the pass's scale table. Returns false if the recorded choice no longer
// holds for the current angles (a cosine too small for the stored form, or
// the growth bound now crossed in either direction).
bool apply_rotation_scaling(PassPlan &p, bool at_plan_time, int rb, double max_growth) {
    if (at_plan_time) {
        p.scaled.clear();
        p.scale_dropped.clear();
        p.scale_table = -1;
        if (!p.reg || p.load_mask != 1u || p.store_mask != 1u || !p.slot_ids.empty() || p.stages.empty() ||
            p.stages.back().kind != STAGE_REG)
            return true;
        for (const TileStage &st : p.stages) {
            if (st.kind != STAGE_REG) continue;
            for (uint32_t i = st.op_begin; i < st.op_end; ++i) {
                TileOp &t = p.ops[i];
                if (t.kind != TOP_DENSE1 || !(t.flags & TF_REAL) || (t.flags & TF_DOT) || t.smask != 1u ||
                    t.gcmask != 0 || t.lcmask != 0 || t.ocmask != 0)
                    continue;
                if (std::fabs(p.coef[t.coef_off].real()) < kMinCos) continue;
                t.flags |= TF_SCALED;
                p.scaled.push_back(t.coef_off);
            }
        }
        if (p.scaled.empty()) return true;
        p.scale_growth_limit = max_growth;
        double growth = 1.0;
        for (uint64_t off : p.scaled) growth /= std::fabs(p.coef[off].real());
        if (!(growth <= max_growth)) {
            // too much growth for these angles: the full form throughout; a
            // refresh whose angles would allow the scaled form asks for a new plan
            for (const TileStage &st : p.stages)
                for (uint32_t i = st.op_begin; i < st.op_end; ++i) p.ops[i].flags &= ~static_cast<uint32_t>(TF_SCALED);
            p.scale_dropped.swap(p.scaled);
            return true;
        }
        // the carrier: the last one-state phase table, or a new one at the end
        for (const TileStage &st : p.stages) {
            if (st.kind != STAGE_REG) continue;
            for (uint32_t i = st.op_begin; i < st.op_end; ++i)
                if (p.ops[i].kind == TOP_RPHASE && p.ops[i].smask == 1u) p.scale_table = static_cast<int64_t>(p.ops[i].coef_off);
        }
        if (p.scale_table < 0) {
            TileOp t{};
            t.kind = TOP_RPHASE;
            t.k = static_cast<uint32_t>(rb);
            t.smask = 1u;
            for (int j = 0; j < kMaxTileOpTargets; ++j) t.lpos[j] = t.gpos[j] = -1;
            t.coef_off = p.coef.size();
            const size_t E = static_cast<size_t>(1) << rb;
            p.coef.insert(p.coef.end(), E, cd(1.0));
            p.coef_src.insert(p.coef_src.end(), E, PassPlan::CoefSrc{PassPlan::kNoSrc, 0});
            PassPlan::PhaseTable rec;
            rec.off = t.coef_off;
            rec.rb = rb;
            p.tables.push_back(rec);
            p.scale_table = static_cast<int64_t>(t.coef_off);
            LQ_ABORT_IF(p.stages.back().op_end != p.ops.size(), "last stage does not end the pass");
            p.ops.push_back(t);
            p.stages.back().op_end = static_cast<uint32_t>(p.ops.size());
        }
    }
    auto growth_of = [&](const std::vector<uint64_t> &offs) {
        double g = 1.0;
        for (uint64_t off : offs) g /= std::fabs(p.coef[off].real());
        return g;
    };
    if (!p.scale_dropped.empty() && growth_of(p.scale_dropped) <= p.scale_growth_limit) return false;
    if (p.scaled.empty()) return true;
    if (!(growth_of(p.scaled) <= p.scale_growth_limit)) return false;
    double prod = 1.0;
    for (uint64_t off : p.scaled) {
        const double c = p.coef[off].real(), s = p.coef[off + 2].real();
        if (std::fabs(c) < kMinCos) return false;
        const double t = s / c;
        p.coef[off] = cd(1.0);
        p.coef[off + 1] = cd(-t);
        p.coef[off + 2] = cd(t);
        p.coef[off + 3] = cd(1.0);
        prod *= c;
    }
    // the table's size from its record
    int trb = 0;
    for (const auto &tb : p.tables)
        if (static_cast<int64_t>(tb.off) == p.scale_table) trb = tb.rb;
    LQ_ABORT_IF(trb <= 0, "scale table not found");
    for (int e = 0; e < (1 << trb); ++e) p.coef[static_cast<size_t>(p.scale_table) + static_cast<size_t>(e)] *= prod;
    return true;
}
} // namespace

std::vector<PassPlan> plan_passes(const std::vector<LOp> &ops, const PlanConfig &cfg) {
    const int N = static_cast<int>(ops.size());
    const int n = cfg.n;
    std::vector<std::vector<int>> succ(N);
    std::vector<int> indeg(N, 0), mark(N, -1);
    std::vector<uint64_t> nm(N), dm(N);
    struct BitDeps {
        int lastN = -1;
        std::vector<int> D, RN, RD;
    };
    std::vector<BitDeps> bd(static_cast<size_t>(n));
    for (int j = 0; j < N; ++j) {
        nm[j] = ops[j].nmask();
        dm[j] = ops[j].dmask() & ~nm[j];
        LQ_ABORT_IF(((nm[j] | dm[j]) >> n) != 0, "operation acts on a wire outside the state");
        const bool dot = ops[j].is_dot();
        auto dep = [&](int i) {
            if (i >= 0 && mark[i] != j) {
                mark[i] = j;
                succ[i].push_back(j);
                ++indeg[j];
            }
        };
        for (int b = 0; b < n; ++b) {
            const uint64_t bit = 1ull << b;
            BitDeps &s = bd[b];
            if (nm[j] & bit) {
                dep(s.lastN);
                if (!dot) {
                    for (int i : s.D) dep(i);
                    for (int i : s.RN) dep(i);
                    for (int i : s.RD) dep(i);
                    s.lastN = j;
                    s.D.clear();
                    s.RN.clear();
                    s.RD.clear();
                } else {
                    for (int i : s.D) dep(i);
                    s.RN.push_back(j);
                }
            } else if (dm[j] & bit) {
                dep(s.lastN);
                if (!dot) {
                    for (int i : s.RN) dep(i);
                    s.D.push_back(j);
                } else {
                    s.RD.push_back(j);
                }
            }
        }
    }

    std::set<int> ready;
    for (int j = 0; j < N; ++j)
        if (indeg[j] == 0) ready.insert(j);

    const int ktmax = std::min(cfg.kt_max, n);
    const int low = std::max(0, std::min(cfg.low_bits, ktmax));
    const uint64_t low_mask = (low >= 64) ? ~0ull : ((1ull << low) - 1);

    // Lookahead choice of a pass's tile bits. run_with(T) replays, without
    // committing, the schedule the pass would make with tile bits T (ready ops
    // in program order, within the cost budget and dot slots) and returns the
    // cost it fits; ops held back only by bits outside T are collected as the
    // frontier. Starting from the low bits, the missing bits of the frontier
    // op whose addition lets the most work run per added bit are added until
    // the tile is full or the budget is met. On a chain of two-qubit gates
    // this grows contiguous blocks of wires, which advance several layers per
    // pass, where adding the first ready op's bits picks every other wire and
    // advances two.
    std::vector<int> touched;
    double budget = cfg.budget; // this pass's
    auto run_with = [&](uint64_t T, std::vector<int> *frontier, bool *full) {
        std::priority_queue<int, std::vector<int>, std::greater<int>> q(ready.begin(), ready.end());
        double cost = 0.0;
        int dots = 0, done = 0;
        *full = false;
        while (!q.empty()) {
            const int i = q.top();
            q.pop();
            if ((nm[i] & ~T) != 0 || (dots >= cfg.max_dots && ops[i].is_dot())) {
                if (frontier) frontier->push_back(i);
                continue;
            }
            if (done > 0 && cost + ops[i].cost() > budget) {
                *full = true;
                break;
            }
            cost += ops[i].cost();
            dots += ops[i].is_dot() ? 1 : 0;
            ++done;
            for (int s2 : succ[i]) {
                touched.push_back(s2);
                if (--indeg[s2] == 0) q.push(s2);
            }
        }
        for (int s2 : touched) ++indeg[s2];
        touched.clear();
        return cost;
    };
    auto choose_bits = [&]() {
        uint64_t T = low_mask;
        std::vector<int> frontier;
        while (popc(T) < ktmax) {
            frontier.clear();
            bool full = false;
            const double base = run_with(T, &frontier, &full);
            if (full) break;
            uint64_t best = 0;
            double best_score = 0.0;
            int best_bits = 0;
            std::vector<uint64_t> seen;
            for (int i : frontier) {
                const uint64_t m = nm[i] & ~T;
                if (m == 0 || popc(T | m) > ktmax || std::find(seen.begin(), seen.end(), m) != seen.end()) continue;
                seen.push_back(m);
                bool f2 = false;
                const double score = (run_with(T | m, nullptr, &f2) - base) / popc(m);
                if (score > best_score + 1e-12 || (best != 0 && std::fabs(score - best_score) <= 1e-12 && popc(m) < best_bits)) {
                    best = m;
                    best_score = score;
                    best_bits = popc(m);
                }
            }
            if (best == 0) break;
            T |= best;
        }
        return T;
    };

    std::vector<PassPlan> passes;
    int remaining = N;
    while (remaining > 0) {
        std::vector<int> sel;
        uint64_t T = 0;
        budget = (passes.empty() && cfg.first_budget > 0) ? cfg.first_budget : cfg.budget;
        const uint64_t T0 = cfg.lookahead ? choose_bits() : low_mask;
        for (int attempt = 0; attempt < 2 && sel.empty(); ++attempt) {
            T = attempt == 0 ? T0 : 0ull;
            double cost = 0.0;
            int dots = 0;
            while (true) {
                int pick = -1;
                const bool dots_full = dots >= cfg.max_dots;
                for (int i : ready)
                    if ((nm[i] & ~T) == 0 && !(dots_full && ops[i].is_dot())) {
                        pick = i;
                        break;
                    }
                if (pick < 0)
                    for (int i : ready)
                        if (popc(T | nm[i]) <= ktmax && !(dots_full && ops[i].is_dot())) {
                            pick = i;
                            break;
                        }
                if (pick < 0) break;
                if (!sel.empty() && cost + ops[pick].cost() > budget) break;
                T |= nm[pick];
                cost += ops[pick].cost();
                dots += ops[pick].is_dot() ? 1 : 0;
                sel.push_back(pick);
                ready.erase(pick);
                --remaining;
                for (int s : succ[pick])
                    if (--indeg[s] == 0) ready.insert(s);
            }
        }
        LQ_ABORT_IF(sel.empty(), "operation needs more tile bits than fit in LDS");
        // pad the tile with the lowest free bits
        for (int b = 0; b < n && popc(T) < ktmax; ++b) T |= 1ull << b;

        PassPlan p;
        int pos[64];
        for (int b = 0; b < 64; ++b) pos[b] = -1;
        for (int b = 0; b < n; ++b)
            if ((T >> b) & 1ull) {
                pos[b] = static_cast<int>(p.tbits.size());
                p.tbits.push_back(b);
            }
        auto local = [&](uint64_t m) {
            uint32_t r = 0;
            for (int b = 0; b < n; ++b)
                if (((m >> b) & 1ull) && pos[b] >= 0) r |= 1u << pos[b];
            return r;
        };
        const int kt = static_cast<int>(p.tbits.size());
        p.reg = cfg.reg_rb > 0 && kt == 8 + cfg.reg_rb && cfg.ns <= kMaxStates;
        std::vector<int> order = sel;
        std::vector<StagePlan> stages;
        if (p.reg) {
            stages = plan_stages(ops, sel, local, cfg.reg_rb, kt);
            order.clear();
            for (const auto &st : stages) order.insert(order.end(), st.ops.begin(), st.ops.end());
        }
        std::vector<int> op_stage;
        for (size_t si = 0; si < stages.size(); ++si)
            for (size_t k2 = 0; k2 < stages[si].ops.size(); ++k2) op_stage.push_back(static_cast<int>(si));
        size_t stage_cursor = 0;
        for (size_t oi = 0; oi < order.size(); ++oi) {
            const int idx = order[oi];
            const LOp &o = ops[idx];
            if (p.reg) {
                const int si = op_stage[oi];
                if (p.stages.size() <= static_cast<size_t>(si)) {
                    // open the stage
                    TileStage ts{};
                    ts.kind = stages[si].reg ? STAGE_REG : STAGE_LDS;
                    ts.op_begin = ts.op_end = static_cast<uint32_t>(p.ops.size());
                    int r = 0, h = 0;
                    for (int b = 0; b < kt; ++b) {
                        if ((stages[si].R >> b) & 1u) {
                            if (r < kMaxRegBits) ts.rbits[r++] = b;
                        } else if (h < 8) {
                            ts.hbits[h++] = b;
                        }
                    }
                    p.stages.push_back(ts);
                    stage_cursor = static_cast<size_t>(si);
                }
            }
            TileOp t{};
            t.kind = o.kind;
            t.flags = o.flags;
            t.k = static_cast<uint32_t>(o.k);
            t.smask = o.smask;
            t.sa = o.sa;
            t.sb = o.sb;
            t.lcmask = local(o.cmask);
            t.lcval = local(o.cval & o.cmask);
            t.ocmask = o.cmask & ~T;
            t.ocval = o.cval & o.cmask & ~T;
            for (int j = 0; j < kMaxTileOpTargets; ++j) {
                t.lpos[j] = j < o.k ? pos[o.bits[j]] : -1;
                t.gpos[j] = j < o.k ? o.bits[j] : 0;
            }
            if (o.kind == TOP_PARITY) {
                t.lxmask = local(o.pmask);
                t.omask = o.pmask & ~T;
            } else if (o.kind == TOP_PAULI) {
                t.lxmask = local(o.xmask);
                t.lzmask = local(o.yzmask);
                t.omask = o.yzmask & ~T;
                t.ny = static_cast<uint32_t>(o.ny);
            }
            t.pc_dim = o.pc_dim;
            t.coef_off = p.coef.size();
            p.coef.insert(p.coef.end(), o.coef.begin(), o.coef.end());
            for (size_t c = 0; c < o.coef.size(); ++c)
                p.coef_src.push_back({static_cast<uint32_t>(idx), static_cast<uint32_t>(c)});
            if (o.is_dot()) {
                t.slot = static_cast<uint32_t>(p.slot_ids.size());
                p.slot_ids.push_back(o.slot);
                p.load_mask |= (1u << o.sa) | (1u << o.sb);
            } else {
                p.load_mask |= o.smask;
                p.store_mask |= o.smask;
            }
            if (p.reg && p.stages[stage_cursor].kind == STAGE_REG)
                to_register_form(t, o, p.stages[stage_cursor], pos, cfg.reg_rb);
            p.ops.push_back(t);
            if (p.reg) p.stages[stage_cursor].op_end = static_cast<uint32_t>(p.ops.size());
        }
        if (p.reg) finish_reg_pass(p, kt);
        if (p.reg && cfg.scale_rotations) apply_rotation_scaling(p, true, kt - 8, cfg.max_scale_growth);
        passes.push_back(std::move(p));
    }
    return passes;
}

std::string plan_structure_key(const std::vector<LOp> &ops, const PlanConfig &cfg) {
    std::string key;
    auto put = [&key](const void *p, size_t n) { key.append(static_cast<const char *>(p), n); };
    const double b = cfg.budget;
    const int32_t hdr[8] = {cfg.n, cfg.ns, cfg.kt_max, cfg.low_bits, cfg.reg_rb, cfg.max_dots,
                            cfg.scale_rotations ? 1 : 0, cfg.lookahead ? 1 : 0};
    put(hdr, sizeof(hdr));
    put(&b, sizeof(b));
    put(&cfg.first_budget, sizeof(cfg.first_budget));
    put(&cfg.max_scale_growth, sizeof(cfg.max_scale_growth));
    const cd one(1.0), mone(-1.0);
    for (const LOp &o : ops) {
        // value tests made while planning (to_register_form)
        uint32_t vflags = 0;
        if (o.kind == TOP_DENSE1 && (o.flags & TF_REAL) && o.coef.size() == 4 && std::fabs(o.coef[0].real()) < kMinCos)
            vflags = 4;
        if (o.kind == TOP_DIAG && !o.is_dot()) {
            if (o.k == 1 && o.coef.size() == 2 && o.coef[0] == one && o.coef[1] == mone) vflags = 1;
            else if (o.k == 1 && o.coef.size() == 2 && o.coef[0] == mone && o.coef[1] == one) vflags = 2;
            else if (o.k == 0 && o.coef.size() == 1 && o.coef[0] == mone) vflags = 3;
        }
        const uint32_t w[8] = {o.kind, o.flags, static_cast<uint32_t>(o.k), o.smask, o.sa, o.sb,
                               static_cast<uint32_t>(o.slot), vflags};
        put(w, sizeof(w));
        put(o.bits, sizeof(int) * static_cast<size_t>(std::max(0, std::min(o.k, kMaxTileOpTargets))));
        const uint64_t m[7] = {o.cmask, o.cval, o.xmask, o.yzmask, o.pmask, o.pc_dim,
                               static_cast<uint64_t>(o.coef.size()) | (static_cast<uint64_t>(o.ny) << 32)};
        put(m, sizeof(m));
    }
    return key;
}

bool refresh_plan_coefs(std::vector<PassPlan> &passes, const std::vector<LOp> &ops) {
    for (PassPlan &p : passes) {
        LQ_ABORT_IF(p.coef_src.size() != p.coef.size(), "plan without coefficient sources");
        for (size_t i = 0; i < p.coef.size(); ++i) {
            const PassPlan::CoefSrc &s = p.coef_src[i];
            if (s.op == PassPlan::kNoSrc) continue;
            p.coef[i] = ops[s.op].coef[s.idx];
        }
        for (const PassPlan::PhaseTable &t : p.tables) {
            const int E = 1 << t.rb;
            for (int e = 0; e < E; ++e) {
                cd v(1.0);
                for (const auto &ph : t.phases) v *= p.coef[ph.first + (((e >> ph.second) & 1) ? 1 : 0)];
                p.coef[t.off + static_cast<uint64_t>(e)] = v;
            }
        }
        if (!apply_rotation_scaling(p, false, 0, p.scale_growth_limit)) return false;
    }
    return true;
}

} // namespace lqamd
