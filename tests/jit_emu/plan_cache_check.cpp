// CPU check of plan reuse: a plan made for one program, with its
// coefficients refilled from a program of the same structure and other
// angles, must equal the plan made directly for that program (every op,
// stage and coefficient, bit for bit). Test infrastructure.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "lq_gates.hpp"
#include "lq_plan.hpp"

using namespace lqamd;

static std::vector<LOp> build(int n, int ns, unsigned seed_structure, unsigned seed_angles, bool split) {
    std::mt19937 rs(seed_structure), ra(seed_angles);
    std::uniform_real_distribution<double> U(-3.1, 3.1);
    const GateId gs[] = {GateId::RX, GateId::RY, GateId::RZ, GateId::Rot, GateId::PhaseShift, GateId::Hadamard,
                         GateId::PauliZ, GateId::CZ, GateId::CNOT, GateId::IsingZZ, GateId::CRY, GateId::MultiRZ,
                         GateId::SWAP, GateId::T, GateId::IsingXY};
    std::vector<LOp> prog;
    int slot = 0;
    for (int i = 0; i < 400; ++i) {
        const GateId g = gs[rs() % (sizeof(gs) / sizeof(gs[0]))];
        const auto &info = gate_info(g);
        const int nw = info.nwires > 0 ? info.nwires : 3;
        std::vector<uint32_t> w;
        while (static_cast<int>(w.size()) < nw) {
            const uint32_t x = rs() % n;
            bool dup = false;
            for (auto y : w) dup |= y == x;
            if (!dup) w.push_back(x);
        }
        std::vector<double> p(std::max(info.nparams, 1));
        for (auto &x : p) x = U(ra);
        LOp op = lower_gate(n, g, w.data(), nw, nullptr, nullptr, 0, (rs() & 1) != 0, p.data());
        op.smask = (1u << ns) - 1u;
        if (!try_merge_1q(prog, op)) prog.push_back(op);
        if (rs() % 17 == 0) {
            // a generator dot <state ns-1| G |state 0>
            double sc;
            const uint32_t q = rs() % n;
            LOp d = lower_generator(n, GateId::RY, &q, 1, nullptr, nullptr, 0, &sc);
            d.flags = (d.flags & ~TF_ZERO_UNCTRL) | TF_DOT;
            d.sa = 0;
            d.sb = static_cast<uint32_t>(ns - 1);
            d.slot = slot++;
            prog.push_back(d);
        }
    }
    if (split) split_1q_rotations(prog);
    return prog;
}

static bool same(const std::vector<PassPlan> &a, const std::vector<PassPlan> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
        const PassPlan &x = a[i], &y = b[i];
        if (x.tbits != y.tbits || x.load_mask != y.load_mask || x.store_mask != y.store_mask || x.reg != y.reg ||
            x.end_transpose != y.end_transpose || x.slot_ids != y.slot_ids || x.ops.size() != y.ops.size() ||
            x.stages.size() != y.stages.size() || x.coef.size() != y.coef.size())
            return false;
        if (std::memcmp(x.ops.data(), y.ops.data(), x.ops.size() * sizeof(TileOp)) != 0) return false;
        if (std::memcmp(x.stages.data(), y.stages.data(), x.stages.size() * sizeof(TileStage)) != 0) return false;
        for (size_t c = 0; c < x.coef.size(); ++c)
            if (x.coef[c] != y.coef[c]) return false;
    }
    return true;
}

int main() {
    int fails = 0, cases = 0, refused = 0;
    for (int n : {14, 20}) {
        for (int ns : {1, 2}) {
            for (int split = 0; split < 3; ++split) {
                for (unsigned s = 1; s <= 3; ++s) {
                    PlanConfig cfg;
                    cfg.n = n;
                    cfg.ns = ns;
                    cfg.kt_max = 12;
                    cfg.low_bits = 4;
                    cfg.budget = split ? 52 : 24;
                    cfg.lookahead = split != 0; // as the engine plans specialised passes
                    cfg.reg_rb = 4;
                    cfg.max_dots = 32;
                    cfg.scale_rotations = split != 0;
                    // split 2: the complex64 growth bound, low enough that some
                    // passes keep the full form for one set of angles only
                    if (split == 2) cfg.max_scale_growth = 1e3;
                    const std::vector<LOp> a = build(n, ns, s, 100 + s, split != 0);
                    const std::vector<LOp> b = build(n, ns, s, 200 + s, split != 0);
                    ++cases;
                    if (plan_structure_key(a, cfg) != plan_structure_key(b, cfg)) {
                        std::printf("n=%d ns=%d split=%d seed=%u: keys differ\n", n, ns, split, s);
                        ++fails;
                        continue;
                    }
                    std::vector<PassPlan> reused = plan_passes(a, cfg);
                    const bool refreshed = refresh_plan_coefs(reused, b);
                    const std::vector<PassPlan> direct = plan_passes(b, cfg);
                    // a refused refresh is answered by a new plan: only a refresh
                    // that succeeds must reproduce the direct plan
                    const bool ok = refreshed ? same(reused, direct) : split == 2;
                    if (!refreshed) ++refused;
                    // and the coefficients really changed
                    const std::vector<PassPlan> pa = plan_passes(a, cfg);
                    bool changed = false;
                    for (size_t i = 0; i < pa.size() && !changed; ++i)
                        for (size_t c = 0; c < pa[i].coef.size() && !changed; ++c) changed = pa[i].coef[c] != direct[i].coef[c];
                    std::printf("n=%d ns=%d split=%d seed=%u: %zu passes %s%s\n", n, ns, split, s, direct.size(),
                                refreshed ? (ok ? "same" : "DIFFERENT") : "refresh refused",
                                changed ? "" : " (coefficients unchanged!)");
                    fails += !ok || !changed;
                }
            }
        }
    }
    // a structural change must change the key
    PlanConfig cfg;
    cfg.n = 14;
    cfg.reg_rb = 4;
    cfg.max_dots = 32;
    std::vector<LOp> a = build(14, 1, 5, 7, true), b = a;
    b[3].bits[0] = (b[3].bits[0] + 1) % 14;
    if (plan_structure_key(a, cfg) == plan_structure_key(b, cfg)) {
        std::printf("a moved target did not change the key\n");
        ++fails;
    }
    // only the growth-bound cases may refuse, and some of them must
    if (refused == 0) {
        std::printf("no refresh was refused under the growth bound\n");
        ++fails;
    }
    std::printf("%d cases (%d refused refreshes), %d failures\n", cases, refused, fails);
    return fails ? 1 : 0;
}
