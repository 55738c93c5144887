// CPU check of the generated pass kernels: plan a program with the engine's
// planner, generate each pass's kernel source, run it under the CPU emulation
// of emu_rt.hpp / emu_main.cpp (one host thread per GPU thread, real barriers
// and wave shuffles), and compare the states and dot products with a naive
// application of the same ops (before the rotation split) in program order.
//
//   emu_check CASE DTYPE N GRID EMU_DIR WORK_DIR
//   CASE 0: hardware-efficient layers (Rot + CZ ladders) with named-generator dots
//   CASE 1: the same plus controlled, parity, phase and SWAP gates
//   CASE 2: case 1 on two states with dots between them (adjoint-like)
//   CASE 3: hardware-efficient layers, then four named expectation values
//   CASE 4: hardware-efficient layers on three states with dots of the first
//           against each of the others (observables sharing an adjoint sweep)
//   CASE 5: the same on four states
// Exit status 0 when every amplitude and dot matches (complex128 1e-10,
// complex64 1e-3 relative).
#include <dlfcn.h>
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include "lq_engine.hpp"
#include "lq_gates.hpp"
#include "lq_jit.hpp"
#include "lq_plan.hpp"
using namespace lqamd;
using St = std::vector<cd>;

static void ref_apply(std::vector<St> &S, int n, const LOp &o, std::vector<cd> &dots) {
    const uint64_t N = 1ull << n;
    auto ctl = [&](uint64_t i) { return (i & o.cmask) == (o.cval & o.cmask); };
    const bool dot = o.is_dot();
    cd acc = 0;
    if (o.kind == TOP_DENSE1 || o.kind == TOP_DENSE2 || o.kind == TOP_DENSEK) {
        const int k = o.k, K = 1 << k;
        uint64_t tm = 0;
        for (int j = 0; j < k; ++j) tm |= 1ull << o.bits[j];
        auto dep = [&](int r) { uint64_t x = 0; for (int j = 0; j < k; ++j) if ((r >> (k - 1 - j)) & 1) x |= 1ull << o.bits[j]; return x; };
        for (uint64_t i = 0; i < N; ++i) {
            if (i & tm) continue;
            if (!ctl(i)) continue;
            if (dot) {
                std::vector<cd> a(K), out(K, 0.0);
                for (int c = 0; c < K; ++c) a[c] = S[o.sa][i | dep(c)];
                for (int r = 0; r < K; ++r) for (int c = 0; c < K; ++c) out[r] += o.coef[r * K + c] * a[c];
                for (int r = 0; r < K; ++r) acc += std::conj(S[o.sb][i | dep(r)]) * out[r];
                continue;
            }
            for (size_t s = 0; s < S.size(); ++s) {
                if (!((o.smask >> s) & 1u)) continue;
                std::vector<cd> a(K), out(K, 0.0);
                for (int c = 0; c < K; ++c) a[c] = S[s][i | dep(c)];
                for (int r = 0; r < K; ++r) for (int c = 0; c < K; ++c) out[r] += o.coef[r * K + c] * a[c];
                for (int r = 0; r < K; ++r) S[s][i | dep(r)] = out[r];
            }
        }
    } else {
        for (uint64_t i = 0; i < N; ++i) {
            if (!ctl(i)) continue;
            cd f;
            if (o.kind == TOP_PARITY) f = o.coef[__builtin_popcountll(i & o.pmask) & 1];
            else {
                uint64_t idx = 0;
                for (int j = 0; j < o.k; ++j) idx |= ((i >> o.bits[j]) & 1ull) << (o.k - 1 - j);
                if (o.kind == TOP_PCPHASE) f = o.coef[idx < o.pc_dim ? 0 : 1];
                else f = o.coef[idx];
            }
            if (dot) { acc += std::conj(S[o.sb][i]) * f * S[o.sa][i]; continue; }
            for (size_t s = 0; s < S.size(); ++s) if ((o.smask >> s) & 1u) S[s][i] *= f;
        }
    }
    if (dot) dots[o.slot] = acc;
}

static std::string transform(std::string s) {
    auto rep = [&](const std::string &a, const std::string &b) {
        size_t p = 0; int c = 0;
        while ((p = s.find(a, p)) != std::string::npos) { s.replace(p, a.size(), b); p += b.size(); ++c; }
        return c;
    };
    // the kernel states its dynamic LDS bytes; emulate exactly that much
    const size_t lp = s.find("// lds_bytes ");
    const unsigned long lds_bytes = lp == std::string::npos ? 0 : std::stoul(s.substr(lp + 13));
    rep("extern __shared__ cplx lds[];",
        "static LdsArr<cplx, " + std::to_string(lds_bytes) + "ul / sizeof(cplx)> lds;");
    for (int d = 32; d <= 96; d += 32)
        rep("__shared__ double2 dred[4][" + std::to_string(d) + "];", "static double2 dred[4][" + std::to_string(d) + "];");
    rep("asm volatile(\"\" : \"+s\"(z));", "");
    rep("typedef __attribute__((address_space(4))) const double2 kcoef;", "typedef const double2 kcoef;");
    rep("typedef __attribute__((address_space(4))) const unsigned long long kgeo;", "typedef const unsigned long long kgeo;");
    rep("typedef real nvec __attribute__((ext_vector_type(2)));", "struct nvec { real x, y; };");
    return "#define LQ_EMU 1\n#include \"emu_rt.hpp\"\n" + s;
}

using Launch = void (*)(void *, void *, const double2 *, double2 *, unsigned long long, unsigned, void *, void *);

int main(int argc, char **argv) {
    const int which = atoi(argv[1]), dtype = atoi(argv[2]);
    const int n = argc > 3 ? atoi(argv[3]) : 14;
    const unsigned grid_cap = argc > 4 ? atoi(argv[4]) : 3;
    const std::string emu_dir = argc > 5 ? argv[5] : ".";
    const std::string work = argc > 6 ? argv[6] : "/tmp";
    std::mt19937 rng(5 + which);
    std::uniform_real_distribution<double> U(0, 6.28);
    std::vector<LOp> prog;
    int ns = 1;
    auto gate = [&](GateId g, std::vector<uint32_t> w, std::vector<double> p, std::vector<uint32_t> c = {}, std::vector<uint8_t> cv = {}) {
        LOp op = lower_gate(n, g, w.data(), (int)w.size(), c.data(), cv.data(), (int)c.size(), false, p.data());
        op.smask = (1u << ns) - 1u;
        if (!try_merge_1q(prog, op)) prog.push_back(op);
    };
    int nslots = 0;
    auto dot = [&](GateId g, uint32_t w, uint32_t sa, uint32_t sb) {
        double sc;
        LOp d = lower_generator(n, g, &w, 1, nullptr, nullptr, 0, &sc);
        d.flags = (d.flags & ~TF_ZERO_UNCTRL) | TF_DOT; d.sa = sa; d.sb = sb; d.slot = nslots++;
        prog.push_back(d);
    };
    if (which == 2) ns = 2;
    if (which == 4) ns = 3; // lambda and two H|lambda> (observables sharing an adjoint sweep)
    if (which == 5) ns = 4;
    if (which == 3) {
        // hardware-efficient layers, then <Z_0>, <X_{n-1}>, <Y_8>, <H_3> as the
        // engine lowers named expectation values (one dot each)
        for (int l = 0; l < 3; ++l) {
            for (int w = 0; w < n; ++w) gate(GateId::Rot, {(uint32_t)w}, {U(rng), U(rng), U(rng)});
            for (int s = 0; s < 2; ++s) for (int w = s; w < n - 1; w += 2) gate(GateId::CZ, {(uint32_t)w, (uint32_t)w + 1}, {});
        }
        const GateId obs[4] = {GateId::PauliZ, GateId::PauliX, GateId::PauliY, GateId::Hadamard};
        const int ow[4] = {0, n - 1, 8, 3};
        for (int i = 0; i < 4; ++i) {
            LOp d = lower_matrix(gate_matrix(obs[i], nullptr, false, 1), {n - 1 - ow[i]}, 0, 0, true);
            d.flags = (d.flags & ~TF_ZERO_UNCTRL) | TF_DOT; d.sa = 0; d.sb = 0; d.slot = nslots++;
            prog.push_back(d);
        }
    }
    for (int l = 0; l < (which == 3 ? 0 : 3); ++l) {
        for (int w = 0; w < n; ++w) gate(GateId::Rot, {(uint32_t)w}, {U(rng), U(rng), U(rng)});
        for (int s = 0; s < 2; ++s) for (int w = s; w < n - 1; w += 2) gate(GateId::CZ, {(uint32_t)w, (uint32_t)w + 1}, {});
        if (which >= 1 && which != 4 && which != 5) {
            gate(GateId::CRY, {1, 7}, {U(rng)});
            gate(GateId::IsingZZ, {0, (uint32_t)n - 1}, {U(rng)});
            gate(GateId::RX, {3}, {U(rng)}, {9}, {0});
            gate(GateId::CNOT, {(uint32_t)n - 2, 2}, {});
            gate(GateId::MultiRZ, {2, 5, 11}, {U(rng)});
            gate(GateId::PhaseShift, {4}, {U(rng)});
            gate(GateId::SWAP, {6, 12}, {});
        }
        if (which >= 4) {
            // generator dots of every wire's rotation against each observable state, as the
            // sweep of several observables makes them (passes fill more than 32 slots)
            for (int w = 0; w < n; ++w)
                for (int sb = 1; sb < ns; ++sb) dot(w % 2 ? GateId::RY : GateId::RZ, (uint32_t)w, 0, sb);
        }
        if (which == 2 || which >= 4 || l == 2) {
            for (int sb = std::min(1, ns - 1); sb < ns; ++sb) {
                dot(GateId::RX, (uint32_t)n - 1, 0, sb);
                dot(GateId::RZ, 0, 0, sb);
                dot(GateId::RY, 5, 0, sb);
            }
        }
    }
    PlanConfig cfg; cfg.n = n; cfg.ns = ns; cfg.kt_max = ns > 2 ? 11 : 12; cfg.low_bits = getenv("EMU_LOW_BITS") ? atoi(getenv("EMU_LOW_BITS")) : (getenv("EMU_NO_SPLIT") ? (dtype ? 4 : 5) : 4); cfg.reg_rb = ns > 2 ? 3 : 4; cfg.max_dots = ns > 2 ? std::min(kMaxJitPassDots, kMaxPassDots * (ns - 1)) : kMaxPassDots; cfg.scale_rotations = !getenv("EMU_NO_SCALE");
    cfg.max_scale_growth = dtype ? 1e300 : 1e20; // as the engine configures specialised passes
    // the engine's specialised-pass planning: lookahead tile bits, budget 52 (one state; 78 for two,
    // 26 (ns + 1) for more), 4 low bits
    cfg.lookahead = !(getenv("EMU_LOOKAHEAD") && atoi(getenv("EMU_LOOKAHEAD")) == 0);
    cfg.budget = getenv("EMU_NO_SPLIT") ? 24 : (ns == 1 ? 52 : ns == 2 ? 78 : 26.0 * (ns + 1));
    // EMU_BASIS=k (one state): the state is the basis state |k> not yet
    // written (a deferred reset); as in the engine, the program's leading
    // one-qubit gates are folded into its product factors and the first pass
    // forms the product state, ignoring memory
    const long long basis = (getenv("EMU_BASIS") && ns == 1) ? atoll(getenv("EMU_BASIS")) : -1;
    std::vector<cd> fac;
    std::vector<LOp> run_prog = prog;
    if (basis >= 0) {
        fac.assign(2 * static_cast<size_t>(n), cd(0.0, 0.0));
        for (int q = 0; q < n; ++q) fac[2 * q + ((basis >> q) & 1)] = 1.0;
        const size_t folded = absorb_into_product(run_prog, fac);
        printf("folded %zu leading one-qubit ops into the product state\n", folded);
    }
    // as the engine does for specialised passes: 2x2 gates as phase / real rotation / phase
    std::vector<LOp> split = run_prog;
    if (!getenv("EMU_NO_SPLIT")) split_1q_rotations(split);
    auto passes = plan_passes(split, cfg);
    // initial states: random normalised
    std::vector<St> S(ns, St(1ull << n));
    for (auto &s : S) for (auto &x : s) x = cd(U(rng) - 3.14, U(rng) - 3.14);
    if (basis >= 0) {
        for (auto &x : S[0]) x = 0.0;
        S[0][static_cast<size_t>(basis)] = 1.0;
    }
    std::vector<St> R = S;
    std::vector<cd> rdots(nslots), edots(nslots);
    for (const auto &o : prog) ref_apply(R, n, o, rdots);
    // emulated run
    const size_t esz = dtype ? 16 : 8;
    std::vector<std::vector<char>> buf(std::max(2, ns), std::vector<char>(esz << n));
    for (int s = 0; s < ns; ++s)
        for (uint64_t i = 0; i < (1ull << n); ++i) {
            // with a deferred reset the memory holds unrelated values
            const cd v = basis >= 0 ? cd(U(rng), -U(rng)) : S[s][i];
            if (dtype) { double *p = (double *)buf[s].data(); p[2 * i] = v.real(); p[2 * i + 1] = v.imag(); }
            else { float *p = (float *)buf[s].data(); p[2 * i] = (float)v.real(); p[2 * i + 1] = (float)v.imag(); }
        }
    for (size_t p = 0; p < passes.size(); ++p) {
        // EMU_LIFT=1: bit positions read from the table after the coefficients
        const bool lift = getenv("EMU_LIFT") && atoi(getenv("EMU_LIFT")) != 0;
        std::vector<uint64_t> geo;
        const std::string src = jit_pass_source(passes[p], n, dtype, ns, lift ? &geo : nullptr);
        if (src.empty()) { printf("pass %zu not specialisable\n", p); return 2; }
        const std::string base = work + "/emu_pass_" + std::to_string(which) + "_" + std::to_string(dtype) + "_" + std::to_string(p);
        { std::ofstream f(base + ".cpp"); f << transform(src); }
        const std::string cmd = "g++ -O1 -std=c++20 -fPIC -shared -w -I" + emu_dir + " " + base + ".cpp " + emu_dir +
                                "/emu_main.cpp -o " + base + ".so -lpthread";
        if (system(cmd.c_str()) != 0) { printf("compile failed for pass %zu\n", p); return 3; }
        void *h = dlopen((base + ".so").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) { printf("dlopen: %s\n", dlerror()); return 4; }
        Launch L = (Launch)dlsym(h, "emu_launch");
        // coefficients, then the deferred-preparation marker, the bit-position
        // table (two entries per slot) and the product factors, as the engine
        // writes them
        const size_t nc = passes[p].coef.size();
        const size_t gs = (geo.size() + 1) / 2;
        std::vector<double2> coef(nc + 1 + gs + 2 * static_cast<size_t>(n), double2{0.0, 0.0});
        if (gs) std::memcpy(&coef[nc + 1], geo.data(), geo.size() * sizeof(uint64_t));
        for (size_t i = 0; i < nc; ++i) coef[i] = {passes[p].coef[i].real(), passes[p].coef[i].imag()};
        const bool form = p == 0 && basis >= 0;
        if (form && !((passes[p].load_mask & 1u) && (passes[p].store_mask & 1u))) {
            printf("first pass cannot form the product state\n");
            return 5;
        }
        coef[nc] = form ? double2{1.0, 0.0} : double2{-1.0, 0.0};
        if (form)
            for (size_t i = 0; i < fac.size(); ++i) coef[nc + 1 + gs + i] = {fac[i].real(), fac[i].imag()};
        const uint64_t ntiles = 1ull << (n - (int)passes[p].tbits.size());
        std::vector<double2> part(std::max<size_t>(1, passes[p].slot_ids.size() * ntiles), {0, 0});
        const unsigned grid = (unsigned)ntiles;
        (void)grid_cap;
        auto t0 = std::chrono::steady_clock::now();
        L(buf[0].data(), ns > 1 ? buf[1].data() : nullptr, coef.data(), part.data(), ntiles, grid,
          ns > 2 ? buf[2].data() : nullptr, ns > 3 ? buf[3].data() : nullptr);
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (size_t sl = 0; sl < passes[p].slot_ids.size(); ++sl) {
            cd sum = 0;
            for (uint64_t t = 0; t < ntiles; ++t) sum += cd(part[sl * ntiles + t].x, part[sl * ntiles + t].y);
            edots[passes[p].slot_ids[sl]] = sum;
        }
        printf("pass %zu ops %zu stages %zu dots %zu ntiles %llu grid %u: %.2f s\n", p, passes[p].ops.size(), passes[p].stages.size(),
               passes[p].slot_ids.size(), (unsigned long long)ntiles, grid, dt);
        fflush(stdout);
        dlclose(h);
    }
    double err = 0, nrm = 0;
    for (int s = 0; s < ns; ++s)
        for (uint64_t i = 0; i < (1ull << n); ++i) {
            cd e = dtype ? cd(((double *)buf[s].data())[2 * i], ((double *)buf[s].data())[2 * i + 1])
                         : cd(((float *)buf[s].data())[2 * i], ((float *)buf[s].data())[2 * i + 1]);
            err = std::max(err, std::abs(e - R[s][i]));
            nrm = std::max(nrm, std::abs(R[s][i]));
        }
    double derr = 0;
    for (int i = 0; i < nslots; ++i) derr = std::max(derr, std::abs(edots[i] - rdots[i]) / std::max(1.0, std::abs(rdots[i])));
    printf("case %d dtype %d n %d: passes %zu max |state err| %.3e (max |amp| %.2f), max rel dot err %.3e over %d dots\n", which, dtype,
           n, passes.size(), err, nrm, derr, nslots);
    const double tol = dtype ? 1e-10 : 1e-3;
    return (err <= tol * std::max(1.0, nrm) && derr <= tol) ? 0 : 1;
}
