// Pass specialisation (see lq_jit.hpp).
#include "lq_jit.hpp"

#include <hip/hiprtc.h>
#include <sys/resource.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lq_common.hpp"

namespace lqamd {

// ---------------------------------------------------------------------------
// Code generation
// ---------------------------------------------------------------------------
namespace {

const char *kPrelude = R"(
typedef LQ_REAL real;
typedef LQ_CPLX cplx;
__device__ __forceinline__ cplx mk(real x, real y) { cplx r; r.x = x; r.y = y; return r; }
// coefficients are read through the constant address space: they do not
// change while a pass runs, so these are scalar loads whatever the kernel's size
typedef __attribute__((address_space(4))) const double2 kcoef;
__device__ __forceinline__ cplx ldc(const double2 *__restrict__ c, int i) {
    const double2 d = ((const kcoef *)c)[i];
    return mk((real)d.x, (real)d.y);
}
__device__ __forceinline__ cplx cm(const cplx a, const cplx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cplx cm2(const cplx a, const cplx b, const cplx c, const cplx d) {
    return mk(a.x * b.x - a.y * b.y + c.x * d.x - c.y * d.y, a.x * b.y + a.y * b.x + c.x * d.y + c.y * d.x);
}
__device__ __forceinline__ cplx add(const cplx a, const cplx b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zm(const double2 a, const double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cplx ng(const cplx a) { return mk(-a.x, -a.y); }
__device__ __forceinline__ void d1(cplx &a, cplx &b, const cplx m0, const cplx m1, const cplx m2, const cplx m3) {
    const cplx x = a, y = b;
    a = cm2(m0, x, m1, y);
    b = cm2(m2, x, m3, y);
}
__device__ __forceinline__ void d2(cplx &a, cplx &b, cplx &c, cplx &d, const cplx *m) {
    const cplx x0 = a, x1 = b, x2 = c, x3 = d;
    a = add(cm2(m[0], x0, m[1], x1), cm2(m[2], x2, m[3], x3));
    b = add(cm2(m[4], x0, m[5], x1), cm2(m[6], x2, m[7], x3));
    c = add(cm2(m[8], x0, m[9], x1), cm2(m[10], x2, m[11], x3));
    d = add(cm2(m[12], x0, m[13], x1), cm2(m[14], x2, m[15], x3));
}
__device__ __forceinline__ void acc(double &r, double &i, const cplx b, const cplx v) {
    const double bx = b.x, by = b.y, vx = v.x, vy = v.y;
    r += bx * vx + by * vy;
    i += bx * vy - by * vx;
}
// [[1, -t], [t, 1]] on (a, b): two FMAs per component
__device__ __forceinline__ void d1t(cplx &a, cplx &b, const real t) {
  const real ax = a.x, ay = a.y;
  a.x = fma(-t, b.x, ax); a.y = fma(-t, b.y, ay);
  b.x = fma(t, ax, b.x); b.y = fma(t, ay, b.y);
}
// 2x2 with real coefficients: half the multiply-adds of the complex form
__device__ __forceinline__ void d1r(cplx &a, cplx &b, const real m0, const real m1, const real m2, const real m3) {
    const cplx x = a, y = b;
    a = mk(m0 * x.x + m1 * y.x, m0 * x.y + m1 * y.y);
    b = mk(m2 * x.x + m3 * y.x, m2 * x.y + m3 * y.y);
}
__device__ __forceinline__ real ldr(const double2 *__restrict__ c, int i) { return (real)((const kcoef *)c)[i].x; }
// predicated forms for ops whose controls depend on the thread: computed by
// every lane and selected, so the wave never diverges
__device__ __forceinline__ cplx sel(const bool g, const cplx a, const cplx b) { return mk(g ? a.x : b.x, g ? a.y : b.y); }
__device__ __forceinline__ void d1p(cplx &a, cplx &b, const cplx m0, const cplx m1, const cplx m2, const cplx m3,
                                    const bool g) {
    cplx x = a, y = b;
    d1(x, y, m0, m1, m2, m3);
    a = sel(g, x, a);
    b = sel(g, y, b);
}
__device__ __forceinline__ void d2p(cplx &a, cplx &b, cplx &c, cplx &d, const cplx *m, const bool g) {
    cplx w = a, x = b, y = c, z = d;
    d2(w, x, y, z, m);
    a = sel(g, w, a);
    b = sel(g, x, b);
    c = sel(g, y, c);
    d = sel(g, z, d);
}
__device__ __forceinline__ void accp(double &r, double &i, const cplx b, const cplx v, const bool g) {
    double tr = 0.0, ti = 0.0;
    acc(tr, ti, b, v);
    r += g ? tr : 0.0;
    i += g ? ti : 0.0;
}
__device__ __forceinline__ unsigned pc(unsigned long long x) { return (unsigned)__popcll(x); }
// the amplitude of the lane whose index differs in the bits of m
__device__ __forceinline__ cplx shx(const cplx a, const int m) {
    return mk((real)__shfl_xor(a.x, m, 64), (real)__shfl_xor(a.y, m, 64));
}
// Exchange between the lanes 2^bit apart (bit 4 or 5): afterwards the lanes
// with that lane bit clear hold in b their partner's a, the others hold in a
// their partner's b (a register bit swapped with a lane bit). One
// v_permlane16_swap / v_permlane32_swap per 32-bit word: it swaps the odd
// 16-lane rows (upper 32 lanes) of its first operand with the even rows
// (lower 32 lanes) of its second.
#ifdef LQ_EMU
__device__ __forceinline__ void pswap(cplx &a, cplx &b, const int bit) {
    const bool hb = ((threadIdx.x >> bit) & 1u) != 0u;
    const cplx x = shx(hb ? a : b, 1 << bit);
    if (hb) a = x;
    else b = x;
}
#else
__device__ __forceinline__ void pswap(cplx &a, cplx &b, const int bit) {
    unsigned wa[sizeof(cplx) / 4], wb[sizeof(cplx) / 4];
    __builtin_memcpy(wa, &a, sizeof(cplx));
    __builtin_memcpy(wb, &b, sizeof(cplx));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(cplx) / 4); ++i) {
        if (bit == 5) {
            const auto r = __builtin_amdgcn_permlane32_swap(wa[i], wb[i], false, false);
            wa[i] = r[0];
            wb[i] = r[1];
        } else {
            const auto r = __builtin_amdgcn_permlane16_swap(wa[i], wb[i], false, false);
            wa[i] = r[0];
            wb[i] = r[1];
        }
    }
    __builtin_memcpy(&a, wa, sizeof(cplx));
    __builtin_memcpy(&b, wb, sizeof(cplx));
}
#endif
// a zero the compiler cannot see through: coefficient offsets depend on it,
// so coefficient loads stay next to their use in each tile instead of being
// hoisted out of the tile loop (and kept in registers across it)
__device__ __forceinline__ unsigned opaque0() {
    unsigned z = 0;
    asm volatile("" : "+s"(z));
    return z;
}
typedef real nvec __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cplx ldnt(const cplx *p) {
    const nvec t = __builtin_nontemporal_load(reinterpret_cast<const nvec *>(p));
    return mk(t.x, t.y);
}
__device__ __forceinline__ void stnt(cplx *p, const cplx v) {
    nvec t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<nvec *>(p));
}
__device__ __forceinline__ unsigned swz(unsigned l) { return LQ_SWZ; }
template <int D> __device__ __forceinline__ void dflush(double r, double i, double2 (*dred)[D], unsigned slot) {
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_xor(r, off, 64);
        i += __shfl_xor(i, off, 64);
    }
    if ((threadIdx.x & 63u) == 0) dred[threadIdx.x >> 6][slot] = make_double2(r, i);
}
)";

// row length of a pass's per-wave dot buffer: the slots rounded up to 32
int dred_len(int nslots) { return std::max(32, (nslots + 31) / 32 * 32); }

std::string hex64(uint64_t v) {
    std::ostringstream s;
    s << "0x" << std::hex << v << "ull";
    return s.str();
}

// Expression depositing bits of `src` (bit j -> position pos[j], j < cnt)
std::string deposit(const char *src, const int *pos, int cnt, bool wide) {
    std::ostringstream e;
    bool any = false;
    for (int j = 0; j < cnt;) {
        int len = 1;
        while (j + len < cnt && pos[j + len] == pos[j] + len) ++len;
        if (any) e << " | ";
        const uint64_t m = (len >= 64) ? ~0ull : ((1ull << len) - 1);
        if (wide)
            e << "((((unsigned long long)" << src << " >> " << j << ") & " << hex64(m) << ") << " << pos[j] << ")";
        else
            e << "(((" << src << " >> " << j << ") & " << m << "u) << " << pos[j] << ")";
        any = true;
        j += len;
    }
    if (!any) return wide ? "0ull" : "0u";
    return e.str();
}

struct Mapping {
    int r[8];
    int h[8];
};

class Gen {
  public:
    Gen(const PassPlan &p, int n, int dtype, int ns, bool lift)
        : p_(p), dtype_(dtype), ns_(ns), nt_(jit_config().nontemporal), lift_(lift) {
        kt_ = static_cast<int>(p.tbits.size());
        rb_ = kt_ - 8;
        E_ = 1 << rb_;
        geom_ = make_geom(n, p.tbits);
        tsize_ = 1u << kt_;
        const size_t per_state = static_cast<size_t>(dtype ? 16 : 8) << kt_;
        lds32_ = jit_config().lds32;
        split_ = lds32_ && per_state > 32768;
        LQ_ABORT_IF(per_state > 65536, "kernel generator: tile too large");
    }

    std::vector<uint64_t> &geo() { return geo_; }

    std::string run() {
        std::ostringstream &o = o_;
        o << "#define LQ_GEO " << p_.coef.size() + 1 << "\n";
        o << "#define LQ_REAL " << (dtype_ ? "double" : "float") << "\n";
        o << "#define LQ_CPLX " << (dtype_ ? "double2" : "float2") << "\n";
        o << "#define LQ_SWZ " << (dtype_ ? "(l ^ (((l >> 3) ^ (l >> 6) ^ (l >> 9)) & 7u))" : "(l ^ (((l >> 4) ^ (l >> 8)) & 15u))")
          << "\n";
        o << kPrelude;
        if (lift_) {
            // positions of the pass's bits in the state index are read at run
            // time from the table after the coefficients, so passes of the
            // same shape share one kernel whatever the qubit count
            o << "typedef __attribute__((address_space(4))) const unsigned long long kgeo;\n";
            o << "#define GQ(i) (((const kgeo *)(coef + LQ_GEO))[i])\n";
        }
        o << "// lds_bytes " << jit_lds_bytes(kt_, dtype_, ns_) << "\n";
        o << "extern \"C\" __global__ __launch_bounds__(256) ";
        if (const int w = jit_waves_per_eu(dtype_, ns_); w > 0) o << "__attribute__((amdgpu_waves_per_eu(" << w << "))) ";
        o << "void lqpass(cplx *__restrict__ sv0, cplx *__restrict__ sv1,"
             " const double2 *__restrict__ coef, double2 *__restrict__ partials, unsigned long long ntiles,"
             " cplx *__restrict__ sv2, cplx *__restrict__ sv3) {\n";
        o << "  extern __shared__ cplx lds[];\n";
        const int nslots = static_cast<int>(p_.slot_ids.size());
        if (nslots) o << "  __shared__ double2 dred[4][" << dred_len(nslots) << "];\n";
        o << "  const unsigned tid = threadIdx.x;\n";
        Mapping m0{};
        for (int i = 0; i < rb_; ++i) m0.r[i] = 8 + i;
        for (int j = 0; j < 8; ++j) m0.h[j] = j;
        for (int s = 0; s < ns_; ++s) o << "  cplx v" << s << "[" << E_ << "];\n";
        // workgroups are dealt round-robin to the 8 XCDs: let XCD x walk
        // the x-th eighth of the tiles in order, so neighbouring segments
        // are streamed by one XCD at a time (ntiles is a power of two)
        o << "  const unsigned long long bid = (unsigned long long)blockIdx.x + (unsigned long long)blockIdx.y * gridDim.x;\n";
        o << "  if (bid >= ntiles) return;\n";
        o << "  const unsigned long long tile = ntiles >= 8ull ? (bid & 7ull) * (ntiles >> 3) + (bid >> 3) : bid;\n";
        // load straight into the first stage's mapping when its thread
        // bits start with the contiguous low bits (the loads coalesce as
        // in the default mapping): no entry transpose
        Mapping lm = m0;
        const TileStage &s0 = p_.stages.front();
        if (s0.kind == STAGE_REG && s0.entry == ENTRY_TRANSPOSE) {
            Mapping m{};
            for (int i = 0; i < rb_; ++i) m.r[i] = s0.rbits[i];
            for (int j = 0; j < 8; ++j) m.h[j] = s0.hbits[j];
            if (coalesces(m)) lm = m;
        }
        o << "  {\n";
        emit_base_gl("tile", "gl0", lm);
        if (ns_ == 1) {
            // the entry after the pass's coefficients is (1, 0) when the state
            // is a product state not yet written (a deferred reset, with the
            // leading one-qubit gates folded in); the pass then forms its
            // input from the factors that follow instead of reading it
            o << "  const double2 bmark = ((const kcoef *)coef)[" << p_.coef.size() << "];\n";
            o << "  if (bmark.x > 0.0) {\n";
            emit_product(lm);
            o << "  } else {\n";
            emit_loads("v", "gl0", lm);
            o << "  }\n";
        } else {
            emit_loads("v", "gl0", lm);
        }
        o << "  }\n";
        emit_tile_body(m0, nslots, lm);
        o << "}\n";
        // the product factors follow the position table
        return "#define LQ_FOFF " + std::to_string(p_.coef.size() + 1 + (geo_.size() + 1) / 2) + "\n" + o.str();
    }

  private:
    const PassPlan &p_;
    int dtype_, ns_;
    bool nt_;
    bool lift_;                // positions in the index read at run time (GQ table)
    std::vector<uint64_t> geo_; // that table, in order of use

    // a 64-bit constant that depends on where the pass's bits sit in the index
    std::string K(uint64_t v) {
        if (!lift_) return hex64(v);
        geo_.push_back(v);
        return "GQ(" + std::to_string(geo_.size() - 1) + ")";
    }
    // a bit position of the index (a shift amount)
    std::string Sh(int v) {
        if (!lift_) return std::to_string(v);
        geo_.push_back(static_cast<uint64_t>(v));
        return "(unsigned)GQ(" + std::to_string(geo_.size() - 1) + ")";
    }
    // the tile's base index: the tile number with zeros at the tile bits
    std::string base_expr(const char *tile) {
        if (!lift_) return deposit(tile, geom_.ob, geom_.nout, true);
        std::ostringstream e;
        e << "[&]() { unsigned long long x = (unsigned long long)" << tile << " << " << geom_.c << ";";
        for (int j = geom_.c; j < kt_; ++j) {
            e << " { const unsigned s = " << Sh(geom_.tb[j]) << "; x = ((x >> s) << (s + 1u)) | (x & ((1ull << s) - 1ull)); }";
        }
        e << " return x; }()";
        return e.str();
    }
    // bits of tid (bit j -> index position pos[j])
    std::string tid_expr(const int *pos) {
        if (!lift_) return deposit("tid", pos, 8, true);
        std::ostringstream e;
        bool any = false;
        for (int j = 0; j < 8;) {
            int len = 1;
            while (j + len < 8 && pos[j + len] == pos[j] + len) ++len;
            if (any) e << " | ";
            e << "((((unsigned long long)tid >> " << j << ") & " << ((1u << len) - 1u) << "ull) << " << Sh(pos[j]) << ")";
            any = true;
            j += len;
        }
        return e.str();
    }
    std::string fac(const std::string &i) { return "((const kcoef *)coef)[LQ_FOFF + " + i + "]"; }

    void emit_base_gl(const char *tile, const char *gl, const Mapping &m) {
        o_ << "  const unsigned long long " << (std::string(gl) == "gl0" ? "b0" : "bn") << " = "
           << base_expr(tile) << ";\n";
        int pos[8];
        for (int j = 0; j < 8; ++j) pos[j] = geom_.tb[m.h[j]];
        o_ << "  const unsigned long long " << gl << "_ = " << (std::string(gl) == "gl0" ? "b0" : "bn") << " | "
           << tid_expr(pos) << ";\n";
    }
    // v0 = the product state prod_q f_q[bit q] over the tile, f_q[b] at
    // factor 2q + b: the factors of the bits fixed per thread (outside and
    // thread bits) multiply into one per-thread value, the register bits'
    // factors into a 2^rb table (FP64 in both precisions)
    void emit_product(const Mapping &m) {
        std::ostringstream &o = o_;
        int rg[kMaxRegBits];
        uint64_t regs = 0;
        for (int i = 0; i < rb_; ++i) {
            rg[i] = geom_.tb[m.r[i]];
            regs |= 1ull << rg[i];
        }
        o << "    double2 pw[" << E_ << "];\n";
        o << "    pw[0] = make_double2(1.0, 0.0);\n";
        if (lift_) {
            // the bits other than the register bits: a count, then up to 40
            // positions (a fixed block, so the table's layout does not
            // depend on the qubit count)
            const size_t at = geo_.size();
            geo_.push_back(0);
            for (int q = 0; q < geom_.n; ++q) {
                if ((regs >> q) & 1ull) continue;
                geo_.push_back(static_cast<uint64_t>(q));
            }
            geo_[at] = geo_.size() - at - 1;
            geo_.resize(at + 41, 0);
            o << "    for (unsigned i = 0; i < (unsigned)GQ(" << at << "); ++i) {\n";
            o << "      const unsigned q = (unsigned)GQ(" << at + 1 << " + i);\n";
            o << "      pw[0] = zm(pw[0], ((gl0_ >> q) & 1ull) ? " << fac("2u * q + 1u") << " : " << fac("2u * q") << ");\n";
            o << "    }\n";
        } else {
            for (int q = 0; q < geom_.n; ++q) {
                if ((regs >> q) & 1ull) continue;
                o << "    pw[0] = zm(pw[0], ((gl0_ >> " << q << ") & 1ull) ? " << fac(std::to_string(2 * q + 1)) << " : "
                  << fac(std::to_string(2 * q)) << ");\n";
            }
        }
        for (int i = 0; i < rb_; ++i) {
            const std::string q = Sh(rg[i]);
            o << "    { const double2 f0 = " << fac("2u * " + q) << ", f1 = " << fac("2u * " + q + " + 1u") << ";\n";
            for (int e = (1 << i) - 1; e >= 0; --e) {
                o << "      pw[" << (e | (1 << i)) << "] = zm(pw[" << e << "], f1);\n";
                o << "      pw[" << e << "] = zm(pw[" << e << "], f0);\n";
            }
            o << "    }\n";
        }
        for (int e = 0; e < E_; ++e) o << "    v0[" << e << "] = mk((real)pw[" << e << "].x, (real)pw[" << e << "].y);\n";
    }
    void emit_loads(const char *dst, const char *gl, const Mapping &m) {
        for (int s = 0; s < ns_; ++s) {
            if (!((p_.load_mask >> s) & 1u)) continue;
            for (int e = 0; e < E_; ++e) {
                const std::string addr = std::string("&sv") + std::to_string(s) + "[" + gl + "_ | " + K(elem_global(m, e)) + "]";
                if (nt_)
                    o_ << "  " << dst << s << "[" << e << "] = ldnt(" << addr << ");\n";
                else
                    o_ << "  " << dst << s << "[" << e << "] = *(" << addr << ");\n";
            }
        }
    }
    // thread bits of m begin with the tile's 4 contiguous low bits (the 16
    // amplitudes of a run are held by consecutive lanes)
    //
    // Taking a stage's mapping for the loads or stores saves an LDS transpose
    // but can spread each memory instruction over farther segments; measured
    // on the 30-qubit passes it pays only in passes that would otherwise make
    // four or more transposes (memory-bound lighter passes got slower).
    bool coalesces(const Mapping &m) const {
        const int L = 4; // 16 lanes: 256 bytes of complex128, one 128-byte line of complex64
        if (geom_.c < L) return false;
        for (int j = 0; j < L; ++j)
            if (m.h[j] != j) return false;
        const int mode = jit_config().direct_io;
        if (mode != 1) return mode == 2;
        int nt = p_.end_transpose ? 1 : 0;
        for (const auto &st : p_.stages) nt += st.entry == ENTRY_TRANSPOSE;
        return nt >= 4;
    }
    static bool same_mapping(const Mapping &a, const Mapping &b) {
        for (int j = 0; j < 8; ++j)
            if (a.h[j] != b.h[j]) return false;
        for (int i = 0; i < kMaxRegBits; ++i)
            if (a.r[i] != b.r[i]) return false;
        return true;
    }
    bool same_regs(const Mapping &a, const Mapping &b) const {
        for (int i = 0; i < rb_; ++i)
            if (a.r[i] != b.r[i]) return false;
        return true;
    }
    // the 16 lanes of a run hold the tile's four contiguous low bits in order
    bool lanes_coalesce(const Mapping &m) const {
        if (geom_.c < 4) return false;
        for (int j = 0; j < 4; ++j)
            if (m.h[j] != j) return false;
        return true;
    }

    // A change of register bits made by swapping register bits with lane
    // bits (thread bits 0..5) through wave shuffles: no LDS and no barrier.
    // It applies when every bit entering the registers is a lane bit; the
    // bit leaving register i takes the lane position of the bit entering, so
    // the thread order of the result differs from the target's (only the
    // register order must match: ops address registers by index).
    struct LaneSwaps {
        int n = 0;
        int reg[kMaxRegBits], lane[kMaxRegBits];
        Mapping result{};
    };
    bool plan_lane_swaps(const Mapping &to, LaneSwaps *ls) const {
        const int mode = jit_config().lane_swaps;
        if (mode == 0) return false;
        int outs[kMaxRegBits], nout = 0, lanes[kMaxRegBits], nin = 0;
        for (int i = 0; i < rb_; ++i) {
            bool stays = false;
            for (int k = 0; k < rb_; ++k) stays = stays || cur_.r[i] == to.r[k];
            if (!stays) outs[nout++] = i;
        }
        for (int k = 0; k < rb_; ++k) {
            bool held = false;
            for (int i = 0; i < rb_; ++i) held = held || cur_.r[i] == to.r[k];
            if (held) continue;
            int j = -1;
            for (int q = 0; q < 8; ++q)
                if (cur_.h[q] == to.r[k]) j = q;
            if (j < 0 || j >= 6) return false;
            if (mode == 1 && j < 4) return false; // only the row swaps (lane bits 4, 5)
            lanes[nin++] = j;
        }
        if (nin != nout) return false;
        Mapping m = cur_;
        for (int p = 0; p < nin; ++p) {
            ls->reg[p] = outs[p];
            ls->lane[p] = lanes[p];
            std::swap(m.r[outs[p]], m.h[lanes[p]]);
        }
        ls->n = nin;
        for (int k = 0; k < rb_; ++k) m.r[k] = to.r[k];
        ls->result = m;
        return true;
    }
    void emit_lane_swaps(const Mapping &to, const LaneSwaps &ls) {
        std::ostringstream &o = o_;
        Mapping m = cur_;
        for (int p = 0; p < ls.n; ++p) {
            const int i = ls.reg[p], j = ls.lane[p];
            o << "  { // register bit " << i << " <-> lane bit " << j << "\n";
            if (j < 4) o << "    const bool hb = ((tid >> " << j << ") & 1u) != 0u;\n";
            for (int s = 0; s < ns_; ++s) {
                if (!((p_.load_mask >> s) & 1u)) continue;
                for (int e = 0; e < E_; ++e) {
                    if ((e >> i) & 1) continue;
                    const std::string a = "v" + std::to_string(s) + "[" + std::to_string(e) + "]";
                    const std::string b = "v" + std::to_string(s) + "[" + std::to_string(e | (1 << i)) + "]";
                    if (j >= 4) {
                        o << "    pswap(" << a << ", " << b << ", " << j << ");\n";
                        continue;
                    }
                    o << "    { const cplx x = shx(sel(hb, " << a << ", " << b << "), " << (1 << j) << ");\n";
                    o << "      " << a << " = sel(hb, x, " << a << "); " << b << " = sel(hb, " << b << ", x); }\n";
                }
            }
            o << "  }\n";
            std::swap(m.r[i], m.h[j]);
        }
        // the target's register order: a renaming of the register arrays
        int from[kMaxRegBits];
        bool same = true;
        for (int k = 0; k < rb_; ++k) {
            from[k] = -1;
            for (int i = 0; i < rb_; ++i)
                if (m.r[i] == to.r[k]) from[k] = i;
            LQ_ABORT_IF(from[k] < 0, "kernel generator: register bits lost in a lane swap");
            same = same && from[k] == k;
        }
        if (!same) {
            for (int s = 0; s < ns_; ++s) {
                if (!((p_.load_mask >> s) & 1u)) continue;
                o << "  {\n    cplx w[" << E_ << "];\n";
                for (int e = 0; e < E_; ++e) {
                    int em = 0;
                    for (int k = 0; k < rb_; ++k)
                        if ((e >> k) & 1) em |= 1 << from[k];
                    o << "    w[" << e << "] = v" << s << "[" << em << "];\n";
                }
                for (int e = 0; e < E_; ++e) o << "    v" << s << "[" << e << "] = w[" << e << "];\n";
                o << "  }\n";
            }
        }
        cur_ = ls.result;
    }
    // stages, store and dot partials of the tile `tile` held in v* (loaded in mapping lm)
    void emit_tile_body(const Mapping &m0, int nslots, const Mapping &lm) {
        std::ostringstream &o = o_;
        o << "  {\n";
        o << "  const unsigned long long base = " << base_expr("tile") << ";\n";
        o << "  const unsigned cz = opaque0();\n";
        if (nslots) {
            const int d = dred_len(nslots);
            o << "  for (unsigned i = tid; i < " << 4 * d << "u; i += 256u) dred[i / " << d << "u][i % " << d
              << "u] = make_double2(0.0, 0.0);\n";
            o << "  __syncthreads();\n";
        }
        cur_ = lm;
        emit_gl("gl0", m0);
        for (size_t si = 0; si < p_.stages.size(); ++si) {
            const TileStage &st = p_.stages[si];
            Mapping m{};
            for (int i = 0; i < rb_; ++i) m.r[i] = st.rbits[i];
            for (int j = 0; j < 8; ++j) m.h[j] = st.hbits[j];
            if (st.entry == ENTRY_TRANSPOSE && !same_regs(cur_, m)) {
                LaneSwaps ls;
                if (plan_lane_swaps(m, &ls)) {
                    emit_lane_swaps(m, ls);
                } else {
                    transpose(m);
                    cur_ = m;
                }
            }
            o << "  {\n";
            emit_gl("gt", cur_);
            for (uint32_t oi = st.op_begin; oi < st.op_end; ++oi) emit_op(p_.ops[oi], static_cast<int>(oi));
            o << "  }\n";
        }
        if (p_.store_mask) {
            // store straight from the last stage's mapping when it coalesces
            const char *gs = "gl0";
            LaneSwaps ls;
            if (p_.end_transpose && coalesces(cur_)) {
                emit_gl("gls", cur_);
                gs = "gls";
            } else if (!same_regs(cur_, m0) && plan_lane_swaps(m0, &ls) && lanes_coalesce(ls.result)) {
                // to the store's register bits by lane swaps; the lanes keep
                // the low bits, so the stores coalesce in that mapping
                emit_lane_swaps(m0, ls);
                emit_gl("gls", cur_);
                gs = "gls";
            } else if (same_regs(cur_, m0) && !same_mapping(cur_, m0) && lanes_coalesce(cur_)) {
                emit_gl("gls", cur_);
                gs = "gls";
            } else {
                if (!same_mapping(cur_, m0)) transpose(m0);
                cur_ = m0;
            }
            for (int s = 0; s < ns_; ++s) {
                if (!((p_.store_mask >> s) & 1u)) continue;
                for (int e = 0; e < E_; ++e) {
                    if (nt_)
                        o << "  stnt(&sv" << s << "[" << gs << " | " << K(elem_global(cur_, e)) << "], v" << s << "[" << e << "]);\n";
                    else
                        o << "  sv" << s << "[" << gs << " | " << K(elem_global(cur_, e)) << "] = v" << s << "[" << e << "];\n";
                }
            }
        }
        if (nslots) {
            o << "  __syncthreads();\n";
            o << "  if (tid < " << nslots << "u) {\n";
            o << "    double2 t = dred[0][tid];\n";
            o << "    for (int w = 1; w < 4; ++w) { t.x += dred[w][tid].x; t.y += dred[w][tid].y; }\n";
            o << "    partials[(unsigned long long)tid * ntiles + tile] = t;\n";
            o << "  }\n";
        }
        o << "  }\n";
    }
    int kt_, rb_, E_;
    uint32_t tsize_;
    TileGeom geom_;
    Mapping cur_{};
    bool lds_used_ = false;
    bool lds32_ = false; // states pass through LDS one at a time
    bool split_ = false; // and each in two halves (complex128 tiles of 2^12)
    std::ostringstream o_;

    uint64_t elem_global(const Mapping &m, int e) const {
        uint64_t g = 0;
        for (int i = 0; i < rb_; ++i)
            if ((e >> i) & 1) g |= 1ull << geom_.tb[m.r[i]];
        return g;
    }
    uint32_t elem_local(const Mapping &m, int e) const {
        uint32_t l = 0;
        for (int i = 0; i < rb_; ++i)
            if ((e >> i) & 1) l |= 1u << m.r[i];
        return l;
    }
    uint64_t regmask() const {
        uint64_t g = 0;
        for (int i = 0; i < rb_; ++i) g |= 1ull << geom_.tb[cur_.r[i]];
        return g;
    }
    int reg_index_of_global(int gb) const {
        for (int i = 0; i < rb_; ++i)
            if (geom_.tb[cur_.r[i]] == gb) return i;
        return -1;
    }
    void emit_gl(const char *name, const Mapping &m) {
        int pos[8];
        for (int j = 0; j < 8; ++j) pos[j] = geom_.tb[m.h[j]];
        o_ << "  const unsigned long long " << name << " = base | " << tid_expr(pos) << ";\n";
    }
    void transpose(const Mapping &to) {
        std::ostringstream &o = o_;
        if (!lds32_) {
            if (lds_used_) o << "  __syncthreads();\n";
            o << "  {\n";
            o << "    const unsigned lt = " << deposit("tid", cur_.h, 8, false) << ";\n";
            for (int s = 0; s < ns_; ++s) {
                if (!((p_.load_mask >> s) & 1u)) continue;
                for (int e = 0; e < E_; ++e)
                    o << "    lds[" << s * tsize_ << "u + swz(lt | " << elem_local(cur_, e) << "u)] = v" << s << "["
                      << e << "];\n";
            }
            o << "  }\n  __syncthreads();\n  {\n";
            o << "    const unsigned lt = " << deposit("tid", to.h, 8, false) << ";\n";
            for (int s = 0; s < ns_; ++s) {
                if (!((p_.load_mask >> s) & 1u)) continue;
                for (int e = 0; e < E_; ++e)
                    o << "    v" << s << "[" << e << "] = lds[" << s * tsize_ << "u + swz(lt | " << elem_local(to, e)
                      << "u)];\n";
            }
            o << "  }\n";
            lds_used_ = true;
            return;
        }
        // 32 KiB of LDS: one state at a time; complex128 tiles in two halves
        // by a tile bit b. With b a register bit on both sides, every thread
        // writes and reads half of its amplitudes per round; the values read
        // go to fresh variables (the old ones of the other half are still to
        // be written). With b a thread bit at the same thread-index position
        // on both sides, a thread writes and reads all its amplitudes in one
        // round, in place. Otherwise the transpose goes through an
        // intermediate mapping that shares register bits with both sides.
        if (!split_) {
            for (int s = 0; s < ns_; ++s) {
                if (!((p_.load_mask >> s) & 1u)) continue;
                if (lds_used_) o << "  __syncthreads();\n";
                lds_used_ = true;
                half_phase(cur_, s, -1, 0, true, "v");
                o << "  __syncthreads();\n";
                half_phase(to, s, -1, 0, false, "v");
            }
            return;
        }
        int b = -1;
        for (int i = 0; i < rb_ && b < 0; ++i)
            for (int j = 0; j < rb_; ++j)
                if (cur_.r[i] == to.r[j]) b = cur_.r[i];
        if (b < 0) {
            for (int j = 0; j < 8; ++j)
                if (cur_.h[j] == to.h[j]) {
                    split_transpose(to, cur_.h[j], false);
                    return;
                }
            Mapping mid{};
            mid.r[0] = cur_.r[0];
            for (int i = 1; i < rb_; ++i) mid.r[i] = to.r[i - 1];
            int h = 0;
            for (int tb = 0; tb < kt_; ++tb) {
                bool reg = false;
                for (int i = 0; i < rb_; ++i) reg |= mid.r[i] == tb;
                if (!reg) mid.h[h++] = tb;
            }
            split_transpose(mid, cur_.r[0], true);
            cur_ = mid;
            for (int i = 0; i < rb_ && b < 0; ++i)
                for (int j = 0; j < rb_; ++j)
                    if (cur_.r[i] == to.r[j]) b = cur_.r[i];
        }
        split_transpose(to, b, true);
    }
    void split_transpose(const Mapping &to, int b, bool fresh) {
        std::ostringstream &o = o_;
        for (int s = 0; s < ns_; ++s) {
            if (!((p_.load_mask >> s) & 1u)) continue;
            if (fresh) o << "  {\n  cplx w" << s << "[" << E_ << "];\n";
            for (int r = 0; r < 2; ++r) {
                if (lds_used_) o << "  __syncthreads();\n";
                lds_used_ = true;
                half_phase(cur_, s, b, r, true, "v");
                o << "  __syncthreads();\n";
                half_phase(to, s, b, r, false, fresh ? "w" : "v");
            }
            if (fresh) {
                for (int e = 0; e < E_; ++e) o << "  v" << s << "[" << e << "] = w" << s << "[" << e << "];\n";
                o << "  }\n";
            }
        }
    }
    // tile-local index without bit b (b < 0: unchanged)
    static uint32_t squeeze(uint32_t l, int b) {
        if (b < 0) return l;
        return ((l >> (b + 1)) << b) | (l & ((1u << b) - 1u));
    }
    // write (or read into `arr`) the amplitudes of state s whose tile bit b
    // equals r (b < 0: all of them)
    void half_phase(const Mapping &m, int s, int b, int r, bool write, const char *arr) {
        std::ostringstream &o = o_;
        int rb_pos = -1; // register index holding bit b
        for (int i = 0; i < rb_; ++i)
            if (m.r[i] == b) rb_pos = i;
        o << "  {\n";
        o << "    const unsigned lt = " << deposit("tid", m.h, 8, false) << ";\n";
        if (b < 0) {
            o << "    const unsigned lq = lt;\n";
        } else {
            o << "    const unsigned lq = ((lt >> " << b + 1 << ") << " << b << ") | (lt & " << ((1u << b) - 1u) << "u);\n";
        }
        const bool thread_cond = b >= 0 && rb_pos < 0;
        if (thread_cond) o << "    if (((lt >> " << b << ") & 1u) == " << r << "u) {\n";
        for (int e = 0; e < E_; ++e) {
            if (b >= 0 && rb_pos >= 0 && ((e >> rb_pos) & 1) != r) continue;
            const uint32_t el = squeeze(elem_local(m, e), b);
            if (write)
                o << "    lds[swz(lq | " << el << "u)] = v" << s << "[" << e << "];\n";
            else
                o << "    " << arr << s << "[" << e << "] = lds[swz(lq | " << el << "u)];\n";
        }
        if (thread_cond) o << "    }\n";
        o << "  }\n";
    }

    // split a (mask, value) condition on global bits into the register part
    // (checked per element at generation time) and the thread part
    bool elem_ok(uint64_t cm, uint64_t cv, int e) const {
        for (int i = 0; i < rb_; ++i) {
            const int gb = geom_.tb[cur_.r[i]];
            if ((cm >> gb) & 1ull)
                if (static_cast<uint64_t>((e >> i) & 1) != ((cv >> gb) & 1ull)) return false;
        }
        return true;
    }
    // returns the guard expression for the thread part ("" when none)
    std::string thread_guard(uint64_t cm, uint64_t cv) {
        const uint64_t tm = cm & ~regmask();
        if (!tm) return "";
        const std::string a = K(tm);
        return "((gt & " + a + ") == " + K(cv & tm) + ")";
    }

    std::vector<int> states_of(uint32_t smask) const {
        std::vector<int> v;
        for (int s = 0; s < ns_; ++s)
            if ((smask >> s) & 1u) v.push_back(s);
        return v;
    }

    void emit_op(const TileOp &t, int id) {
        std::ostringstream &o = o_;
        const bool dot = (t.flags & TF_DOT) != 0;
        const std::string guard = thread_guard(t.gcmask, t.gcval);
        const auto sts = states_of(t.smask);
        const std::string A = "v" + std::to_string(t.sa), B = "v" + std::to_string(t.sb);
        o << "  { // op " << id << " kind " << t.kind << "\n";
        if (dot) o << "    double ar = 0.0, ai = 0.0;\n";
        const bool pred = !guard.empty();
        if (pred) o << "    const bool gd = " << guard << ";\n";
        const std::string accf = pred ? "accp(" : "acc(";
        const std::string gsuf = pred ? ", gd)" : ")";
        const uint64_t off = t.coef_off;
        switch (t.kind) {
        case TOP_DENSE1: {
            const int I = t.lpos[0];
            if ((t.flags & TF_SCALED) && !dot && !pred) {
                // [[1, -t], [t, 1]]: the rotation's cosine is in a phase table of the pass
                o << "    const real tt = ldr(coef, cz + " << off + 2 << ");\n";
                for (int e = 0; e < E_; ++e) {
                    if ((e >> I) & 1) continue;
                    const int e1 = e | (1 << I);
                    for (int s : sts) o << "    d1t(v" << s << "[" << e << "], v" << s << "[" << e1 << "], tt);\n";
                }
                break;
            }
            if ((t.flags & TF_REAL) && !dot && !pred) {
                o << "    const real r0 = ldr(coef, cz + " << off << "), r1 = ldr(coef, cz + " << off + 1
                  << "), r2 = ldr(coef, cz + " << off + 2 << "), r3 = ldr(coef, cz + " << off + 3 << ");\n";
                for (int e = 0; e < E_; ++e) {
                    if ((e >> I) & 1) continue;
                    const int e1 = e | (1 << I);
                    for (int s : sts) o << "    d1r(v" << s << "[" << e << "], v" << s << "[" << e1 << "], r0, r1, r2, r3);\n";
                }
                break;
            }
            o << "    const cplx m0 = ldc(coef, cz + " << off << "), m1 = ldc(coef, cz + " << off + 1 << "), m2 = ldc(coef, cz + " << off + 2
              << "), m3 = ldc(coef, cz + " << off + 3 << ");\n";
            for (int e = 0; e < E_; ++e) {
                if ((e >> I) & 1) continue;
                if (!elem_ok(t.gcmask, t.gcval, e)) continue;
                const int e1 = e | (1 << I);
                if (dot) {
                    o << "    " << accf << "ar, ai, " << B << "[" << e << "], cm2(m0, " << A << "[" << e << "], m1, " << A
                      << "[" << e1 << "])" << gsuf << ";\n";
                    o << "    " << accf << "ar, ai, " << B << "[" << e1 << "], cm2(m2, " << A << "[" << e << "], m3, " << A
                      << "[" << e1 << "])" << gsuf << ";\n";
                } else {
                    for (int s : sts)
                        o << "    " << (pred ? "d1p(" : "d1(") << "v" << s << "[" << e << "], v" << s << "[" << e1
                          << "], m0, m1, m2, m3" << gsuf << ";\n";
                }
            }
            break;
        }
        case TOP_DENSE2: {
            const int I0 = t.lpos[0], I1 = t.lpos[1];
            o << "    cplx M[16];\n";
            for (int i = 0; i < 16; ++i) o << "    M[" << i << "] = ldc(coef, cz + " << off + i << ");\n";
            for (int e = 0; e < E_; ++e) {
                if (((e >> I0) & 1) || ((e >> I1) & 1)) continue;
                if (!elem_ok(t.gcmask, t.gcval, e)) continue;
                const int idx[4] = {e, e | (1 << I1), e | (1 << I0), e | (1 << I0) | (1 << I1)};
                if (dot) {
                    for (int r = 0; r < 4; ++r) {
                        o << "    " << accf << "ar, ai, " << B << "[" << idx[r] << "], add(cm2(M[" << 4 * r << "], " << A
                          << "[" << idx[0] << "], M[" << 4 * r + 1 << "], " << A << "[" << idx[1] << "]), cm2(M["
                          << 4 * r + 2 << "], " << A << "[" << idx[2] << "], M[" << 4 * r + 3 << "], " << A << "["
                          << idx[3] << "]))" << gsuf << ";\n";
                    }
                } else {
                    for (int s : sts)
                        o << "    " << (pred ? "d2p(" : "d2(") << "v" << s << "[" << idx[0] << "], v" << s << "["
                          << idx[1] << "], v" << s << "[" << idx[2] << "], v" << s << "[" << idx[3] << "], M" << gsuf
                          << ";\n";
                }
            }
            break;
        }
        case TOP_DIAG:
        case TOP_PCPHASE:
        case TOP_PARITY: emit_phase(t, sts, dot, A, B, pred); break;
        case TOP_SIGN: {
            for (int e = 0; e < E_; ++e) {
                if (!elem_ok(t.gcmask, t.gcval, e)) continue;
                for (int s : sts) {
                    const std::string v = "v" + std::to_string(s) + "[" + std::to_string(e) + "]";
                    if (pred) o << "    " << v << " = sel(gd, ng(" << v << "), " << v << ");\n";
                    else o << "    " << v << " = ng(" << v << ");\n";
                }
            }
            break;
        }
        case TOP_CZSET: emit_czset(t, sts); break;
        case TOP_TPHASE: {
            // product of one-bit factors on bits fixed per thread, then one multiply per amplitude
            for (uint32_t j = 0; j < t.k; ++j) {
                const std::string pick = "(((gt >> " + Sh(t.gpos[j]) + ") & 1ull) ? ldc(coef, cz + " +
                                         std::to_string(off + 2 * j + 1) + ") : ldc(coef, cz + " +
                                         std::to_string(off + 2 * j) + "))";
                if (j == 0) o << "    cplx f = " << pick << ";\n";
                else o << "    f = cm(" << pick << ", f);\n";
            }
            for (int e = 0; e < E_; ++e)
                for (int s : sts) o << "    v" << s << "[" << e << "] = cm(f, v" << s << "[" << e << "]);\n";
            break;
        }
        case TOP_RPHASE: {
            for (int e = 0; e < E_; ++e)
                for (int s : sts)
                    o << "    v" << s << "[" << e << "] = cm(ldc(coef, cz + " << off + e << "), v" << s << "[" << e << "]);\n";
            break;
        }
        default: LQ_ABORT("kernel generator: unsupported op");
        }
        if (dot) o << "    dflush(ar, ai, dred, " << t.slot << "u);\n";
        o << "  }\n";
    }

    // DIAG (k <= 2 in register passes), PCPHASE, PARITY: a phase per element
    // chosen by bits of its global index.
    void emit_phase(const TileOp &t, const std::vector<int> &sts, bool dot, const std::string &A,
                    const std::string &B, bool pred) {
        std::ostringstream &o = o_;
        const uint64_t off = t.coef_off;
        // per element: phase index = thread part (runtime) | register part (constant)
        std::vector<std::string> phase_of(E_);
        if (t.kind == TOP_PARITY) {
            const uint64_t tm = t.gpmask & ~regmask();
            o << "    const cplx d0 = ldc(coef, cz + " << off << "), d1_ = ldc(coef, cz + " << off + 1 << ");\n";
            if (tm) {
                o << "    const bool pt = (pc(gt & " << K(tm) << ") & 1u) != 0u;\n";
                o << "    const cplx q0 = pt ? d1_ : d0, q1 = pt ? d0 : d1_;\n";
            } else {
                o << "    const cplx q0 = d0, q1 = d1_;\n";
            }
            for (int e = 0; e < E_; ++e) {
                int par = 0;
                for (int i = 0; i < rb_; ++i)
                    if (((t.gpmask >> geom_.tb[cur_.r[i]]) & 1ull) && ((e >> i) & 1)) par ^= 1;
                phase_of[e] = par ? "q1" : "q0";
            }
        } else {
            const int k = static_cast<int>(t.k);
            // thread-part index
            std::ostringstream ti;
            bool any = false;
            int reg_of[kMaxTileOpTargets];
            for (int j = 0; j < k; ++j) {
                reg_of[j] = reg_index_of_global(t.gpos[j]);
                if (reg_of[j] < 0) {
                    if (any) ti << " | ";
                    ti << "((unsigned)((gt >> " << Sh(t.gpos[j]) << ") & 1ull) << " << (k - 1 - j) << ")";
                    any = true;
                }
            }
            o << "    const unsigned ti = " << (any ? ti.str() : "0u") << ";\n";
            const int nd = t.kind == TOP_PCPHASE ? 2 : (1 << k);
            for (int i = 0; i < nd; ++i) o << "    const cplx d" << i << " = ldc(coef, cz + " << off + i << ");\n";
            std::vector<int> done(1 << std::max(k, 1), 0);
            for (int e = 0; e < E_; ++e) {
                uint32_t rpart = 0;
                for (int j = 0; j < k; ++j)
                    if (reg_of[j] >= 0 && ((e >> reg_of[j]) & 1)) rpart |= 1u << (k - 1 - j);
                const std::string name = "p" + std::to_string(rpart);
                if (!done[rpart]) {
                    done[rpart] = 1;
                    if (t.kind == TOP_PCPHASE) {
                        o << "    const cplx " << name << " = ((ti | " << rpart << "u) < " << t.pc_dim << "ull) ? d0 : d1;\n";
                    } else if (!any) {
                        o << "    const cplx " << name << " = d" << rpart << ";\n";
                    } else {
                        o << "    const unsigned x" << rpart << " = ti | " << rpart << "u;\n";
                        o << "    const cplx " << name << " = ";
                        for (int i = 0; i < nd - 1; ++i) o << "(x" << rpart << " == " << i << "u) ? d" << i << " : ";
                        o << "d" << nd - 1 << ";\n";
                    }
                }
                phase_of[e] = name;
            }
        }
        for (int e = 0; e < E_; ++e) {
            if (!elem_ok(t.gcmask, t.gcval, e)) continue;
            if (dot) {
                o << "    " << (pred ? "accp(" : "acc(") << "ar, ai, " << B << "[" << e << "], cm(" << phase_of[e] << ", "
                  << A << "[" << e << "])" << (pred ? ", gd)" : ")") << ";\n";
            } else {
                for (int s : sts) {
                    const std::string v = "v" + std::to_string(s) + "[" + std::to_string(e) + "]";
                    if (pred) o << "    " << v << " = sel(gd, cm(" << phase_of[e] << ", " << v << "), " << v << ");\n";
                    else o << "    " << v << " = cm(" << phase_of[e] << ", " << v << ");\n";
                }
            }
        }
    }

    // CZ runs: negate where the number of set (a, a + d) pairs is odd. Pairs
    // inside the register bits are resolved per element at generation time;
    // pairs with one register bit contribute that bit times a per-thread
    // parity; pairs outside the register bits give a per-thread parity.
    void emit_czset(const TileOp &t, const std::vector<int> &sts) {
        std::ostringstream &o = o_;
        const uint64_t rm = regmask();
        std::vector<uint64_t> tmask(rb_, 0);        // thread bits paired with register bit i
        std::vector<std::pair<uint32_t, uint64_t>> tt; // (d, mask of low bits a) both in thread part
        std::vector<uint32_t> qpairs;                // register-register pairs as (i << 8 | j)
        for (uint32_t z = 0; z < t.nz; ++z) {
            const uint32_t d = t.sd[z];
            uint64_t ttm = 0;
            for (int a = 0; a < 64; ++a) {
                if (!((t.sm[z] >> a) & 1ull)) continue;
                const int b = a + static_cast<int>(d);
                const bool ra = (rm >> a) & 1ull, rbb = (rm >> b) & 1ull;
                if (ra && rbb) qpairs.push_back((reg_index_of_global(a) << 8) | reg_index_of_global(b));
                else if (ra) tmask[reg_index_of_global(a)] |= 1ull << b;
                else if (rbb) tmask[reg_index_of_global(b)] |= 1ull << a;
                else ttm |= 1ull << a;
            }
            if (ttm) tt.emplace_back(d, ttm);
        }
        o << "    const unsigned p0 = (0u";
        for (const auto &x : tt) {
            const std::string d = Sh(static_cast<int>(x.first));
            o << " + pc(gt & (gt >> " << d << ") & " << K(x.second) << ")";
        }
        o << ") & 1u;\n";
        for (int i = 0; i < rb_; ++i)
            if (tmask[i]) o << "    const unsigned l" << i << " = pc(gt & " << K(tmask[i]) << ") & 1u;\n";
        for (int e = 0; e < E_; ++e) {
            int q = 0;
            for (uint32_t pr : qpairs)
                if (((e >> (pr >> 8)) & 1) && ((e >> (pr & 255)) & 1)) q ^= 1;
            std::ostringstream cond;
            cond << "(p0";
            for (int i = 0; i < rb_; ++i)
                if (tmask[i] && ((e >> i) & 1)) cond << " ^ l" << i;
            cond << (q ? " ^ 1u" : "") << ")";
            for (int s : sts)
                o << "    if (" << cond.str() << ") v" << s << "[" << e << "] = ng(v" << s << "[" << e << "]);\n";
        }
    }
};
} // namespace

std::string jit_pass_source(const PassPlan &p, int n, int dtype, int ns, std::vector<uint64_t> *geo) {
    if (!p.reg || p.stages.empty()) return "";
    const int kt = static_cast<int>(p.tbits.size());
    // three and four states (the adjoint's lambda and several H lambda) in
    // tiles of 11 bits, eight amplitudes per state per thread, transposes
    // through the 32 KiB LDS buffer one state at a time
    if (kt < 11 || kt > 13 || ns > 4) return "";
    if (p.slot_ids.size() > static_cast<size_t>(ns > 2 ? kMaxJitPassDots : kMaxPassDots)) return "";
    if (ns > 2 && (kt != 11 || !jit_config().lds32)) return "";
    for (const auto &st : p.stages)
        if (st.kind != STAGE_REG) return "";
    for (const auto &t : p.ops) {
        switch (t.kind) {
        case TOP_DENSE1:
        case TOP_DENSE2:
        case TOP_PARITY:
        case TOP_SIGN:
        case TOP_CZSET:
        case TOP_RPHASE:
        case TOP_TPHASE: break;
        case TOP_DIAG:
            if (t.k > 2) return "";
            break;
        case TOP_PCPHASE:
            if (t.k > 12) return "";
            break;
        default: return "";
        }
        if (t.flags & TF_ZERO_UNCTRL) return "";
    }
    Gen g(p, n, dtype, ns, geo != nullptr);
    std::string src = g.run();
    if (geo) geo->swap(g.geo());
    return src;
}

// ---------------------------------------------------------------------------
// Compile service
// ---------------------------------------------------------------------------
JitConfig &jit_config() {
    static JitConfig c = [] {
        JitConfig j;
        if (const char *s = std::getenv("LQ_JIT")) j.enabled = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_JIT_SYNC")) j.sync = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_JIT_MIN_QUBITS")) j.min_qubits = std::atoi(s);
        if (const char *s = std::getenv("LQ_JIT_THREADS")) j.threads = std::max(1, std::atoi(s));
        if (const char *s = std::getenv("LQ_JIT_NICE")) j.nice = std::atoi(s);
        if (const char *s = std::getenv("LQ_JIT_NT")) j.nontemporal = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_JIT_SPLIT")) j.split_rotations = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_JIT_LDS32")) j.lds32 = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_JIT_SCALE")) j.scale_rotations = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_JIT_WAVES")) j.waves_per_eu = std::atoi(s);
        if (const char *s = std::getenv("LQ_JIT_DIRECT")) j.direct_io = std::atoi(s);
        if (const char *s = std::getenv("LQ_JIT_LIFT")) j.lift = std::atoi(s) != 0;
        if (const char *s = std::getenv("LQ_JIT_LANES")) j.lane_swaps = std::atoi(s);
        return j;
    }();
    return c;
}

int jit_waves_per_eu(int dtype, int ns) {
    const int w = jit_config().waves_per_eu;
    if (w >= 0) return w;
    // measured on the 30-qubit HEA passes with 32 KiB transposes: 3 waves per
    // SIMD for complex128 (2: 2% slower, 4: 10% slower), 4 for complex64;
    // two-state passes hold twice the registers and get no request
    if (ns != 1) return 0;
    return dtype ? 3 : 4;
}

size_t jit_lds_bytes(int kt, int dtype, int ns) {
    const size_t per_state = static_cast<size_t>(dtype ? 16 : 8) << kt;
    if (!jit_config().lds32) return per_state * static_cast<size_t>(ns);
    return std::min<size_t>(per_state, 32768);
}

namespace {
enum class JState { Pending, Ready, Failed };

struct JEntry {
    JState state = JState::Pending;
    std::vector<char> code;
    std::string log;
    std::unordered_map<int, std::pair<hipModule_t, hipFunction_t>> loaded; // per device
};

class Service {
  public:
    static Service &get() {
        static Service *s = new Service(); // lives until exit; see stop()
        return *s;
    }

    hipFunction_t lookup(const std::string &src) { return resolve(enqueue(src)); }

    // A structural key of the pass stands for its source, so a pass seen
    // before needs no source generation.
    hipFunction_t lookup_key(const std::string &key, const PassPlan &p, int n, int dtype, int ns,
                             const std::vector<uint64_t> **geo) {
        const KeyEntry *hit = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = keys_.find(key);
            if (it != keys_.end()) hit = &it->second;
        }
        if (!hit) {
            KeyEntry k;
            std::vector<uint64_t> g;
            const std::string src = jit_pass_source(p, n, dtype, ns, jit_config().lift ? &g : nullptr);
            k.e = src.empty() ? nullptr : enqueue(src);
            k.geo = std::move(g);
            std::lock_guard<std::mutex> lk(mu_);
            hit = &keys_.emplace(key, std::move(k)).first->second; // entries are never erased
        }
        if (geo) *geo = &hit->geo;
        return hit->e ? resolve(hit->e) : nullptr;
    }

    void wait() {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return pending_ == 0 && running_ == 0; });
    }

    void stats(int *ready, int *failed, int *pending) {
        std::lock_guard<std::mutex> lk(mu_);
        if (ready) *ready = ready_;
        if (failed) *failed = failed_;
        if (pending) *pending = pending_ + running_;
    }

    // Drop queued work and wait for compiles in flight; later lookups find
    // only kernels already compiled. Runs at process exit at the latest.
    void stop() {
        std::unique_lock<std::mutex> lk(mu_);
        stopping_ = true;
        pending_ -= static_cast<int>(queue_.size());
        queue_.clear();
        cv_.notify_all();
        done_cv_.wait(lk, [this] { return running_ == 0; });
    }

  private:
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::unordered_map<std::string, std::shared_ptr<JEntry>> map_;
    struct KeyEntry {
        std::shared_ptr<JEntry> e; // null: not specialisable
        std::vector<uint64_t> geo; // the pass's position table
    };
    std::unordered_map<std::string, KeyEntry> keys_;

    // Entry for `src`, queueing its compile (or compiling now in sync mode)
    // the first time; null after stop().
    std::shared_ptr<JEntry> enqueue(const std::string &src) {
        std::shared_ptr<JEntry> e;
        {
            std::unique_lock<std::mutex> lk(mu_);
            auto it = map_.find(src);
            if (it != map_.end()) return it->second;
            if (stopping_) return nullptr;
            e = std::make_shared<JEntry>();
            map_.emplace(src, e);
            if (!jit_config().sync) {
                queue_.push_back(src);
                ++pending_;
                start_workers();
                cv_.notify_one();
                return e;
            }
            ++running_;
        }
        compile(src, *e);
        std::lock_guard<std::mutex> lk(mu_);
        --running_;
        done_cv_.notify_all();
        return e;
    }

    // The kernel of a compiled entry, loaded on the current device; null
    // while the compile is pending or after it failed.
    hipFunction_t resolve(const std::shared_ptr<JEntry> &e) {
        if (!e) return nullptr;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return nullptr;
        std::lock_guard<std::mutex> lk(mu_);
        if (e->state != JState::Ready) return nullptr;
        auto it = e->loaded.find(dev);
        if (it != e->loaded.end()) return it->second.second;
        hipModule_t mod = nullptr;
        hipFunction_t fn = nullptr;
        if (hipModuleLoadData(&mod, e->code.data()) != hipSuccess ||
            hipModuleGetFunction(&fn, mod, "lqpass") != hipSuccess) {
            e->state = JState::Failed;
            ++failed_;
            return nullptr;
        }
        e->loaded.emplace(dev, std::make_pair(mod, fn));
        ++ready_;
        return fn;
    }
    std::deque<std::string> queue_;
    int pending_ = 0, running_ = 0, ready_ = 0, failed_ = 0;
    bool started_ = false, stopping_ = false;

    std::once_flag exit_hook_;

    Service() = default;

    // The compiler's libraries are loaded by the first compile and register
    // their own exit-time destructors then; a hook registered after them runs
    // before them (exit handlers run in reverse order), so no compile is still
    // inside those libraries when they are torn down.
    void hook_exit() {
        std::call_once(exit_hook_, [] { std::atexit([] { Service::get().stop(); }); });
    }

    void start_workers() {
        if (started_) return;
        started_ = true;
        for (int i = 0; i < jit_config().threads; ++i) std::thread([this] { worker(); }).detach();
    }

    void worker() {
        // compiles are background work: they yield the CPU to the threads
        // that run circuits and requests (an unprivileged thread may lower
        // its own priority; a refusal leaves it as it was)
        if (jit_config().nice > 0)
            (void)setpriority(PRIO_PROCESS, static_cast<id_t>(syscall(SYS_gettid)), jit_config().nice);
        while (true) {
            std::string src;
            std::shared_ptr<JEntry> e;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stopping_ || !queue_.empty(); });
                if (stopping_) return;
                src = queue_.front();
                queue_.pop_front();
                e = map_[src];
                --pending_;
                ++running_;
            }
            compile(src, *e);
            std::lock_guard<std::mutex> lk(mu_);
            --running_;
            done_cv_.notify_all();
        }
    }

    static void compile(const std::string &src, JEntry &e) {
        hiprtcProgram prog;
        JState st = JState::Failed;
        std::vector<char> code;
        std::string log;
        if (hiprtcCreateProgram(&prog, src.c_str(), "lq_pass.hip", 0, nullptr, nullptr) == HIPRTC_SUCCESS) {
            const char *opts[] = {"--offload-arch=gfx950", "-O3",        "-ffp-contract=fast", "-Xclang",
                                  "-target-feature",       "-Xclang",    "-packed-fp32-ops"};
            const hiprtcResult r = hiprtcCompileProgram(prog, 7, opts);
            size_t ls = 0;
            if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
                log.resize(ls);
                hiprtcGetProgramLog(prog, &log[0]);
            }
            size_t cs = 0;
            if (r == HIPRTC_SUCCESS && hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0) {
                code.resize(cs);
                if (hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS) st = JState::Ready;
            }
            hiprtcDestroyProgram(&prog);
        }
        get().hook_exit();
        std::lock_guard<std::mutex> lk(get().mu_);
        e.code.swap(code);
        e.log.swap(log);
        e.state = st;
        if (st == JState::Failed) ++get().failed_;
    }
};
} // namespace

hipFunction_t jit_lookup(const std::string &src) { return Service::get().lookup(src); }

hipFunction_t jit_lookup_plan(const PassPlan &p, int n, int dtype, int ns, const std::vector<uint64_t> **geo) {
    // every field the generated source depends on (ops and stages are
    // zero-initialised PODs, so their bytes are a faithful key)
    std::string key;
    // (the coefficient count places the deferred-preparation marker)
    const int32_t hdr[10] = {n, dtype, ns, p.reg ? 1 : 0, static_cast<int32_t>(p.tbits.size()),
                             static_cast<int32_t>(p.load_mask), static_cast<int32_t>(p.store_mask),
                             static_cast<int32_t>(p.end_transpose), static_cast<int32_t>(p.slot_ids.size()),
                             static_cast<int32_t>(p.coef.size())};
    key.append(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    key.append(reinterpret_cast<const char *>(p.tbits.data()), p.tbits.size() * sizeof(int));
    key.append(reinterpret_cast<const char *>(p.ops.data()), p.ops.size() * sizeof(TileOp));
    key.append(reinterpret_cast<const char *>(p.stages.data()), p.stages.size() * sizeof(TileStage));
    return Service::get().lookup_key(key, p, n, dtype, ns, geo);
}
void jit_wait() { Service::get().wait(); }
void jit_shutdown() { Service::get().stop(); }
void jit_stats(int *ready, int *failed, int *pending) { Service::get().stats(ready, failed, pending); }

void jit_launch(hipFunction_t f, int dtype, void *sv0, void *sv1, void *sv2, void *sv3, const void *coef,
                void *partials, uint64_t ntiles, size_t lds_bytes, hipStream_t s) {
    void *a0 = sv0, *a1 = sv1, *a3 = partials, *a5 = sv2, *a6 = sv3;
    const void *a2 = coef;
    unsigned long long nt = ntiles;
    void *args[] = {&a0, &a1, &a2, &a3, &nt, &a5, &a6};
    const uint32_t gx = static_cast<uint32_t>(ntiles < 65536 ? ntiles : 65536);
    const uint32_t gy = static_cast<uint32_t>((ntiles + gx - 1) / gx);
    const hipError_t e = hipModuleLaunchKernel(f, gx, gy, 1, kBlock, 1, 1, static_cast<unsigned>(lds_bytes), s, args, nullptr);
    if (e != hipSuccess) {
        const std::string m = std::string("hipModuleLaunchKernel: ") + hipGetErrorString(e);
        LQ_ABORT(m.c_str());
    }
}

} // namespace lqamd
