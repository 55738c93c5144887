// Register-resident tile kernel for gfx950.
//
// One workgroup of 256 threads (4 wave64) owns a tile of 2^(8+RB) amplitudes
// per state. Each thread keeps E = 2^RB amplitudes per state in VGPRs. In a
// register stage the thread's amplitudes differ in the stage's RB register
// bits, so 1- and 2-qubit gates on those bits are pure register arithmetic
// (no LDS traffic, no barrier); diagonal gates, controls and CZ runs are
// evaluated from each amplitude's global index. Changing the register bits
// between stages is one LDS transpose (XOR-swizzled addresses). Loads and
// stores use the mapping "thread bits = tile bits 0..7", whose low tile bits
// are contiguous in HBM, so a wave reads and writes runs of consecutive
// amplitudes. Dot products (expvals, adjoint generator terms) are reduced per
// wave into LDS and written once per tile.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "lq_devcommon.hpp"
#include "lq_kernels.hpp"

namespace lqamd {
namespace {

using namespace dev;

template <typename T, int NS, int RB> struct Regs {
    typename VecT<T>::type v[NS][1 << RB];
};

// global index of register element e (e is a compile-time constant after unrolling)
template <int RB> __device__ __forceinline__ uint64_t gidx(uint64_t gbt, const uint64_t (&rg)[kMaxRegBits], int e) {
    uint64_t g = gbt;
#pragma unroll
    for (int i = 0; i < RB; ++i)
        if ((e >> i) & 1) g |= rg[i];
    return g;
}

// pick state s of element e from the register file (s is uniform, runtime)
template <typename T, int NS, int RB>
__device__ __forceinline__ typename VecT<T>::type pick(const Regs<T, NS, RB> &R, uint32_t s, int e) {
    typename VecT<T>::type x = R.v[0][e];
#pragma unroll
    for (int k = 1; k < NS; ++k)
        if (s == static_cast<uint32_t>(k)) x = R.v[k][e];
    return x;
}

template <int V> using IC = std::integral_constant<int, V>;

// Call f(IC<sa>, IC<sb>) with the dot's ket and bra states as constants (one
// uniform branch instead of a select per amplitude).
template <int NS, typename F> __device__ __forceinline__ void with_pair(uint32_t sa, uint32_t sb, F &&f) {
    if constexpr (NS == 1) {
        f(IC<0>{}, IC<0>{});
    } else {
        if (sa == 0) {
            if (sb == 1) f(IC<0>{}, IC<1>{});
            else f(IC<0>{}, IC<0>{});
        } else {
            if (sb == 0) f(IC<1>{}, IC<0>{});
            else f(IC<1>{}, IC<1>{});
        }
    }
}

// One-bit phase diag(d0, d1), no controls, on register bit I of the stage.
template <typename T, int NS, int RB, int I>
__device__ __forceinline__ void reg_phase_reg(Regs<T, NS, RB> &R, const TileOp &op, typename VecT<T>::type d0,
                                              typename VecT<T>::type d1, double &accr, double &acci) {
    constexpr int E = 1 << RB;
    if (op.flags & TF_DOT) {
        with_pair<NS>(op.sa, op.sb, [&](auto A, auto B) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                acc_dot(accr, acci, R.v[decltype(B)::value][e], cmul(((e >> I) & 1) ? d1 : d0, R.v[decltype(A)::value][e]));
        });
        return;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (!((op.smask >> s) & 1u)) continue;
#pragma unroll
        for (int e = 0; e < E; ++e) R.v[s][e] = cmul(((e >> I) & 1) ? d1 : d0, R.v[s][e]);
    }
}

// One-bit phase, no controls, on a bit that is fixed per thread.
template <typename T, int NS, int RB>
__device__ __forceinline__ void reg_phase_thread(Regs<T, NS, RB> &R, const TileOp &op, typename VecT<T>::type f,
                                                 double &accr, double &acci) {
    constexpr int E = 1 << RB;
    if (op.flags & TF_DOT) {
        with_pair<NS>(op.sa, op.sb, [&](auto A, auto B) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                acc_dot(accr, acci, R.v[decltype(B)::value][e], cmul(f, R.v[decltype(A)::value][e]));
        });
        return;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (!((op.smask >> s) & 1u)) continue;
#pragma unroll
        for (int e = 0; e < E; ++e) R.v[s][e] = cmul(f, R.v[s][e]);
    }
}

template <typename T, int NS, int RB, int I>
__device__ __forceinline__ void reg_dense1(Regs<T, NS, RB> &R, const TileOp &op, const double2 *coef, uint64_t gbt,
                                           const uint64_t (&rg)[kMaxRegBits], double &accr, double &acci) {
    using V = typename VecT<T>::type;
    constexpr int E = 1 << RB;
    const V m00 = ld_coef<T>(coef, op.coef_off + 0), m01 = ld_coef<T>(coef, op.coef_off + 1);
    const V m10 = ld_coef<T>(coef, op.coef_off + 2), m11 = ld_coef<T>(coef, op.coef_off + 3);
    const uint64_t gcm = op.gcmask, gcv = op.gcval;
    if (op.flags & TF_DOT) {
        with_pair<NS>(op.sa, op.sb, [&](auto A, auto B) {
            constexpr int sa = decltype(A)::value, sb = decltype(B)::value;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (e & (1 << I)) continue;
                const int e1 = e | (1 << I);
                if (gcm != 0 && (gidx<RB>(gbt, rg, e) & gcm) != gcv) continue;
                const V a0 = R.v[sa][e], a1 = R.v[sa][e1];
                acc_dot(accr, acci, R.v[sb][e], cmul2(m00, a0, m01, a1));
                acc_dot(accr, acci, R.v[sb][e1], cmul2(m10, a0, m11, a1));
            }
        });
        return;
    }
    // real matrices (RY, X of a CNOT, H, ...) take half the arithmetic; the
    // test is on the op's coefficients, uniform over the workgroup
    const bool real = m00.y == T(0) && m01.y == T(0) && m10.y == T(0) && m11.y == T(0);
    if (gcm == 0) {
        if (real) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (!((op.smask >> s) & 1u)) continue;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (e & (1 << I)) continue;
                    const int e1 = e | (1 << I);
                    const V a0 = R.v[s][e], a1 = R.v[s][e1];
                    R.v[s][e] = rmul2(m00, a0, m01, a1);
                    R.v[s][e1] = rmul2(m10, a0, m11, a1);
                }
            }
            return;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((op.smask >> s) & 1u)) continue;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (e & (1 << I)) continue;
                const int e1 = e | (1 << I);
                const V a0 = R.v[s][e], a1 = R.v[s][e1];
                R.v[s][e] = cmul2(m00, a0, m01, a1);
                R.v[s][e1] = cmul2(m10, a0, m11, a1);
            }
        }
        return;
    }
    if (real) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (e & (1 << I)) continue;
            const int e1 = e | (1 << I);
            const bool ok = (gidx<RB>(gbt, rg, e) & gcm) == gcv;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (!((op.smask >> s) & 1u)) continue;
                const V a0 = R.v[s][e], a1 = R.v[s][e1];
                const V n0 = rmul2(m00, a0, m01, a1), n1 = rmul2(m10, a0, m11, a1);
                R.v[s][e] = ok ? n0 : a0;
                R.v[s][e1] = ok ? n1 : a1;
            }
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (e & (1 << I)) continue;
        const int e1 = e | (1 << I);
        const bool ok = (gidx<RB>(gbt, rg, e) & gcm) == gcv;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((op.smask >> s) & 1u)) continue;
            const V a0 = R.v[s][e], a1 = R.v[s][e1];
            const V n0 = cmul2(m00, a0, m01, a1), n1 = cmul2(m10, a0, m11, a1);
            R.v[s][e] = ok ? n0 : a0;
            R.v[s][e1] = ok ? n1 : a1;
        }
    }
}

template <typename T, int NS, int RB, int I0, int I1>
__device__ __forceinline__ void reg_dense2(Regs<T, NS, RB> &R, const TileOp &op, const double2 *coef, uint64_t gbt,
                                           const uint64_t (&rg)[kMaxRegBits], double &accr, double &acci) {
    using V = typename VecT<T>::type;
    constexpr int E = 1 << RB;
    V m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = ld_coef<T>(coef, op.coef_off + i);
    const uint64_t gcm = op.gcmask, gcv = op.gcval;
    const bool dot = (op.flags & TF_DOT) != 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (e & ((1 << I0) | (1 << I1))) continue;
        const int idx[4] = {e, e | (1 << I1), e | (1 << I0), e | (1 << I0) | (1 << I1)};
        const bool ok = gcm == 0 || (gidx<RB>(gbt, rg, e) & gcm) == gcv;
        if (dot) {
            if (!ok) continue;
            V a[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) a[c] = pick(R, op.sa, idx[c]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const V o = cmul2(m[4 * r], a[0], m[4 * r + 1], a[1]);
                const V o2 = cmul2(m[4 * r + 2], a[2], m[4 * r + 3], a[3]);
                acc_dot(accr, acci, pick(R, op.sb, idx[r]), mk<T>(o.x + o2.x, o.y + o2.y));
            }
            continue;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((op.smask >> s) & 1u)) continue;
            V a[4], o[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) a[c] = R.v[s][idx[c]];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const V p = cmul2(m[4 * r], a[0], m[4 * r + 1], a[1]);
                const V q = cmul2(m[4 * r + 2], a[2], m[4 * r + 3], a[3]);
                o[r] = mk<T>(p.x + q.x, p.y + q.y);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) R.v[s][idx[r]] = ok ? o[r] : a[r];
        }
    }
}

// Diagonal kinds: DIAG (k <= 2), PARITY, PCPHASE, SIGN, CZSET
template <typename T, int NS, int RB>
__device__ __forceinline__ void reg_diag(Regs<T, NS, RB> &R, const TileOp &op, const double2 *coef, uint64_t gbt,
                                         const uint64_t (&rg)[kMaxRegBits], double &accr, double &acci) {
    using V = typename VecT<T>::type;
    constexpr int E = 1 << RB;
    const uint32_t kind = op.kind;
    const bool dot = (op.flags & TF_DOT) != 0;
    const uint64_t gcm = op.gcmask, gcv = op.gcval;
    if (kind == TOP_SIGN || kind == TOP_CZSET) {
        const uint32_t nz = op.nz;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint64_t g = gidx<RB>(gbt, rg, e);
            bool neg;
            if (kind == TOP_SIGN) {
                neg = (g & gcm) == gcv;
            } else {
                uint32_t p = 0;
                for (uint32_t i = 0; i < nz; ++i) p += __popcll(g & (g >> op.sd[i]) & op.sm[i]);
                neg = (p & 1u) != 0;
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (!((op.smask >> s) & 1u)) continue;
                R.v[s][e] = neg ? cneg(R.v[s][e]) : R.v[s][e];
            }
        }
        return;
    }
    if (kind == TOP_TPHASE) {
        // product of one-bit factors on bits fixed per thread
        V f = ld_coef<T>(coef, op.coef_off + ((gbt >> op.gpos[0]) & 1ull));
        for (uint32_t j = 1; j < op.k; ++j)
            f = cmul(ld_coef<T>(coef, op.coef_off + 2 * j + ((gbt >> op.gpos[j]) & 1ull)), f);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((op.smask >> s) & 1u)) continue;
#pragma unroll
            for (int e = 0; e < E; ++e) R.v[s][e] = cmul(f, R.v[s][e]);
        }
        return;
    }
    if (kind == TOP_RPHASE) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const V f = ld_coef<T>(coef, op.coef_off + e);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (!((op.smask >> s) & 1u)) continue;
                R.v[s][e] = cmul(f, R.v[s][e]);
            }
        }
        return;
    }
    if (kind == TOP_DIAG && op.k == 1 && gcm == 0) {
        const V d0 = ld_coef<T>(coef, op.coef_off), d1 = ld_coef<T>(coef, op.coef_off + 1);
        const uint64_t b = 1ull << op.gpos[0];
        int r = -1;
#pragma unroll
        for (int i = 0; i < RB; ++i)
            if (rg[i] == b) r = i;
        switch (r) {
        case 0: reg_phase_reg<T, NS, RB, 0>(R, op, d0, d1, accr, acci); return;
        case 1: if constexpr (RB > 1) reg_phase_reg<T, NS, RB, 1>(R, op, d0, d1, accr, acci); return;
        case 2: if constexpr (RB > 2) reg_phase_reg<T, NS, RB, 2>(R, op, d0, d1, accr, acci); return;
        case 3: if constexpr (RB > 3) reg_phase_reg<T, NS, RB, 3>(R, op, d0, d1, accr, acci); return;
        default: reg_phase_thread<T, NS, RB>(R, op, (gbt & b) ? d1 : d0, accr, acci); return;
        }
    }
    const int k = static_cast<int>(op.k);
    V d[4];
    const int nd = kind == TOP_DIAG ? (1 << k) : 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) d[i] = i < nd ? ld_coef<T>(coef, op.coef_off + i) : d[0];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint64_t g = gidx<RB>(gbt, rg, e);
        const bool ok = (g & gcm) == gcv;
        uint32_t idx = 0;
        if (kind == TOP_PARITY) {
            idx = __popcll(g & op.gpmask) & 1u;
        } else {
            for (int j = 0; j < k; ++j) idx |= static_cast<uint32_t>((g >> op.gpos[j]) & 1ull) << (k - 1 - j);
            if (kind == TOP_PCPHASE) idx = idx < op.pc_dim ? 0u : 1u;
        }
        // two-level select on the index bits (a select chain on idx became a
        // dynamically indexed private array, i.e. scratch memory)
        const V f01 = (idx & 1u) ? d[1] : d[0];
        const V f23 = (idx & 1u) ? d[3] : d[2];
        const V f = (idx & 2u) ? f23 : f01;
        if (dot) {
            if (ok) acc_dot(accr, acci, pick(R, op.sb, e), cmul(f, pick(R, op.sa, e)));
            continue;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((op.smask >> s) & 1u)) continue;
            const V x = R.v[s][e];
            R.v[s][e] = ok ? cmul(f, x) : x;
        }
    }
}

template <typename T, int NS, int RB, bool WIDE>
__device__ __forceinline__ void reg_apply(Regs<T, NS, RB> &R, const TileOp &op, const double2 *coef, uint64_t gbt,
                                          const uint64_t (&rg)[kMaxRegBits], double &accr, double &acci) {
    const uint32_t kind = op.kind;
    if (kind == TOP_DENSE1) {
        switch (op.lpos[0]) {
        case 0: reg_dense1<T, NS, RB, 0>(R, op, coef, gbt, rg, accr, acci); break;
        case 1: if constexpr (RB > 1) reg_dense1<T, NS, RB, 1>(R, op, coef, gbt, rg, accr, acci); break;
        case 2: if constexpr (RB > 2) reg_dense1<T, NS, RB, 2>(R, op, coef, gbt, rg, accr, acci); break;
        case 3: if constexpr (RB > 3) reg_dense1<T, NS, RB, 3>(R, op, coef, gbt, rg, accr, acci); break;
        default: break;
        }
    } else if (kind == TOP_DENSE2) {
        if constexpr (WIDE) {
#define LQ_D2(a, b)                                                                                          \
    case (a)*4 + (b):                                                                                        \
        if constexpr ((a) < RB && (b) < RB) reg_dense2<T, NS, RB, a, b>(R, op, coef, gbt, rg, accr, acci);   \
        break;
        switch (op.lpos[0] * 4 + op.lpos[1]) {
            LQ_D2(0, 1) LQ_D2(0, 2) LQ_D2(0, 3) LQ_D2(1, 0) LQ_D2(1, 2) LQ_D2(1, 3)
            LQ_D2(2, 0) LQ_D2(2, 1) LQ_D2(2, 3) LQ_D2(3, 0) LQ_D2(3, 1) LQ_D2(3, 2)
        default: break;
        }
#undef LQ_D2
        }
    } else {
        reg_diag<T, NS, RB>(R, op, coef, gbt, rg, accr, acci);
    }
}

// Every global pointer is its own __restrict__ parameter: the program (ops,
// stages, coefficients) is then provably unclobbered by the state stores and
// is read with scalar loads through the scalar cache.
// WIDE: two-qubit dense ops and LDS stages are compiled in; passes without
// them run the narrow form, whose smaller register footprint lets more waves
// share a SIMD.
template <typename T, int NS, int RB, bool WIDE>
__global__ __launch_bounds__(kBlock) void tile_reg_kernel(TileRegArgs a, typename VecT<T>::type *__restrict__ sv0,
                                                         typename VecT<T>::type *__restrict__ sv1,
                                                         const TileOp *__restrict__ ops,
                                                         const TileStage *__restrict__ stages,
                                                         const double2 *__restrict__ coef,
                                                         double2 *__restrict__ partials) {
    using V = typename VecT<T>::type;
    constexpr int E = 1 << RB;
    constexpr int EB = sizeof(V);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V *lds = reinterpret_cast<V *>(smem_raw);
    __shared__ double2 dred[kBlock / 64][kMaxPassDots];

    const uint32_t tid = threadIdx.x;
    const uint64_t tile = static_cast<uint64_t>(blockIdx.x) + static_cast<uint64_t>(blockIdx.y) * gridDim.x;
    if (tile >= a.ntiles) return; // uniform over the block
    const int kt = a.g.kt;        // == 8 + RB
    const int c = a.g.c;
    const uint32_t tsize = 1u << kt;

    uint64_t base = 0;
    for (int j = 0; j < a.g.nout; ++j) base |= ((tile >> j) & 1ull) << a.g.ob[j];
    auto spread = [&](uint32_t l) -> uint64_t {
        uint64_t s = l & ((1u << c) - 1u);
        for (int b = c; b < kt; ++b) s |= static_cast<uint64_t>((l >> b) & 1u) << a.g.tb[b];
        return s;
    };
    if (a.nslots) {
        for (uint32_t i = tid; i < (kBlock / 64) * kMaxPassDots; i += kBlock)
            dred[i / kMaxPassDots][i % kMaxPassDots] = make_double2(0.0, 0.0);
        __syncthreads(); // the zeros land before any wave writes its dot partials
    }

    // load with the HBM mapping: tile-local l = tid | (e << 8)
    Regs<T, NS, RB> R;
    uint64_t hi[kMaxRegBits];
#pragma unroll
    for (int i = 0; i < kMaxRegBits; ++i) hi[i] = i < RB ? (1ull << a.g.tb[8 + i]) : 0ull;
    const uint64_t sp_tid = spread(tid);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (!((a.load_mask >> s) & 1u)) continue;
        const V *src = s == 0 ? sv0 : sv1;
#pragma unroll
        for (int e = 0; e < E; ++e) R.v[s][e] = src[base | sp_tid | gidx<RB>(0, hi, e)];
    }

    // current register mapping (uniform)
    int cr[kMaxRegBits], ch[8];
#pragma unroll
    for (int i = 0; i < kMaxRegBits; ++i) cr[i] = 8 + i;
#pragma unroll
    for (int j = 0; j < 8; ++j) ch[j] = j;
    auto local_of_thread = [&](const int *h) -> uint32_t {
        uint32_t l = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) l |= ((tid >> j) & 1u) << h[j];
        return l;
    };
    auto local_of_elem = [&](const int *r, int e) -> uint32_t {
        uint32_t l = 0;
#pragma unroll
        for (int i = 0; i < RB; ++i)
            if ((e >> i) & 1) l |= 1u << r[i];
        return l;
    };
    auto write_regs = [&]() {
        const uint32_t lt = local_of_thread(ch);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((a.load_mask >> s) & 1u)) continue;
#pragma unroll
            for (int e = 0; e < E; ++e) lds[s * tsize + swz<EB>(lt | local_of_elem(cr, e))] = R.v[s][e];
        }
    };
    auto read_regs = [&]() {
        const uint32_t lt = local_of_thread(ch);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((a.load_mask >> s) & 1u)) continue;
#pragma unroll
            for (int e = 0; e < E; ++e) R.v[s][e] = lds[s * tsize + swz<EB>(lt | local_of_elem(cr, e))];
        }
    };

    bool in_lds = false;
    for (uint32_t si = 0; si < a.nstages; ++si) {
        const TileStage &st = stages[si];
        if (st.kind == STAGE_REG) {
            if (st.entry == ENTRY_TRANSPOSE) {
                __syncthreads();
                write_regs();
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < kMaxRegBits; ++i) cr[i] = i < RB ? st.rbits[i] : 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) ch[j] = st.hbits[j];
            if (st.entry != ENTRY_NONE) read_regs();
            in_lds = false;
            const uint64_t gbt = base | spread(local_of_thread(ch));
            uint64_t rg[kMaxRegBits];
#pragma unroll
            for (int i = 0; i < kMaxRegBits; ++i) rg[i] = i < RB ? (1ull << a.g.tb[cr[i]]) : 0ull;
            for (uint32_t oi = st.op_begin; oi < st.op_end; ++oi) {
                const TileOp &op = ops[oi];
                double accr = 0.0, acci = 0.0;
                reg_apply<T, NS, RB, WIDE>(R, op, coef, gbt, rg, accr, acci);
                if (op.flags & TF_DOT) {
                    wave_sum(accr, acci);
                    if ((tid & 63u) == 0) dred[tid >> 6][op.slot] = make_double2(accr, acci);
                }
            }
        } else if constexpr (WIDE) {
            if (!in_lds) {
                __syncthreads();
                write_regs();
                __syncthreads();
                in_lds = true;
            }
            for (uint32_t oi = st.op_begin; oi < st.op_end; ++oi) {
                const TileOp &op = ops[oi];
                if ((base & op.ocmask) == op.ocval) {
                    double accr = 0.0, acci = 0.0;
                    lds_apply_op<T, NS, true>(op, lds, tsize, base, coef, accr, acci);
                    if (op.flags & TF_DOT) {
                        wave_sum(accr, acci);
                        if ((tid & 63u) == 0) dred[tid >> 6][op.slot] = make_double2(accr, acci);
                    }
                } else if (op.flags & TF_ZERO_UNCTRL) {
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        if ((op.smask >> s) & 1u)
                            for (uint32_t l = tid; l < tsize; l += kBlock) lds[s * tsize + l] = mk<T>(T(0), T(0));
                }
                __syncthreads();
            }
        }
    }

    if (a.store_mask) {
        if (in_lds || a.end_transpose) {
            if (!in_lds) {
                __syncthreads();
                write_regs();
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < kMaxRegBits; ++i) cr[i] = 8 + i;
#pragma unroll
            for (int j = 0; j < 8; ++j) ch[j] = j;
            read_regs();
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!((a.store_mask >> s) & 1u)) continue;
            V *dst = s == 0 ? sv0 : sv1;
#pragma unroll
            for (int e = 0; e < E; ++e) dst[base | sp_tid | gidx<RB>(0, hi, e)] = R.v[s][e];
        }
    }
    if (a.nslots) {
        __syncthreads();
        for (uint32_t sl = tid; sl < a.nslots; sl += kBlock) {
            double2 t = dred[0][sl];
#pragma unroll
            for (int w = 1; w < kBlock / 64; ++w) {
                t.x += dred[w][sl].x;
                t.y += dred[w][sl].y;
            }
            partials[static_cast<uint64_t>(sl) * a.ntiles + tile] = t;
        }
    }
}

} // namespace

template <typename T> void launch_tile_reg(const TileRegArgs &a, size_t lds_bytes, hipStream_t s) {
    const uint64_t nt = a.ntiles;
    const uint32_t gx = static_cast<uint32_t>(nt < 65536 ? nt : 65536);
    const uint32_t gy = static_cast<uint32_t>((nt + gx - 1) / gx);
    const dim3 grid(gx, gy);
    const int rb = a.g.kt - 8;
    using V = typename VecT<T>::type;
    V *p0 = static_cast<V *>(a.sv[0]);
    V *p1 = a.ns > 1 ? static_cast<V *>(a.sv[1]) : nullptr;
#define LQ_LAUNCH(NS_, RB_)                                                                                    \
    do {                                                                                                       \
        if (a.wide)                                                                                            \
            hipLaunchKernelGGL((tile_reg_kernel<T, NS_, RB_, true>), grid, dim3(kBlock), lds_bytes, s, a, p0,  \
                               p1, a.ops, a.stages, a.coef, a.partials);                                       \
        else                                                                                                   \
            hipLaunchKernelGGL((tile_reg_kernel<T, NS_, RB_, false>), grid, dim3(kBlock), lds_bytes, s, a, p0, \
                               p1, a.ops, a.stages, a.coef, a.partials);                                       \
    } while (0)
    if (a.ns == 1) {
        switch (rb) {
        case 3: LQ_LAUNCH(1, 3); break;
        default: LQ_LAUNCH(1, 4); break;
        }
    } else {
        switch (rb) {
        case 3: LQ_LAUNCH(2, 3); break;
        default: LQ_LAUNCH(2, 4); break;
        }
    }
#undef LQ_LAUNCH
}
template void launch_tile_reg<double>(const TileRegArgs &, size_t, hipStream_t);
template void launch_tile_reg<float>(const TileRegArgs &, size_t, hipStream_t);

} // namespace lqamd
