// Device helpers shared by the tile kernels: complex arithmetic on
// double2/float2, index helpers, the LDS swizzle and the LDS op interpreter.
#pragma once

#include <hip/hip_runtime.h>

#include "lq_tileop.hpp"

namespace lqamd {
namespace dev {

template <typename T> struct VecT;
template <> struct VecT<double> { using type = double2; };
template <> struct VecT<float> { using type = float2; };

template <typename T> __device__ __forceinline__ typename VecT<T>::type mk(T x, T y) {
    typename VecT<T>::type v;
    v.x = x;
    v.y = y;
    return v;
}

template <typename T> __device__ __forceinline__ typename VecT<T>::type ld_coef(const double2 *c, uint64_t i) {
    const double2 d = c[i];
    return mk<T>(static_cast<T>(d.x), static_cast<T>(d.y));
}

template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
// a*b + c*d
template <typename V> __device__ __forceinline__ V cmul2(V a, V b, V c, V d) {
    V r;
    r.x = a.x * b.x - a.y * b.y + c.x * d.x - c.y * d.y;
    r.y = a.x * b.y + a.y * b.x + c.x * d.y + c.y * d.x;
    return r;
}
// a*b + c*d for real a, c (the real parts of a and c are used)
template <typename V> __device__ __forceinline__ V rmul2(V a, V b, V c, V d) {
    V r;
    r.x = a.x * b.x + c.x * d.x;
    r.y = a.x * b.y + c.x * d.y;
    return r;
}
template <typename V> __device__ __forceinline__ V cneg(V a) {
    V r;
    r.x = -a.x;
    r.y = -a.y;
    return r;
}
// acc += conj(b) * v, in double
template <typename V> __device__ __forceinline__ void acc_dot(double &re, double &im, V b, V v) {
    const double bx = b.x, by = b.y, vx = v.x, vy = v.y;
    re += bx * vx + by * vy;
    im += bx * vy - by * vx;
}

__device__ __forceinline__ uint32_t insert_zero(uint32_t p, int q) {
    return ((p >> q) << (q + 1)) | (p & ((1u << q) - 1u));
}

// LDS index swizzle: folds higher index bits into the bits that select the
// 128-byte bank row position, so strided tile accesses spread over banks. A
// bijection on [0, 2^kt): only the low bits change, as a function of the
// unchanged high bits.
template <int ELEM_BYTES> __device__ __forceinline__ uint32_t swz(uint32_t l) {
    if (ELEM_BYTES == 16) return l ^ (((l >> 3) ^ (l >> 6) ^ (l >> 9)) & 7u);
    return l ^ (((l >> 4) ^ (l >> 8)) & 15u);
}

// Wave-level sum of (re, im) into lane 0.
__device__ __forceinline__ void wave_sum(double &re, double &im) {
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
}

// One op on the LDS copy of a tile (kt tile-local bits, tsize = 2^kt). Gate
// ops update the states in op.smask in place; dot ops accumulate
// sum conj(B) (G A) into (accr, acci). Must be called by the whole block;
// the caller places the barrier after it. With SWZ, LDS holds element l at
// swz(l).
template <typename T, int NS, bool SWZ>
__device__ void lds_apply_op(const TileOp &op, typename VecT<T>::type *lds, uint32_t tsize, uint64_t base,
                             const double2 *coef, double &accr, double &acci) {
    using V = typename VecT<T>::type;
    constexpr int EB = sizeof(V);
    auto at = [](uint32_t l) -> uint32_t { return SWZ ? swz<EB>(l) : l; };
    const uint32_t tid = threadIdx.x;
    const uint32_t kind = op.kind;
    const bool dot = (op.flags & TF_DOT) != 0;
    const bool zero_unctrl = (op.flags & TF_ZERO_UNCTRL) != 0;
    const uint32_t smask = op.smask;
    const uint32_t lcm = op.lcmask, lcv = op.lcval;
    const V zero = mk<T>(T(0), T(0));
    const V *A = lds + op.sa * tsize;
    const V *B = lds + op.sb * tsize;

    if (kind == TOP_DENSE1) {
        const int q = op.lpos[0];
        const uint32_t qb = 1u << q;
        const V m00 = ld_coef<T>(coef, op.coef_off + 0), m01 = ld_coef<T>(coef, op.coef_off + 1);
        const V m10 = ld_coef<T>(coef, op.coef_off + 2), m11 = ld_coef<T>(coef, op.coef_off + 3);
        for (uint32_t p = tid; p < (tsize >> 1); p += kBlock) {
            const uint32_t i0 = insert_zero(p, q), i1 = i0 | qb;
            const uint32_t j0 = at(i0), j1 = at(i1);
            const bool ok = (i0 & lcm) == lcv;
            if (dot) {
                if (ok) {
                    const V a0 = A[j0], a1 = A[j1];
                    acc_dot(accr, acci, B[j0], cmul2(m00, a0, m01, a1));
                    acc_dot(accr, acci, B[j1], cmul2(m10, a0, m11, a1));
                }
            } else {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (!((smask >> s) & 1u)) continue;
                    V *S = lds + s * tsize;
                    if (ok) {
                        const V a0 = S[j0], a1 = S[j1];
                        S[j0] = cmul2(m00, a0, m01, a1);
                        S[j1] = cmul2(m10, a0, m11, a1);
                    } else if (zero_unctrl) {
                        S[j0] = zero;
                        S[j1] = zero;
                    }
                }
            }
        }
    } else if (kind == TOP_DENSE2) {
        const int q0 = op.lpos[0], q1 = op.lpos[1];
        const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
        const uint32_t b0 = 1u << q0, b1 = 1u << q1;
        V m[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) m[e] = ld_coef<T>(coef, op.coef_off + e);
        for (uint32_t p = tid; p < (tsize >> 2); p += kBlock) {
            const uint32_t i00 = insert_zero(insert_zero(p, lo), hi);
            const uint32_t idx[4] = {at(i00), at(i00 | b1), at(i00 | b0), at(i00 | b0 | b1)};
            const bool ok = (i00 & lcm) == lcv;
            if (dot) {
                if (ok) {
                    V v[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = A[idx[e]];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const V o = cmul2(m[4 * r], v[0], m[4 * r + 1], v[1]);
                        const V o2 = cmul2(m[4 * r + 2], v[2], m[4 * r + 3], v[3]);
                        acc_dot(accr, acci, B[idx[r]], mk<T>(o.x + o2.x, o.y + o2.y));
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (!((smask >> s) & 1u)) continue;
                    V *S = lds + s * tsize;
                    if (ok) {
                        V v[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) v[e] = S[idx[e]];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const V o = cmul2(m[4 * r], v[0], m[4 * r + 1], v[1]);
                            const V o2 = cmul2(m[4 * r + 2], v[2], m[4 * r + 3], v[3]);
                            S[idx[r]] = mk<T>(o.x + o2.x, o.y + o2.y);
                        }
                    } else if (zero_unctrl) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) S[idx[e]] = zero;
                    }
                }
            }
        }
    } else if (kind == TOP_DENSEK) {
        const int k = static_cast<int>(op.k);
        const uint32_t dim = 1u << k;
        int srt[kMaxTileOpTargets];
        for (int j = 0; j < k; ++j) srt[j] = op.lpos[j];
        for (int i = 1; i < k; ++i)
            for (int j = i; j > 0 && srt[j - 1] > srt[j]; --j) {
                const int t = srt[j];
                srt[j] = srt[j - 1];
                srt[j - 1] = t;
            }
        auto lsp = [&](uint32_t col) -> uint32_t {
            uint32_t o = 0;
            for (int j = 0; j < k; ++j) o |= ((col >> (k - 1 - j)) & 1u) << op.lpos[j];
            return o;
        };
        const uint32_t R = dim > static_cast<uint32_t>(kBlock) ? dim / kBlock : 1u;
        const uint32_t per_round = kBlock * R;
        const uint32_t rounds = tsize > per_round ? tsize / per_round : 1u;
        const double2 *M = coef + op.coef_off;
        for (int s = 0; s < (dot ? 1 : NS); ++s) {
            if (!dot && !((smask >> s) & 1u)) continue;
            V *S = dot ? const_cast<V *>(A) : lds + s * tsize;
            for (uint32_t r = 0; r < rounds; ++r) {
                V outv[4];
                uint32_t outi[4];
                bool outok[4];
                for (uint32_t rr = 0; rr < R; ++rr) {
                    const uint32_t e = r * per_round + rr * kBlock + tid;
                    outok[rr] = false;
                    outi[rr] = 0;
                    outv[rr] = zero;
                    if (e >= tsize) continue;
                    const uint32_t g = e >> k, row = e & (dim - 1u);
                    uint32_t gb = g;
                    for (int j = 0; j < k; ++j) gb = insert_zero(gb, srt[j]);
                    outi[rr] = gb | lsp(row);
                    if ((gb & lcm) != lcv) continue;
                    outok[rr] = true;
                    double ox = 0.0, oy = 0.0;
                    for (uint32_t col = 0; col < dim; ++col) {
                        const double2 mm = M[static_cast<uint64_t>(row) * dim + col];
                        const V v = S[at(gb | lsp(col))];
                        const T mx = static_cast<T>(mm.x), my = static_cast<T>(mm.y);
                        ox += static_cast<double>(mx * v.x - my * v.y);
                        oy += static_cast<double>(mx * v.y + my * v.x);
                    }
                    outv[rr] = mk<T>(static_cast<T>(ox), static_cast<T>(oy));
                }
                if (dot) {
                    for (uint32_t rr = 0; rr < R; ++rr)
                        if (outok[rr]) acc_dot(accr, acci, B[at(outi[rr])], outv[rr]);
                } else {
                    __syncthreads();
                    for (uint32_t rr = 0; rr < R; ++rr) {
                        const uint32_t e = r * per_round + rr * kBlock + tid;
                        if (e >= tsize) continue;
                        if (outok[rr]) S[at(outi[rr])] = outv[rr];
                        else if (zero_unctrl) S[at(outi[rr])] = zero;
                    }
                    __syncthreads();
                }
            }
        }
    } else if (kind == TOP_DIAG || kind == TOP_PCPHASE) {
        const int k = static_cast<int>(op.k);
        uint32_t idx_out = 0;
        for (int j = 0; j < k; ++j)
            if (op.lpos[j] < 0) idx_out |= static_cast<uint32_t>((base >> op.gpos[j]) & 1ull) << (k - 1 - j);
        const double2 *D = coef + op.coef_off;
        // DIAG k=0 holds one entry; only PCPHASE has a second one
        const V d0 = ld_coef<T>(D, 0);
        const V d1 = (kind == TOP_PCPHASE) ? ld_coef<T>(D, 1) : d0;
        for (uint32_t l = tid; l < tsize; l += kBlock) {
            const bool ok = (l & lcm) == lcv;
            uint32_t idx = idx_out;
            for (int j = 0; j < k; ++j)
                if (op.lpos[j] >= 0) idx |= ((l >> op.lpos[j]) & 1u) << (k - 1 - j);
            V d;
            if (kind == TOP_PCPHASE) d = (idx < op.pc_dim) ? d0 : d1;
            else d = (k == 0) ? d0 : ld_coef<T>(D, idx);
            const uint32_t j = at(l);
            if (dot) {
                if (ok) acc_dot(accr, acci, B[j], cmul(d, A[j]));
            } else {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (!((smask >> s) & 1u)) continue;
                    V *S = lds + s * tsize;
                    if (ok) S[j] = cmul(d, S[j]);
                    else if (zero_unctrl) S[j] = zero;
                }
            }
        }
    } else if (kind == TOP_PARITY) {
        const uint32_t pout = __popcll(base & op.omask) & 1u;
        const V d0 = ld_coef<T>(coef, op.coef_off), d1 = ld_coef<T>(coef, op.coef_off + 1);
        for (uint32_t l = tid; l < tsize; l += kBlock) {
            const bool ok = (l & lcm) == lcv;
            const uint32_t par = (__popc(l & op.lxmask) & 1u) ^ pout;
            const V d = par ? d1 : d0;
            const uint32_t j = at(l);
            if (dot) {
                if (ok) acc_dot(accr, acci, B[j], cmul(d, A[j]));
            } else {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (!((smask >> s) & 1u)) continue;
                    V *S = lds + s * tsize;
                    if (ok) S[j] = cmul(d, S[j]);
                    else if (zero_unctrl) S[j] = zero;
                }
            }
        }
    } else if (kind == TOP_PAULI) {
        const uint32_t lx = op.lxmask, lz = op.lzmask;
        const uint32_t sout = __popcll(base & op.omask) & 1u;
        const V cp = ld_coef<T>(coef, op.coef_off);
        const V cm = cneg(cp);
        if (dot) {
            for (uint32_t l = tid; l < tsize; l += kBlock) {
                if ((l & lcm) != lcv) continue;
                const V f = ((__popc(l & lz) & 1u) ^ sout) ? cm : cp;
                acc_dot(accr, acci, B[at(l)], cmul(f, A[at(l ^ lx)]));
            }
        } else if (lx == 0) {
            for (uint32_t l = tid; l < tsize; l += kBlock) {
                const bool ok = (l & lcm) == lcv;
                const V f = ((__popc(l & lz) & 1u) ^ sout) ? cm : cp;
                const uint32_t j = at(l);
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (!((smask >> s) & 1u)) continue;
                    V *S = lds + s * tsize;
                    if (ok) S[j] = cmul(f, S[j]);
                    else if (zero_unctrl) S[j] = zero;
                }
            }
        } else {
            const int hb = 31 - __clz(lx);
            for (uint32_t p = tid; p < (tsize >> 1); p += kBlock) {
                const uint32_t i = insert_zero(p, hb), j = i ^ lx;
                const uint32_t ai = at(i), aj = at(j);
                const bool ok = (i & lcm) == lcv;
                const V fi = ((__popc(i & lz) & 1u) ^ sout) ? cm : cp;
                const V fj = ((__popc(j & lz) & 1u) ^ sout) ? cm : cp;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (!((smask >> s) & 1u)) continue;
                    V *S = lds + s * tsize;
                    if (ok) {
                        const V vi = S[ai], vj = S[aj];
                        S[ai] = cmul(fi, vj);
                        S[aj] = cmul(fj, vi);
                    } else if (zero_unctrl) {
                        S[ai] = zero;
                        S[aj] = zero;
                    }
                }
            }
        }
    }
}

} // namespace dev
} // namespace lqamd
