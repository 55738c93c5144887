// Packed small-request kernel: one 256-thread workgroup per request runs the
// request's whole program (forward, H lambda, energy, backward sweep) on its
// lambda and mu. Up to 12 qubits (complex128) both states live in LDS, so a
// request costs no HBM traffic besides its program; up to 16 they live in a
// global scratch area that only this workgroup touches. Ops run in program
// order with a workgroup barrier between them; dot ops reduce per wave, then
// across the four waves, into the request's result slots.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lq_batch.hpp"
#include "lq_devcommon.hpp"

namespace lqamd {
namespace {

using namespace dev;

// LDS: the states in dynamic LDS. Otherwise in `scratch` (only this
// workgroup touches them; the barrier between ops orders the global accesses
// of the workgroup's threads).
template <typename T, bool LDS>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BReq *__restrict__ reqs, const BOp *__restrict__ ops,
                                                       const double2 *__restrict__ coef,
                                                       const PauliTerm *__restrict__ terms,
                                                       double2 *__restrict__ out,
                                                       typename VecT<T>::type *__restrict__ scratch) {
    using V = typename VecT<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double2 red[kBlock / 64];
    const BReq r = reqs[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const uint32_t N = 1u << r.n;
    V *S0 = LDS ? reinterpret_cast<V *>(smem_raw) : scratch + r.state_off;
    V *S1 = S0 + N;
    for (uint32_t i = tid; i < N; i += kBlock) {
        S0[i] = mk<T>(i == 0 ? T(1) : T(0), T(0));
        S1[i] = mk<T>(T(0), T(0));
    }
    __syncthreads();

    for (uint32_t o = r.op_begin; o < r.op_end; ++o) {
        const BOp op = ops[o];
        const bool dot = (op.flags & BOP_F_DOT) != 0;
        const bool s0 = (op.smask & 1u) != 0, s1 = (op.smask & 2u) != 0;
        double ar = 0.0, ai = 0.0;
        switch (op.kind) {
        case BOP_DENSE1: {
            const V m00 = ld_coef<T>(coef, op.arg), m01 = ld_coef<T>(coef, op.arg + 1);
            const V m10 = ld_coef<T>(coef, op.arg + 2), m11 = ld_coef<T>(coef, op.arg + 3);
            const uint32_t tb = 1u << op.t0;
            for (uint32_t p = tid; p < (N >> 1); p += kBlock) {
                const uint32_t i0 = insert_zero(p, op.t0), i1 = i0 | tb;
                if ((i0 & op.cmask) != op.cval) continue;
                if (dot) {
                    const V a0 = S0[i0], a1 = S0[i1];
                    acc_dot(ar, ai, S1[i0], cmul2(m00, a0, m01, a1));
                    acc_dot(ar, ai, S1[i1], cmul2(m10, a0, m11, a1));
                    continue;
                }
                if (s0) {
                    const V a0 = S0[i0], a1 = S0[i1];
                    S0[i0] = cmul2(m00, a0, m01, a1);
                    S0[i1] = cmul2(m10, a0, m11, a1);
                }
                if (s1) {
                    const V a0 = S1[i0], a1 = S1[i1];
                    S1[i0] = cmul2(m00, a0, m01, a1);
                    S1[i1] = cmul2(m10, a0, m11, a1);
                }
            }
            break;
        }
        case BOP_DENSE2: {
            V m[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) m[i] = ld_coef<T>(coef, op.arg + i);
            const uint32_t lo = op.t0 < op.t1 ? op.t0 : op.t1, hi = op.t0 < op.t1 ? op.t1 : op.t0;
            const uint32_t b0 = 1u << op.t0, b1 = 1u << op.t1;
            for (uint32_t p = tid; p < (N >> 2); p += kBlock) {
                const uint32_t i00 = insert_zero(insert_zero(p, lo), hi);
                if ((i00 & op.cmask) != op.cval) continue;
                const uint32_t idx[4] = {i00, i00 | b1, i00 | b0, i00 | b0 | b1};
                if (dot) {
                    V a[4];
#pragma unroll
                    for (int c = 0; c < 4; ++c) a[c] = S0[idx[c]];
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) {
                        const V x = cmul2(m[4 * rr], a[0], m[4 * rr + 1], a[1]);
                        const V y = cmul2(m[4 * rr + 2], a[2], m[4 * rr + 3], a[3]);
                        acc_dot(ar, ai, S1[idx[rr]], mk<T>(x.x + y.x, x.y + y.y));
                    }
                    continue;
                }
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    if (!((op.smask >> s) & 1u)) continue;
                    V *S = s ? S1 : S0;
                    V a[4];
#pragma unroll
                    for (int c = 0; c < 4; ++c) a[c] = S[idx[c]];
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) {
                        const V x = cmul2(m[4 * rr], a[0], m[4 * rr + 1], a[1]);
                        const V y = cmul2(m[4 * rr + 2], a[2], m[4 * rr + 3], a[3]);
                        S[idx[rr]] = mk<T>(x.x + y.x, x.y + y.y);
                    }
                }
            }
            break;
        }
        case BOP_DIAG:
        case BOP_PARITY: {
            const bool par = op.kind == BOP_PARITY;
            const uint32_t nd = par ? 2u : (1u << op.k);
            V d[4];
#pragma unroll
            for (uint32_t i = 0; i < 4; ++i) d[i] = i < nd ? ld_coef<T>(coef, op.arg + i) : d[0];
            for (uint32_t i = tid; i < N; i += kBlock) {
                if ((i & op.cmask) != op.cval) continue;
                uint32_t idx;
                if (par) {
                    idx = __popc(i & op.arg2) & 1u;
                } else if (op.k == 0) {
                    idx = 0;
                } else if (op.k == 1) {
                    idx = (i >> op.t0) & 1u;
                } else {
                    idx = (((i >> op.t0) & 1u) << 1) | ((i >> op.t1) & 1u);
                }
                // select by index bits (a select chain on idx compiles to a
                // dynamically indexed private array in scratch memory)
                const V f01 = (idx & 1u) ? d[1] : d[0];
                const V f23 = (idx & 1u) ? d[3] : d[2];
                const V f = (idx & 2u) ? f23 : f01;
                if (dot) {
                    acc_dot(ar, ai, S1[i], cmul(f, S0[i]));
                    continue;
                }
                if (s0) S0[i] = cmul(f, S0[i]);
                if (s1) S1[i] = cmul(f, S1[i]);
            }
            break;
        }
        case BOP_HAM: {
            for (uint32_t i = tid; i < N; i += kBlock) {
                double sr = 0.0, si = 0.0;
                for (uint32_t t = 0; t < op.arg2; ++t) {
                    const PauliTerm pt = terms[op.arg + t];
                    const V v = S0[i ^ static_cast<uint32_t>(pt.x)];
                    const double sg = (__popc(i & static_cast<uint32_t>(pt.yz)) & 1) ? -1.0 : 1.0;
                    const double cr = sg * pt.cre, ci = sg * pt.cim;
                    const double vx = v.x, vy = v.y;
                    sr += cr * vx - ci * vy;
                    si += cr * vy + ci * vx;
                }
                S1[i] = mk<T>(static_cast<T>(sr), static_cast<T>(si));
            }
            break;
        }
        case BOP_DOTID: {
            for (uint32_t i = tid; i < N; i += kBlock) acc_dot(ar, ai, S1[i], S0[i]);
            break;
        }
        default: break;
        }
        if (dot) {
            wave_sum(ar, ai);
            if ((tid & 63u) == 0) red[tid >> 6] = make_double2(ar, ai);
            __syncthreads();
            if (tid == 0) {
                double2 t = red[0];
#pragma unroll
                for (int w = 1; w < kBlock / 64; ++w) {
                    t.x += red[w].x;
                    t.y += red[w].y;
                }
                out[r.out_off + op.arg2] = t;
            }
        }
        __syncthreads();
    }
}

} // namespace

namespace {
__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256u + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * 256u)
        dst[i] = src[i];
}
} // namespace

void launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    if (n == 0) return;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 1024));
    hipLaunchKernelGGL(copy16_kernel, dim3(blocks), dim3(256), 0, s, static_cast<const uint4 *>(src),
                       static_cast<uint4 *>(dst), n);
}

void launch_batch(const BReq *reqs, const BOp *ops, const double2 *coef, const PauliTerm *terms, double2 *out,
                  void *scratch, uint32_t nreq, int max_n, bool in_lds, int dtype, hipStream_t s) {
    if (nreq == 0) return;
    const size_t elem = dtype ? 16 : 8;
    const size_t lds = in_lds ? 2 * (elem << max_n) : 0;
#define LQ_BATCH(T_, L_)                                                                                      \
    hipLaunchKernelGGL((batch_kernel<T_, L_>), dim3(nreq), dim3(kBlock), lds, s, reqs, ops, coef, terms, out, \
                       static_cast<typename VecT<T_>::type *>(scratch))
    if (dtype) {
        if (in_lds) LQ_BATCH(double, true);
        else LQ_BATCH(double, false);
    } else {
        if (in_lds) LQ_BATCH(float, true);
        else LQ_BATCH(float, false);
    }
#undef LQ_BATCH
}

} // namespace lqamd
