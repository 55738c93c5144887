// HIP kernels for gfx950 (MI355X): the LDS tile-program kernel that applies
// runs of gates (and reads generator inner products) in one HBM pass, plus
// the standalone reductions and fills around it.
//
// Amplitude layout: interleaved complex (double2 / float2), index bit n-1-w
// is wire w (lq/StateVectorLQubit.hpp:950). All sums are accumulated in
// double, also for complex64 states.
#include <hip/hip_runtime.h>

#include "lq_devcommon.hpp"
#include "lq_kernels.hpp"

namespace lqamd {
namespace {

using dev::VecT;
using dev::mk;
using dev::ld_coef;
using dev::cmul;
using dev::acc_dot;

__device__ __forceinline__ void block_reduce_store(double re, double im, double2 *red, double2 *dst) {
    dev::wave_sum(re, im);
    const uint32_t tid = threadIdx.x;
    if ((tid & 63u) == 0) red[tid >> 6] = make_double2(re, im);
    __syncthreads();
    if (tid == 0) {
        double2 s = red[0];
        for (int w = 1; w < kBlock / 64; ++w) {
            s.x += red[w].x;
            s.y += red[w].y;
        }
        *dst = s;
    }
}

// ---------------------------------------------------------------------------
// Small-tile kernel: the whole tile lives in LDS and every op sweeps it.
// Used for tiles of at most 2^10 amplitudes and for passes with more than two
// states; larger tiles run in tile_reg_kernel (lq_kernels_reg.hip).
// ---------------------------------------------------------------------------
template <typename T, int NS>
__global__ __launch_bounds__(kBlock) void tile_pass_kernel(TilePassArgs a) {
    using V = typename VecT<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V *lds = reinterpret_cast<V *>(smem_raw);
    __shared__ double2 red[kBlock / 64];

    const uint32_t tid = threadIdx.x;
    const uint64_t tile = static_cast<uint64_t>(blockIdx.x) + static_cast<uint64_t>(blockIdx.y) * gridDim.x;
    if (tile >= a.ntiles) return; // uniform over the block
    const int kt = a.g.kt;
    const int c = a.g.c;
    const uint32_t tsize = 1u << kt;

    uint64_t base = 0;
    for (int j = 0; j < a.g.nout; ++j) base |= ((tile >> j) & 1ull) << a.g.ob[j];

    // spread(l) = base offset of tile-local index l; separable over the bits
    // of l, so spread(tid + it*kBlock) = sp_tid | spread(it*kBlock).
    auto spread = [&](uint32_t l) -> uint64_t {
        uint64_t s = l & ((1u << c) - 1u);
        for (int b = c; b < kt; ++b) s |= static_cast<uint64_t>((l >> b) & 1u) << a.g.tb[b];
        return s;
    };
    const uint64_t sp_tid = spread(tid < tsize ? tid : 0u);
    const uint32_t iters = (tsize + kBlock - 1) / kBlock;

#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (!((a.load_mask >> s) & 1u)) continue;
        const V *src = reinterpret_cast<const V *>(a.sv[s]);
        for (uint32_t it = 0; it < iters; ++it) {
            const uint32_t l = tid + it * kBlock;
            if (l < tsize) lds[s * tsize + l] = src[base | sp_tid | spread(it * kBlock)];
        }
    }
    __syncthreads();

    const V zero = mk<T>(T(0), T(0));
    for (uint32_t oi = 0; oi < a.nops; ++oi) {
        const TileOp &op = a.ops[oi];
        const bool dot = (op.flags & TF_DOT) != 0;
        double2 *dst = dot ? &a.partials[static_cast<uint64_t>(op.slot) * a.ntiles + tile] : nullptr;
        if ((base & op.ocmask) != op.ocval) {
            // control bits outside the tile do not match: no-op for this tile
            if (dot) {
                if (tid == 0) *dst = make_double2(0.0, 0.0);
            } else if (op.flags & TF_ZERO_UNCTRL) {
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    if ((op.smask >> s) & 1u)
                        for (uint32_t l = tid; l < tsize; l += kBlock) lds[s * tsize + l] = zero;
            }
            __syncthreads();
            continue;
        }
        double accr = 0.0, acci = 0.0;
        dev::lds_apply_op<T, NS, false>(op, lds, tsize, base, a.coef, accr, acci);
        if (dot) block_reduce_store(accr, acci, red, dst);
        __syncthreads();
    }

#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (!((a.store_mask >> s) & 1u)) continue;
        V *dstp = reinterpret_cast<V *>(a.sv[s]);
        for (uint32_t it = 0; it < iters; ++it) {
            const uint32_t l = tid + it * kBlock;
            if (l < tsize) dstp[base | sp_tid | spread(it * kBlock)] = lds[s * tsize + l];
        }
    }
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
// One block per slot; each thread keeps a Neumaier-compensated sum over a
// strided subset, then a fixed-order tree.
__global__ __launch_bounds__(kBlock) void reduce_slots_kernel(const double2 *partials, const SlotRange *ranges,
                                                             double2 *out) {
    __shared__ double2 red[kBlock / 64];
    const SlotRange r = ranges[blockIdx.x];
    double sr = 0.0, cr = 0.0, si = 0.0, ci = 0.0;
    for (uint64_t i = threadIdx.x; i < r.count; i += kBlock) {
        const double2 v = partials[r.off + i];
        double t = sr + v.x;
        cr += (fabs(sr) >= fabs(v.x)) ? (sr - t) + v.x : (v.x - t) + sr;
        sr = t;
        t = si + v.y;
        ci += (fabs(si) >= fabs(v.y)) ? (si - t) + v.y : (v.y - t) + si;
        si = t;
    }
    block_reduce_store(sr + cr, si + ci, red, &out[blockIdx.x]);
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void pauli_sum_kernel(const typename VecT<T>::type *a, typename VecT<T>::type *out,
                                                          const PauliTerm *terms, uint32_t nterms, uint64_t N,
                                                          double2 *partials) {
    using V = typename VecT<T>::type;
    __shared__ double2 red[kBlock / 64];
    double er = 0.0, ei = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride) {
        double sr = 0.0, si = 0.0;
        for (uint32_t t = 0; t < nterms; ++t) {
            const PauliTerm pt = terms[t];
            const V v = a[i ^ pt.x];
            const double sg = (__popcll(i & pt.yz) & 1) ? -1.0 : 1.0;
            const double cr = sg * pt.cre, ci = sg * pt.cim;
            const double vx = v.x, vy = v.y;
            sr += cr * vx - ci * vy;
            si += cr * vy + ci * vx;
        }
        if (MODE == 1) {
            V o;
            o.x = static_cast<T>(sr);
            o.y = static_cast<T>(si);
            out[i] = o;
        } else {
            const V b = a[i];
            const double bx = b.x, by = b.y;
            er += bx * sr + by * si;
            ei += bx * si - by * sr;
        }
    }
    if (MODE == 0) block_reduce_store(er, ei, red, &partials[blockIdx.x]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void norm2_kernel(const typename VecT<T>::type *a, uint64_t N, double2 *partials) {
    __shared__ double2 red[kBlock / 64];
    double s = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride) {
        const double x = a[i].x, y = a[i].y;
        s += x * x + y * y;
    }
    block_reduce_store(s, 0.0, red, &partials[blockIdx.x]);
}

// <a|b> of two states of either precision, accumulated in FP64
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void overlap_kernel(const typename VecT<TA>::type *__restrict__ a,
                                                         const typename VecT<TB>::type *__restrict__ b, uint64_t N,
                                                         double2 *partials) {
    __shared__ double2 red[kBlock / 64];
    double re = 0.0, im = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride) {
        const double ax = a[i].x, ay = a[i].y, bx = b[i].x, by = b[i].y;
        re += ax * bx + ay * by;
        im += ax * by - ay * bx;
    }
    block_reduce_store(re, im, red, &partials[blockIdx.x]);
}

// CSR operator rows over the state index (lq/utils/SparseLinAlg.hpp:42-63):
// v_i = sum_k values[k] a[indices[k]], k in [indptr[i], indptr[i+1]), in FP64.
// MODE 0: per-block partials of sum conj(a_i) v_i (partials[2b]) and
// sum |v_i|^2 (partials[2b+1].x); MODE 1: out[i] = v_i.
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void sparse_rows_kernel(const typename VecT<T>::type *__restrict__ a,
                                                             typename VecT<T>::type *__restrict__ out,
                                                             const int64_t *__restrict__ indptr,
                                                             const int64_t *__restrict__ indices,
                                                             const typename VecT<T>::type *__restrict__ values,
                                                             uint64_t nrows, double2 *partials) {
    __shared__ double2 red[kBlock / 64];
    double re = 0.0, im = 0.0, sq = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < nrows; i += stride) {
        double vr = 0.0, vi = 0.0;
        const int64_t k1 = indptr[i + 1];
        for (int64_t k = indptr[i]; k < k1; ++k) {
            const typename VecT<T>::type h = values[k], x = a[indices[k]];
            const double hx = h.x, hy = h.y, xx = x.x, xy = x.y;
            vr += hx * xx - hy * xy;
            vi += hx * xy + hy * xx;
        }
        if (MODE == 1) {
            out[i] = mk<T>(static_cast<T>(vr), static_cast<T>(vi));
        } else {
            const double ax = a[i].x, ay = a[i].y;
            re += ax * vr + ay * vi;
            im += ax * vi - ay * vr;
            sq += vr * vr + vi * vi;
        }
    }
    if (MODE == 0) {
        block_reduce_store(re, im, red, &partials[2 * static_cast<uint64_t>(blockIdx.x)]);
        __syncthreads(); // red is reused
        block_reduce_store(sq, 0.0, red, &partials[2 * static_cast<uint64_t>(blockIdx.x) + 1]);
    }
}

struct BitList {
    int32_t b[64];
};

template <typename T>
__global__ __launch_bounds__(kBlock) void probs_hist_kernel(const typename VecT<T>::type *a, uint64_t N, BitList bits,
                                                           int nw, double *partials) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *hist = reinterpret_cast<double *>(smem_raw);
    const uint32_t nb = 1u << nw;
    for (uint32_t o = threadIdx.x; o < nb; o += kBlock) hist[o] = 0.0;
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    const uint64_t nchunks = (N + stride - 1) / stride;
    for (uint64_t ch = 0; ch < nchunks; ++ch) {
        const uint64_t i = ch * stride + static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
        double p = 0.0;
        uint32_t o = 0;
        if (i < N) {
            const double x = a[i].x, y = a[i].y;
            p = x * x + y * y;
            for (int j = 0; j < nw; ++j) o |= static_cast<uint32_t>((i >> bits.b[j]) & 1ull) << (nw - 1 - j);
        }
        // wave-uniform bin (high wires): reduce across the wave, one atomic
        const uint32_t o0 = __shfl(o, 0, 64);
        if (__all(o == o0)) {
            for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
            if ((threadIdx.x & 63u) == 0) atomicAdd(&hist[o0], p);
        } else if (i < N) {
            atomicAdd(&hist[o], p);
        }
    }
    __syncthreads();
    for (uint32_t o = threadIdx.x; o < nb; o += kBlock)
        partials[static_cast<uint64_t>(blockIdx.x) * nb + o] = hist[o];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void probs_gather_kernel(const typename VecT<T>::type *a, BitList bits, int nw,
                                                             BitList rest, int nrest, double *out) {
    const uint64_t o = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (o >= (1ull << nw)) return;
    uint64_t bi = 0;
    for (int j = 0; j < nw; ++j) bi |= ((o >> (nw - 1 - j)) & 1ull) << bits.b[j];
    double s = 0.0;
    const uint64_t nr = 1ull << nrest;
    for (uint64_t r = 0; r < nr; ++r) {
        uint64_t idx = bi;
        for (int j = 0; j < nrest; ++j) idx |= ((r >> j) & 1ull) << rest.b[j];
        const double x = a[idx].x, y = a[idx].y;
        s += x * x + y * y;
    }
    out[o] = s;
}

__global__ __launch_bounds__(kBlock) void reduce_rows_kernel(const double *partials, uint32_t nrows, uint32_t ncols,
                                                            double *out) {
    const uint32_t col = blockIdx.x * kBlock + threadIdx.x;
    if (col >= ncols) return;
    double s = 0.0, cmp = 0.0;
    for (uint32_t r = 0; r < nrows; ++r) {
        const double v = partials[static_cast<uint64_t>(r) * ncols + col];
        const double t = s + v;
        cmp += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    out[col] = s + cmp;
}

// One workgroup per segment of 2^(lb + mb) amplitudes: the products over the
// low lb and the next mb bits are tables in LDS, the product over the
// segment's own bits one value; every amplitude is then two multiplies.
template <typename T>
__global__ __launch_bounds__(kBlock) void fill_product_kernel(typename VecT<T>::type *a, int n, int lb, int mb,
                                                             ProductFactors f) {
    __shared__ double2 lo[256], mid[256];
    __shared__ double2 seg;
    const uint32_t tid = threadIdx.x;
    auto prod = [&](uint64_t bits, int from, int count) {
        double2 p = make_double2(1.0, 0.0);
        for (int j = 0; j < count; ++j) {
            const double2 g = f.f[2 * (from + j) + ((bits >> j) & 1ull)];
            p = make_double2(p.x * g.x - p.y * g.y, p.x * g.y + p.y * g.x);
        }
        return p;
    };
    if (tid < (1u << lb)) lo[tid] = prod(tid, 0, lb);
    if (tid < (1u << mb)) mid[tid] = prod(tid, lb, mb);
    if (tid == 0) seg = prod(blockIdx.x, lb + mb, n - lb - mb);
    __syncthreads();
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) << (lb + mb);
    const uint32_t per = 1u << (lb + mb);
    for (uint32_t i = tid; i < per; i += kBlock) {
        const double2 m = mid[i >> lb], l = lo[i & ((1u << lb) - 1u)];
        const double2 h = make_double2(seg.x * m.x - seg.y * m.y, seg.x * m.y + seg.y * m.x);
        a[base + i] = mk<T>(static_cast<T>(h.x * l.x - h.y * l.y), static_cast<T>(h.x * l.y + h.y * l.x));
    }
}

__global__ __launch_bounds__(kBlock) void stream_copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                            uint64_t N) {
    typedef double dvec __attribute__((ext_vector_type(2)));
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock * 4;
    for (uint64_t i = (static_cast<uint64_t>(blockIdx.x) * kBlock * 4) + threadIdx.x; i < N; i += stride) {
        dvec v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * kBlock < N) v[u] = __builtin_nontemporal_load(reinterpret_cast<const dvec *>(src + i + u * kBlock));
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * kBlock < N) __builtin_nontemporal_store(v[u], reinterpret_cast<dvec *>(dst + i + u * kBlock));
    }
}

// Dense 2^k x 2^k operator on k target bits (any k), out of place: one
// thread per output amplitude sums its matrix row against the 2^k amplitudes
// of its group (targets bits[0] = local MSB). Amplitudes whose control bits
// do not match are copied unchanged. M is row-major complex double.
template <typename T>
__global__ __launch_bounds__(kBlock) void dense_wide_kernel(const typename VecT<T>::type *__restrict__ in,
                                                           typename VecT<T>::type *__restrict__ out,
                                                           const double2 *__restrict__ M, BitList bits, int k,
                                                           uint64_t tmask, uint64_t cmask, uint64_t cval, uint64_t N) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    const uint64_t d = 1ull << k;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride) {
        if ((i & cmask) != cval) {
            out[i] = in[i];
            continue;
        }
        uint64_t r = 0;
        for (int j = 0; j < k; ++j) r |= ((i >> bits.b[j]) & 1ull) << (k - 1 - j);
        const uint64_t base = i & ~tmask;
        const double2 *row = M + r * d;
        double ar = 0.0, ai = 0.0;
        for (uint64_t c = 0; c < d; ++c) {
            uint64_t idx = base; // the group's amplitude with target bits c
            for (int j = 0; j < k; ++j) idx |= ((c >> (k - 1 - j)) & 1ull) << bits.b[j];
            const typename VecT<T>::type v = in[idx];
            const double2 m = row[c];
            const double vx = v.x, vy = v.y;
            ar += m.x * vx - m.y * vy;
            ai += m.x * vy + m.y * vx;
        }
        out[i] = mk<T>(static_cast<T>(ar), static_cast<T>(ai));
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scale_kernel(typename VecT<T>::type *a, uint64_t N, double re, double im) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    const typename VecT<T>::type f = mk<T>(static_cast<T>(re), static_cast<T>(im));
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride)
        a[i] = cmul(f, a[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void collapse_kernel(typename VecT<T>::type *a, uint64_t N, int bit, int branch,
                                                         double factor) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    const T f = static_cast<T>(factor);
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride) {
        if (static_cast<int>((i >> bit) & 1ull) != branch) a[i] = mk<T>(T(0), T(0));
        else a[i] = mk<T>(a[i].x * f, a[i].y * f);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void set_subspace_kernel(typename VecT<T>::type *a, const typename VecT<T>::type *src,
                                                             BitList bits, int nw) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= (1ull << nw)) return;
    uint64_t idx = 0;
    for (int j = 0; j < nw; ++j) idx |= ((i >> (nw - 1 - j)) & 1ull) << bits.b[j];
    a[idx] = src[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void axpy_kernel(const typename VecT<T>::type *x, typename VecT<T>::type *y,
                                                     uint64_t N, double re, double im) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    const typename VecT<T>::type f = mk<T>(static_cast<T>(re), static_cast<T>(im));
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride) {
        const typename VecT<T>::type v = cmul(f, x[i]);
        y[i] = mk<T>(y[i].x + v.x, y[i].y + v.y);
    }
}

__global__ __launch_bounds__(kBlock) void c128_to_c64_kernel(const double2 *s, float2 *d, uint64_t N) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride)
        d[i] = make_float2(static_cast<float>(s[i].x), static_cast<float>(s[i].y));
}
__global__ __launch_bounds__(kBlock) void c64_to_c128_kernel(const float2 *s, double2 *d, uint64_t N) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N; i += stride)
        d[i] = make_double2(s[i].x, s[i].y);
}

uint32_t stream_blocks(uint64_t N) {
    const uint64_t b = (N + kBlock - 1) / kBlock;
    return static_cast<uint32_t>(b < 8192 ? (b ? b : 1) : 8192);
}

BitList to_bitlist(const int *bits, int nw) {
    BitList b{};
    for (int j = 0; j < nw && j < 64; ++j) b.b[j] = bits[j];
    return b;
}

} // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename T> void launch_tile_pass(const TilePassArgs &a, size_t lds_bytes, hipStream_t s) {
    const uint64_t nt = a.ntiles;
    const uint32_t gx = static_cast<uint32_t>(nt < 65536 ? nt : 65536);
    const uint32_t gy = static_cast<uint32_t>((nt + gx - 1) / gx);
    const dim3 grid(gx, gy);
    switch (a.ns) {
    case 1: hipLaunchKernelGGL((tile_pass_kernel<T, 1>), grid, dim3(kBlock), lds_bytes, s, a); break;
    case 2: hipLaunchKernelGGL((tile_pass_kernel<T, 2>), grid, dim3(kBlock), lds_bytes, s, a); break;
    case 3: hipLaunchKernelGGL((tile_pass_kernel<T, 3>), grid, dim3(kBlock), lds_bytes, s, a); break;
    default: hipLaunchKernelGGL((tile_pass_kernel<T, 4>), grid, dim3(kBlock), lds_bytes, s, a); break;
    }
}
template void launch_tile_pass<double>(const TilePassArgs &, size_t, hipStream_t);
template void launch_tile_pass<float>(const TilePassArgs &, size_t, hipStream_t);

void launch_reduce_slots(const double2 *partials, const SlotRange *ranges, uint32_t nslots, double2 *out,
                         hipStream_t s) {
    if (nslots == 0) return;
    hipLaunchKernelGGL(reduce_slots_kernel, dim3(nslots), dim3(kBlock), 0, s, partials, ranges, out);
}

template <typename T>
void launch_pauli_sum(const void *a, void *out, const PauliTerm *terms, uint32_t nterms, int n, int mode,
                      double2 *partials, uint32_t nblocks, hipStream_t s) {
    using V = typename VecT<T>::type;
    const uint64_t N = 1ull << n;
    if (mode == 0)
        hipLaunchKernelGGL((pauli_sum_kernel<T, 0>), dim3(nblocks), dim3(kBlock), 0, s, static_cast<const V *>(a),
                           static_cast<V *>(out), terms, nterms, N, partials);
    else
        hipLaunchKernelGGL((pauli_sum_kernel<T, 1>), dim3(stream_blocks(N)), dim3(kBlock), 0, s,
                           static_cast<const V *>(a), static_cast<V *>(out), terms, nterms, N, partials);
}
template void launch_pauli_sum<double>(const void *, void *, const PauliTerm *, uint32_t, int, int, double2 *, uint32_t,
                                       hipStream_t);
template void launch_pauli_sum<float>(const void *, void *, const PauliTerm *, uint32_t, int, int, double2 *, uint32_t,
                                      hipStream_t);

template <typename T> void launch_norm2(const void *a, int n, double2 *partials, uint32_t nblocks, hipStream_t s) {
    using V = typename VecT<T>::type;
    hipLaunchKernelGGL(norm2_kernel<T>, dim3(nblocks), dim3(kBlock), 0, s, static_cast<const V *>(a), 1ull << n,
                       partials);
}
template void launch_norm2<double>(const void *, int, double2 *, uint32_t, hipStream_t);
template void launch_norm2<float>(const void *, int, double2 *, uint32_t, hipStream_t);

void launch_overlap(const void *a, int dtype_a, const void *b, int dtype_b, int n, double2 *partials, uint32_t nblocks,
                    hipStream_t s) {
    const uint64_t N = 1ull << n;
#define LQ_OV(TA, TB)                                                                                          \
    hipLaunchKernelGGL((overlap_kernel<TA, TB>), dim3(nblocks), dim3(kBlock), 0, s,                          \
                       static_cast<const typename VecT<TA>::type *>(a), static_cast<const typename VecT<TB>::type *>(b), \
                       N, partials)
    if (dtype_a && dtype_b) LQ_OV(double, double);
    else if (dtype_a) LQ_OV(double, float);
    else if (dtype_b) LQ_OV(float, double);
    else LQ_OV(float, float);
#undef LQ_OV
}

template <typename T>
void launch_sparse_rows(const void *a, void *out, const int64_t *indptr, const int64_t *indices, const void *values,
                        uint64_t nrows, int mode, double2 *partials, uint32_t nblocks, hipStream_t s) {
    using V = typename VecT<T>::type;
    if (mode == 0)
        hipLaunchKernelGGL((sparse_rows_kernel<T, 0>), dim3(nblocks), dim3(kBlock), 0, s, static_cast<const V *>(a),
                           nullptr, indptr, indices, static_cast<const V *>(values), nrows, partials);
    else
        hipLaunchKernelGGL((sparse_rows_kernel<T, 1>), dim3(nblocks), dim3(kBlock), 0, s, static_cast<const V *>(a),
                           static_cast<V *>(out), indptr, indices, static_cast<const V *>(values), nrows, nullptr);
}
template void launch_sparse_rows<double>(const void *, void *, const int64_t *, const int64_t *, const void *, uint64_t,
                                         int, double2 *, uint32_t, hipStream_t);
template void launch_sparse_rows<float>(const void *, void *, const int64_t *, const int64_t *, const void *, uint64_t,
                                        int, double2 *, uint32_t, hipStream_t);

template <typename T>
void launch_probs_hist(const void *a, int n, const int *bits, int nw, double *partials, uint32_t nblocks,
                       hipStream_t s) {
    using V = typename VecT<T>::type;
    const size_t lds = sizeof(double) << nw;
    hipLaunchKernelGGL(probs_hist_kernel<T>, dim3(nblocks), dim3(kBlock), lds, s, static_cast<const V *>(a), 1ull << n,
                       to_bitlist(bits, nw), nw, partials);
}
template void launch_probs_hist<double>(const void *, int, const int *, int, double *, uint32_t, hipStream_t);
template void launch_probs_hist<float>(const void *, int, const int *, int, double *, uint32_t, hipStream_t);

template <typename T>
void launch_probs_gather(const void *a, int n, const int *bits, int nw, double *out, hipStream_t s) {
    using V = typename VecT<T>::type;
    int rest[64];
    int nr = 0;
    for (int b = 0; b < n; ++b) {
        bool in = false;
        for (int j = 0; j < nw; ++j) in = in || bits[j] == b;
        if (!in) rest[nr++] = b;
    }
    const uint64_t no = 1ull << nw;
    hipLaunchKernelGGL(probs_gather_kernel<T>, dim3(static_cast<uint32_t>((no + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, static_cast<const V *>(a), to_bitlist(bits, nw), nw, to_bitlist(rest, nr), nr, out);
}
template void launch_probs_gather<double>(const void *, int, const int *, int, double *, hipStream_t);
template void launch_probs_gather<float>(const void *, int, const int *, int, double *, hipStream_t);

void launch_reduce_rows(const double *partials, uint32_t nrows, uint32_t ncols, double *out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_rows_kernel, dim3((ncols + kBlock - 1) / kBlock), dim3(kBlock), 0, s, partials, nrows,
                       ncols, out);
}

template <typename T> void launch_fill_product(void *a, int n, const ProductFactors &f, hipStream_t s) {
    using V = typename VecT<T>::type;
    const int lb = n < 8 ? n : 8;
    const int mb = (n - lb) < 8 ? (n - lb) : 8;
    const uint32_t segs = 1u << (n - lb - mb); // n <= 40: at most 2^24 segments
    hipLaunchKernelGGL(fill_product_kernel<T>, dim3(segs), dim3(kBlock), 0, s, static_cast<V *>(a), n, lb, mb, f);
}
template void launch_fill_product<double>(void *, int, const ProductFactors &, hipStream_t);
template void launch_fill_product<float>(void *, int, const ProductFactors &, hipStream_t);

void launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s) {
    const uint64_t N = bytes / 16;
    const uint64_t per = static_cast<uint64_t>(kBlock) * 4;
    const uint64_t b = (N + per - 1) / per;
    const uint32_t blocks = static_cast<uint32_t>(b < 16384 ? (b ? b : 1) : 16384);
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(kBlock), 0, s, static_cast<const double2 *>(src),
                       static_cast<double2 *>(dst), N);
}

template <typename T>
void launch_dense_wide(const void *in, void *out, const double2 *M, const int *bits, int k, uint64_t cmask,
                       uint64_t cval, int n, hipStream_t s) {
    using V = typename VecT<T>::type;
    const uint64_t N = 1ull << n;
    uint64_t tmask = 0;
    for (int j = 0; j < k; ++j) tmask |= 1ull << bits[j];
    hipLaunchKernelGGL(dense_wide_kernel<T>, dim3(stream_blocks(N)), dim3(kBlock), 0, s, static_cast<const V *>(in),
                       static_cast<V *>(out), M, to_bitlist(bits, k), k, tmask, cmask, cval, N);
}
template void launch_dense_wide<double>(const void *, void *, const double2 *, const int *, int, uint64_t, uint64_t,
                                        int, hipStream_t);
template void launch_dense_wide<float>(const void *, void *, const double2 *, const int *, int, uint64_t, uint64_t,
                                       int, hipStream_t);

template <typename T> void launch_scale(void *a, int n, double re, double im, hipStream_t s) {
    using V = typename VecT<T>::type;
    const uint64_t N = 1ull << n;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(stream_blocks(N)), dim3(kBlock), 0, s, static_cast<V *>(a), N, re, im);
}
template void launch_scale<double>(void *, int, double, double, hipStream_t);
template void launch_scale<float>(void *, int, double, double, hipStream_t);

template <typename T> void launch_collapse(void *a, int n, int bit, int branch, double factor, hipStream_t s) {
    using V = typename VecT<T>::type;
    const uint64_t N = 1ull << n;
    hipLaunchKernelGGL(collapse_kernel<T>, dim3(stream_blocks(N)), dim3(kBlock), 0, s, static_cast<V *>(a), N, bit,
                       branch, factor);
}
template void launch_collapse<double>(void *, int, int, int, double, hipStream_t);
template void launch_collapse<float>(void *, int, int, int, double, hipStream_t);

template <typename T>
void launch_set_subspace(void *a, int n, const void *src, const int *bits, int nw, hipStream_t s) {
    using V = typename VecT<T>::type;
    (void)n;
    const uint64_t no = 1ull << nw;
    hipLaunchKernelGGL(set_subspace_kernel<T>, dim3(static_cast<uint32_t>((no + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, static_cast<V *>(a), static_cast<const V *>(src), to_bitlist(bits, nw), nw);
}
template void launch_set_subspace<double>(void *, int, const void *, const int *, int, hipStream_t);
template void launch_set_subspace<float>(void *, int, const void *, const int *, int, hipStream_t);

template <typename T> void launch_axpy(const void *x, void *y, double re, double im, int n, hipStream_t s) {
    using V = typename VecT<T>::type;
    const uint64_t N = 1ull << n;
    hipLaunchKernelGGL(axpy_kernel<T>, dim3(stream_blocks(N)), dim3(kBlock), 0, s, static_cast<const V *>(x),
                       static_cast<V *>(y), N, re, im);
}
template void launch_axpy<double>(const void *, void *, double, double, int, hipStream_t);
template void launch_axpy<float>(const void *, void *, double, double, int, hipStream_t);

void launch_c128_to_c64(const void *src, void *dst, uint64_t count, hipStream_t s) {
    hipLaunchKernelGGL(c128_to_c64_kernel, dim3(stream_blocks(count)), dim3(kBlock), 0, s,
                       static_cast<const double2 *>(src), static_cast<float2 *>(dst), count);
}
void launch_c64_to_c128(const void *src, void *dst, uint64_t count, hipStream_t s) {
    hipLaunchKernelGGL(c64_to_c128_kernel, dim3(stream_blocks(count)), dim3(kBlock), 0, s,
                       static_cast<const float2 *>(src), static_cast<double2 *>(dst), count);
}

} // namespace lqamd
