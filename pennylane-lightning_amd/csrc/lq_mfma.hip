// Dense 32x32 complex operator on the five lowest index bits of a complex64
// state, on the matrix cores (v_mfma_f32_16x16x4_f32).
//
// With the five targets on index bits 0..4, the state is a 32 x 2^(n-5)
// complex matrix whose column j holds the 32 amplitudes j*32 .. j*32+31, and
// the operator is out = U * S. In real form this is one GEMM per column
// block: [Re out; Im out] (64 x N) = W (64 x 64) x [Re S; Im S] (64 x N) with
// W = [[Re U, -Im U], [Im U, Re U]].
//
// Persistent workgroups (4 waves) walk tiles of 128 columns (32 KiB of
// complex64): the next tile's 16-byte loads are in flight while the current
// one runs. A tile goes through LDS as 64 real rows per column (Re 0..31,
// Im 0..31), each wave multiplies its 32 columns as 4 x 2 blocks of 16x16
// over 16 K-steps of 4 (W from LDS, loaded once per workgroup), and the
// results go back through LDS as interleaved complex for coalesced 16-byte
// stores. In place is allowed: a tile is read completely before it is
// written, and no two workgroups share a tile.
//
// Matrix-core layout of v_mfma_f32_16x16x4_f32 (CDNA3/CDNA4 ISA): A (16x4)
// lane l = A[l%16][l/16]; B (4x16) lane l = B[l/16][l%16]; D (16x16) in 4
// registers, register v of lane l = D[4*(l/16) + v][l%16]. The K index of
// step kk for lane l is 16*(l/16) + kk (any order of K works when A and B
// use the same one), so a lane reads 16 consecutive rows of its column.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lq_kernels.hpp"

namespace lqamd {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCols = 128; // columns (groups of 32 amplitudes) per tile
constexpr int kSX = 68;    // floats per column in LDS (64 real rows + 4: the 16 columns of a quarter-wave's
                           // 16-byte reads fall on 64 distinct banks)
constexpr int kSW = 68;    // floats per W row in LDS

__global__ __launch_bounds__(256) void dense5_c64_mfma_kernel(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                              const float *__restrict__ W, uint64_t ntiles) {
    __shared__ __align__(16) float xs[kCols * kSX]; // a tile's real rows, then its results (complex)
    __shared__ __align__(16) float ws[64 * kSW];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t j = lane & 15u, kq = lane >> 4;
    uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;

    for (uint32_t i = tid; i < 64u * 64u; i += 256u) ws[(i >> 6) * kSW + (i & 63u)] = W[i];

    // thread t holds amplitudes 2t + 512*it and the next one: column c = idx / 32, row r = idx % 32
    f32x4 pre[8];
    {
        const f32x4 *src = reinterpret_cast<const f32x4 *>(in + tile * (kCols * 32ull));
#pragma unroll
        for (int it = 0; it < 8; ++it) pre[it] = __builtin_nontemporal_load(src + tid + 256u * it);
    }
    while (true) {
        __syncthreads(); // ws written / the previous tile's results read
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const uint32_t idx = 2u * tid + 512u * it;
            const uint32_t c = idx >> 5, r = idx & 31u;
            *reinterpret_cast<float2 *>(&xs[c * kSX + r]) = make_float2(pre[it].x, pre[it].z);
            *reinterpret_cast<float2 *>(&xs[c * kSX + 32u + r]) = make_float2(pre[it].y, pre[it].w);
        }
        __syncthreads();
        const uint64_t next = tile + gridDim.x;
        if (next < ntiles) {
            const f32x4 *src = reinterpret_cast<const f32x4 *>(in + next * (kCols * 32ull));
#pragma unroll
            for (int it = 0; it < 8; ++it) pre[it] = __builtin_nontemporal_load(src + tid + 256u * it);
        }

        f32x4 acc[4][2];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        // K rows 16kq + 4m .. +3 of the lane's column and W rows as 16-byte
        // LDS reads (a quarter-wave's 16 columns cover all 64 banks)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t k = 16u * kq + 4u * m;
            f32x4 b[2], a[4];
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
                b[cb] = *reinterpret_cast<const f32x4 *>(&xs[(wave * 32u + 16u * cb + j) * kSX + k]);
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) a[rb] = *reinterpret_cast<const f32x4 *>(&ws[(16u * rb + j) * kSW + k]);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb)
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb][t], b[cb][t], acc[rb][cb], 0, 0, 0);
        }
        __syncthreads(); // every wave has read its columns

        // rows 16rb + 4kq + v: real parts for rb 0, 1, imaginary parts of
        // the same amplitudes for rb 2, 3
        float2 *os = reinterpret_cast<float2 *>(xs);
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            const uint32_t col = wave * 32u + 16u * cb + j;
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                const uint32_t a0 = 16u * rb + 4u * kq;
                f32x4 lo, hi;
                lo.x = acc[rb][cb][0];
                lo.y = acc[rb + 2][cb][0];
                lo.z = acc[rb][cb][1];
                lo.w = acc[rb + 2][cb][1];
                hi.x = acc[rb][cb][2];
                hi.y = acc[rb + 2][cb][2];
                hi.z = acc[rb][cb][3];
                hi.w = acc[rb + 2][cb][3];
                *reinterpret_cast<f32x4 *>(&os[col * (kSX / 2) + a0]) = lo;
                *reinterpret_cast<f32x4 *>(&os[col * (kSX / 2) + a0 + 2u]) = hi;
            }
        }
        __syncthreads();
        f32x4 *dst = reinterpret_cast<f32x4 *>(out + tile * (kCols * 32ull));
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const uint32_t idx = 2u * tid + 512u * it;
            const uint32_t c = idx >> 5, r = idx & 31u;
            __builtin_nontemporal_store(*reinterpret_cast<const f32x4 *>(&os[c * (kSX / 2) + r]), dst + tid + 256u * it);
        }
        if (next >= ntiles) break;
        tile = next;
    }
}

} // namespace

void launch_dense5_c64_mfma(const void *in, void *out, const float *W, int n, hipStream_t s) {
    // n >= 12: whole tiles of 128 columns x 32 amplitudes; persistent
    // workgroups, three per CU (50 KiB of LDS each)
    const uint64_t ntiles = 1ull << (n - 12);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || cus < 1)
        cus = 256;
    const uint64_t slots = static_cast<uint64_t>(cus) * 3u;
    const uint32_t grid = static_cast<uint32_t>(ntiles < slots ? ntiles : slots);
    hipLaunchKernelGGL(dense5_c64_mfma_kernel, dim3(grid), dim3(256), 0, s, static_cast<const float2 *>(in),
                       static_cast<float2 *>(out), W, ntiles);
}

} // namespace lqamd
