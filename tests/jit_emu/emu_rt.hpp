// CPU emulation of the device environment the generated pass kernels use
// (test infrastructure: vector types, thread/block indices, barriers, wave
// shuffles, nontemporal accesses as plain ones).
#pragma once
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>
struct double2 { double x, y; };
struct float2 { float x, y; };
inline double2 make_double2(double x, double y) { return {x, y}; }
#define __device__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
struct Dim3 { unsigned x, y, z; };
extern thread_local Dim3 threadIdx;
extern Dim3 blockIdx, gridDim;
void __syncthreads();
double __shfl_xor(double v, int off, int w);
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
// LDS of the size the kernel declares, with every index checked
template <typename T, unsigned long N> struct LdsArr {
    T d[N];
    T &operator[](unsigned long i) {
        if (i >= N) {
            std::fprintf(stderr, "LDS index %lu outside %lu entries\n", i, N);
            std::abort();
        }
        return d[i];
    }
};
