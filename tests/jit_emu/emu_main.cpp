// Emulation runtime linked into each emulated kernel: a workgroup is 256 host
// threads sharing one barrier; blocks of the grid run one after another.
#include "emu_rt.hpp"
thread_local Dim3 threadIdx;
Dim3 blockIdx, gridDim;
static std::barrier<> *g_bar;
static double g_shfl[256];
void __syncthreads() { g_bar->arrive_and_wait(); }
double __shfl_xor(double v, int off, int) {
    const unsigned t = threadIdx.x;
    g_shfl[t] = v;
    g_bar->arrive_and_wait();
    const double r = g_shfl[t ^ static_cast<unsigned>(off)];
    g_bar->arrive_and_wait();
    return r;
}
extern "C" void lqpass(void *, void *, const double2 *, double2 *, unsigned long long, void *, void *);
extern "C" void emu_launch(void *sv0, void *sv1, const double2 *coef, double2 *partials, unsigned long long ntiles,
                           unsigned grid, void *sv2, void *sv3) {
    gridDim = {grid, 1, 1};
    for (unsigned b = 0; b < grid; ++b) {
        blockIdx = {b, 0, 0};
        std::barrier<> bar(256);
        g_bar = &bar;
        std::vector<std::thread> th;
        for (unsigned t = 0; t < 256; ++t)
            th.emplace_back([=] {
                threadIdx = {t, 0, 0};
                lqpass(sv0, sv1, coef, partials, ntiles, sv2, sv3);
            });
        for (auto &x : th) x.join();
    }
}
