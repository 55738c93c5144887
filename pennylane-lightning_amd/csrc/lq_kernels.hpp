// Launch interface of the HIP kernels (lq_kernels.hip). Host code calls these
// launchers; every launch goes on the caller's stream and returns without
// synchronising.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lq_tileop.hpp"

namespace lqamd {

struct TilePassArgs {
    void *sv[kMaxStates];     // state base pointers (complex<T>)
    const TileOp *ops;        // device program for this pass
    const double2 *coef;      // device coefficient buffer (complex double)
    double2 *partials;        // dot partial sums: [slot][tile], slots local to the pass
    uint32_t nops;
    uint32_t load_mask;       // states loaded into LDS
    uint32_t store_mask;      // states written back
    uint32_t ns;              // number of state slots in LDS
    uint64_t ntiles;
    TileGeom g;
};

struct PauliTerm {
    uint64_t x;   // flipped bits
    uint64_t yz;  // bits contributing (-1)^bit (Y and Z factors)
    double cre;   // coefficient times (-i)^ny, real part
    double cim;
};

struct SlotRange {
    uint64_t off;   // offset (complex entries) of the slot's partials
    uint64_t count; // number of partials
};

template <typename T> void launch_tile_pass(const TilePassArgs &a, size_t lds_bytes, hipStream_t s);

// Register-resident tile pass (lq_kernels_reg.hip): kt = 8 + RB, RB in {3, 4},
// one or two states.
struct TileRegArgs {
    void *sv[kMaxStates]; // the generic kernel takes two; specialised passes up to four
    const TileOp *ops;
    const TileStage *stages;
    const double2 *coef;
    double2 *partials;        // [slot][tile]
    uint32_t nstages;
    uint32_t nslots;          // dot slots of the pass (<= kMaxPassDots)
    uint32_t load_mask;
    uint32_t store_mask;
    uint32_t end_transpose;   // last stage's mapping differs from the HBM mapping
    uint32_t ns;
    uint32_t wide;            // the pass has two-qubit dense ops or LDS stages
    uint64_t ntiles;
    TileGeom g;
};
template <typename T> void launch_tile_reg(const TileRegArgs &a, size_t lds_bytes, hipStream_t s);

// out[slot] = sum of partials (Neumaier-compensated, fixed order)
void launch_reduce_slots(const double2 *partials, const SlotRange *ranges, uint32_t nslots,
                         double2 *out, hipStream_t s);

// mode 0: partials[block] = sum_t c_t sum_i conj(a[i]) P_t a[i]   (expectation)
// mode 1: out[i] = sum_t c_t (P_t a)[i]                             (apply)
template <typename T>
void launch_pauli_sum(const void *a, void *out, const PauliTerm *terms, uint32_t nterms, int n,
                      int mode, double2 *partials, uint32_t nblocks, hipStream_t s);

// sum_i |a_i|^2 into per-block partials
template <typename T>
void launch_norm2(const void *a, int n, double2 *partials, uint32_t nblocks, hipStream_t s);
// partials[b] for b < nblocks: per-block sums of conj(a) b (FP64)
void launch_overlap(const void *a, int dtype_a, const void *b, int dtype_b, int n, double2 *partials, uint32_t nblocks,
                    hipStream_t s);

// CSR operator over the state index (values in the state's precision):
// mode 0: partials[2b] = block sums of conj(a_i) (Ha)_i, partials[2b+1].x of
// |(Ha)_i|^2 (FP64); mode 1: out = H a.
template <typename T>
void launch_sparse_rows(const void *a, void *out, const int64_t *indptr, const int64_t *indices, const void *values,
                        uint64_t nrows, int mode, double2 *partials, uint32_t nblocks, hipStream_t s);
// marginal probabilities: out[o] over the wires given as physical bits
// (bits[0] is the output MSB). Small nw: per-block LDS histograms into
// partials[block][o]; large nw: one output per thread.
template <typename T>
void launch_probs_hist(const void *a, int n, const int *bits, int nw, double *partials,
                       uint32_t nblocks, hipStream_t s);
template <typename T>
void launch_probs_gather(const void *a, int n, const int *bits, int nw, double *out, hipStream_t s);
void launch_reduce_rows(const double *partials, uint32_t nrows, uint32_t ncols, double *out,
                        hipStream_t s);

// a[i] = prod_q f[2q + bit q of i] (FP64 products), q < n <= 40
struct ProductFactors {
    double2 f[80];
};
template <typename T> void launch_fill_product(void *a, int n, const ProductFactors &f, hipStream_t s);
// dst = src, bytes a multiple of 16 (nontemporal 16-byte loads and stores)
void launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s);
// out = M in for a dense 2^k x 2^k operator on physical bits[0..k) (bits[0]
// the local MSB; M row-major complex double on the device), amplitudes whose
// controls do not match copied unchanged; in != out
template <typename T>
void launch_dense_wide(const void *in, void *out, const double2 *M, const int *bits, int k, uint64_t cmask,
                       uint64_t cval, int n, hipStream_t s);
// complex64, in place allowed: out = U in for a dense 32x32 operator on index
// bits 0..4 (bit 4 the local MSB), given as the row-major real 64x64 matrix
// W = [[Re U, -Im U], [Im U, Re U]] on the device; matrix cores; n >= 12
void launch_dense5_c64_mfma(const void *in, void *out, const float *W, int n, hipStream_t s);
// Probabilities of 1..kNoAllocProbsMaxWires wires of a caller-owned device
// state into device memory `out` (2^nw doubles) on stream s, with the
// partial histograms in a fixed device array (no allocation; calls that
// overlap on one GPU share it). probs_noalloc_init takes the array's address
// on the current device (the first probs_noalloc call does it otherwise).
constexpr int kNoAllocProbsMaxWires = 10;
void probs_noalloc_init();
void probs_noalloc(const void *state, int n, int dtype, const uint32_t *wires, int nw, double *out, hipStream_t s);
template <typename T> void launch_scale(void *a, int n, double re, double im, hipStream_t s);
// zero amplitudes whose bit `bit` != branch, then scale the rest by `factor`
template <typename T>
void launch_collapse(void *a, int n, int bit, int branch, double factor, hipStream_t s);
// a[dst_index(i)] = src[i] for i < 2^nw with all other bits zero; the rest is
// left as it is (lq/StateVectorLQubit.hpp:1005-1036)
template <typename T>
void launch_set_subspace(void *a, int n, const void *src, const int *bits, int nw, hipStream_t s);
// y += (re + i im) * x
template <typename T> void launch_axpy(const void *x, void *y, double re, double im, int n, hipStream_t s);
// convert between complex128 and complex64 device buffers
void launch_c128_to_c64(const void *src, void *dst, uint64_t count, hipStream_t s);
void launch_c64_to_c128(const void *src, void *dst, uint64_t count, hipStream_t s);

} // namespace lqamd
