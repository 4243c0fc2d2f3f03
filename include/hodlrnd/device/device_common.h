// device_common.h -- device helpers shared by the HIP translation units.
//
// All device code is compiled with -ffp-contract=off: entry evaluation must
// round exactly like the FP64 CPU restatement, and every fused multiply-add
// below is written out explicitly.  Caller code that instantiates the entry
// kernels for its own Kernel_Fun includes this header as well.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "aca_layout.h"

namespace hodlr {

constexpr int kWave = 64;
constexpr int kMaxDimDev = 8;  // largest NDIM

// One problem's kernel.  Positions are global over a batch of problems;
// this problem's points occupy [pos0, pos0 + n) and X, W hold them in tree
// order starting at pos0.
struct KernelDev {
    int family;   // HODLR_KERNEL_*
    int ndim;
    double scale;
    double diag;
    const double *X;   // tree order, n x ndim
    const double *W;   // tree order weights or nullptr
    int64_t pos0;
    const void *obj;   // device copy of a caller's kernel object (custom family), else nullptr
};

// f(r) of the built-in families (user_kernel.hpp:46-48,101-103,167-169)
struct FamilyFn {
    __device__ __forceinline__ static double f(const KernelDev &k, double r) {
        if (k.family == 1) return 1.0 / r;
        if (k.family == 0) return log(r);
        return exp(-r);
    }
};

// A_ij for global tree positions i, j (user_kernel.hpp:50-58,105-113,171-177
// and points_dt.hpp:31-39): squared distance by separate multiply and add,
// then diag (i == j) or scale * f(r) [* w_j]; Fn supplies f.
template <class Fn>
__device__ __forceinline__ double entry_t(const KernelDev &k, int64_t i, int64_t j) {
    if (i == j) return k.diag;
    const double *a = k.X + (i - k.pos0) * k.ndim;
    const double *b = k.X + (j - k.pos0) * k.ndim;
    double s = 0.0;
    for (int d = 0; d < k.ndim; d++) {
        double t = a[d] - b[d];
        double t2 = t * t;
        s = s + t2;
    }
    double r = sqrt(s);
    double f = Fn::f(k, r);
    double v = k.scale * f;
    if (k.W) v = v * k.W[j - k.pos0];
    return v;
}

__device__ __forceinline__ double entry(const KernelDev &k, int64_t i, int64_t j) { return entry_t<FamilyFn>(k, i, j); }

// layout fingerprint of the descriptors passed to the launchers: the library
// refuses a table built against other layouts
constexpr uint32_t kEntryAbiVersion = 2;
constexpr uint32_t entry_abi_layout() {
    return (uint32_t)(sizeof(AcaBlock) | sizeof(FillDest) << 8 | sizeof(DenseBlock) << 16 | sizeof(NearLeaf) << 24) ^
           (uint32_t)(sizeof(KernelDev) * 2654435761u) ^ (uint32_t)(sizeof(AcaStage) << 4);
}

// the launchers of one entry function
struct EntryLaunchers {
    uint32_t version;  // kEntryAbiVersion
    uint32_t layout;   // entry_abi_layout()
    // ACA pass over nblocks blocks; bs selects the form: 8/16/32 = one wave
    // with that rank capacity, 64/256/1024 = workgroup of that size
    hipError_t (*aca)(int bs, const AcaBlock *blocks, int64_t nblocks, const KernelDev *K, double tol,
                      double *pool, int32_t *rank, int32_t *flags, int fill, const FillDest *dest,
                      const int64_t *leaf_start, const int32_t *leaf_cnt, double *UP, double *VP, int32_t *SK,
                      const AcaStage &S, hipStream_t s);
    // dense-fallback blocks into their row node's U panel
    hipError_t (*dense_fill)(const DenseBlock *b, int64_t nb, const KernelDev *K, const int64_t *leaf_start,
                             const int32_t *leaf_cnt, double *UP, hipStream_t s);
    // every leaf's near-field panel
    hipError_t (*near_fill)(const NearLeaf *l, int64_t nl, const int64_t *colsrc, const KernelDev *K, double *NP,
                            hipStream_t s);
    // direct-sum rows of the dense operator
    hipError_t (*dense_rows)(const int64_t *rows, const int32_t *rprob, int64_t nrows, const KernelDev *K,
                             const int64_t *pos_base, const double *x, double *y, hipStream_t s);
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// XCD-aware remap of a 1-D grid: consecutive logical work items land on the
// same XCD (the hardware deals workgroups round-robin over 8 XCDs), so work
// items that share inputs share an L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblocks) {
    const int64_t nx = 8;
    const int64_t per = (nblocks + nx - 1) / nx;
    const int64_t x = bid % nx, q = bid / nx;
    const int64_t full = nblocks % nx == 0 ? nx : nblocks % nx;  // XCDs with `per` items
    int64_t base = x < full ? x * per : full * per + (x - full) * (per - 1);
    return base + q;
}

}  // namespace hodlr
