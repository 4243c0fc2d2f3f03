// device_common.h -- device helpers shared by the HIP translation units.
//
// All device code is compiled with -ffp-contract=off: entry evaluation must
// round exactly like the FP64 CPU restatement, and every fused multiply-add
// below is written out explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
};

// A_ij for global tree positions i, j (user_kernel.hpp:50-58,105-113,171-177
// and points_dt.hpp:31-39): squared distance by separate multiply and add.
__device__ __forceinline__ double entry(const KernelDev &k, int64_t i, int64_t j) {
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
    double f;
    if (k.family == 1) f = 1.0 / r;
    else if (k.family == 0) f = log(r);
    else f = exp(-r);
    double v = k.scale * f;
    if (k.W) v = v * k.W[j - k.pos0];
    return v;
}

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
