// aca.hip -- the library's instance of the entry kernels (built-in kernel
// families) and the assembly kernels that evaluate no entries: staged-factor
// copies, the FP32 operator's zero self-block diagonal, FP64 -> FP32.
#include "../../include/hodlrnd/device/entry_kernels.hpp"

namespace hodlr {

const EntryLaunchers &builtin_entry_launchers() { return *entry_launchers<FamilyFn>(); }

// staged factors into their panels: the same destinations as the fill pass
// (U rows into the row node's leaf-blocked panel, V rows into the column
// node's tile-major panel); one workgroup per piece of a block
__global__ __launch_bounds__(256) void stage_copy_kernel(const StageCopy *__restrict__ items,
                                                         const double *__restrict__ buf,
                                                         const int64_t *__restrict__ leaf_start,
                                                         const int32_t *__restrict__ leaf_cnt,
                                                         double *__restrict__ UP, double *__restrict__ VP) {
    const StageCopy C = items[blockIdx.x];
    const int64_t nu = (int64_t)C.r * C.m;
    for (int64_t e = C.e0 + threadIdx.x; e < C.e1; e += blockDim.x) {
        const double v = buf[C.src + e];
        if (e < nu) {
            const int q = (int)(e / C.m), i = (int)(e % C.m);
            const int64_t p = C.ord[C.row0l + i];
            const int64_t ls = leaf_start[p];
            const int64_t nc = leaf_cnt[p];
            UP[C.up_base + (int64_t)C.R * (ls - C.row0) + (p - ls) + (int64_t)(C.coff + q) * nc] = v;
        } else {
            const int64_t ev = e - nu;
            const int q = (int)(ev / C.n), j = (int)(ev % C.n);
            const int64_t row = C.ord[C.col0l + j] - C.col0;
            const int col = C.voff + q;
            const int tile = col / kProjTile;
            const int tw = min(kProjTile, C.T - tile * kProjTile);
            VP[C.vp_base + (int64_t)tile * kProjTile * C.n + row * tw + (col - tile * kProjTile)] = v;
        }
    }
}

// staged pivot positions of skeleton blocks into the operator's pivot
// slots; one wave per block
__global__ __launch_bounds__(64) void stage_pivots_kernel(const PivotCopy *__restrict__ items,
                                                          const double *__restrict__ buf,
                                                          int32_t *__restrict__ SK) {
    const PivotCopy C = items[blockIdx.x];
    const int32_t *pv = reinterpret_cast<const int32_t *>(buf + C.src);
    for (int q = threadIdx.x; q < C.r; q += blockDim.x) {
        SK[C.sk_row + q] = pv[q];
        SK[C.sk_col + q] = pv[C.r + q];
    }
}

hipError_t launch_stage_pivots(const PivotCopy *items, int64_t n, const double *buf, int32_t *SK, hipStream_t s) {
    if (n == 0) return hipSuccess;
    stage_pivots_kernel<<<dim3((unsigned)n), 64, 0, s>>>(items, buf, SK);
    return hipGetLastError();
}

// zero the diagonal of every leaf's self block (the first nc columns of its
// near-field panel); the FP32 operator then holds only off-diagonal entries
__global__ void zero_self_diag_kernel(const NearLeaf *__restrict__ leaves, int64_t nl, float *__restrict__ NP) {
    for (int64_t l = blockIdx.x; l < nl; l += gridDim.x) {
        const NearLeaf L = leaves[l];
        for (int i = threadIdx.x; i < L.nc; i += blockDim.x) NP[L.np_base + (int64_t)i * L.nc + i] = 0.0f;
    }
}

hipError_t launch_zero_self_diag(const NearLeaf *l, int64_t nl, float *NP, hipStream_t s) {
    if (nl == 0) return hipSuccess;
    zero_self_diag_kernel<<<dim3((unsigned)std::min<int64_t>(nl, 65536)), 64, 0, s>>>(l, nl, NP);
    return hipGetLastError();
}

__global__ void convert_f64_f32_kernel(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

hipError_t launch_stage_copy(const StageCopy *items, int64_t n, const double *buf, const int64_t *leaf_start,
                             const int32_t *leaf_cnt, double *UP, double *VP, hipStream_t s) {
    if (n == 0) return hipSuccess;
    stage_copy_kernel<<<dim3((unsigned)n), 256, 0, s>>>(items, buf, leaf_start, leaf_cnt, UP, VP);
    return hipGetLastError();
}

hipError_t launch_convert(const double *a, float *b, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    convert_f64_f32_kernel<<<dim3((unsigned)g), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

}  // namespace hodlr
