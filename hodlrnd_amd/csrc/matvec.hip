// matvec.hip -- the fast HODLR matvec (gfx950), FP64 or FP32, 1..32 RHS.
//
// The reference applies every node's blocks one GEMV at a time
// (HDD_tree.hpp:86-107, HDD_node.hpp:223-254, LowRank_matrix.hpp:40-63).
// Here the operator is stored in three streaming arenas and applied by
// three launches, each reading its arena exactly once:
//   proj:   t = V^T x for every compressed block.  The V factors of all
//           blocks that share a column node b form one panel (n_b x T_b,
//           T_b padded to a multiple of 4) stored as 256-column tiles, each
//           a contiguous row-major block; a workgroup streams one tile, each
//           lane four adjacent columns (one vector load per row), and
//           x_b[j] is a broadcast.  Tall tiles are cut
//           into row chunks whose partial sums are added by a second small
//           pass (deterministic order).
//   tcopy:  dense-fallback blocks take t = x_b (their U panel holds K).
//   panel:  y = sum over segments of M x_seg, one workgroup per leaf row
//           tile.  A leaf's segments are its near-field blocks (self and
//           neighbours) and one slice of the U panel of each ancestor, all
//           column-major with leading dimension = leaf size, so G = 64/h
//           column groups of h lanes read G*h contiguous elements per step.
//           Each leaf row is written once (no atomics).
#include "aca_layout.h"
#include "device_common.h"

namespace hodlr {

// 4 consecutive elements, 16-byte (float) or 32-byte (double) aligned
template <typename T>
struct Vec4 {
    T v[4];
};
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ Vec4<T> load4(const T *p) {
    Vec4<T> r;
    if constexpr (sizeof(T) == 4) {
        const f32x4 a = *reinterpret_cast<const f32x4 *>(p);
        r.v[0] = a[0]; r.v[1] = a[1]; r.v[2] = a[2]; r.v[3] = a[3];
    } else {
        const f64x2 a = *reinterpret_cast<const f64x2 *>(p);
        const f64x2 b = *reinterpret_cast<const f64x2 *>(p + 2);
        r.v[0] = a[0]; r.v[1] = a[1]; r.v[2] = b[0]; r.v[3] = b[1];
    }
    return r;
}
template <typename T>
__device__ __forceinline__ T load_stream(const T *p) {
    return *p;
}

// One workgroup per (V panel, 256-column tile, row chunk).  Lane l owns
// columns 4l..4l+3 of the tile (one 16/32-byte load per row); the 4 waves
// take rows w, w+4, ... and are summed through LDS at the end.  Panels are
// padded to a multiple of 4 columns, so every row of a tile starts aligned.
template <typename T, int K>
__global__ __launch_bounds__(256) void proj_kernel(const ProjItem *__restrict__ items, int64_t nitems,
                                                   const T *__restrict__ VP, const T *__restrict__ x,
                                                   T *__restrict__ t, T *__restrict__ part,
                                                   const int64_t *__restrict__ tidx) {
    constexpr int U = 4;  // rows in flight per wave per step
    __shared__ T red[kProjTile][K];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    // items are sorted by size, largest first: the hardware's round-robin
    // dispatch over XCDs then balances bytes per XCD
    const ProjItem P = items[blockIdx.x];
    const int c = lane * 4;
    const bool live = c < P.w;
    T acc[4][K];
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[e][k] = T(0);
    const T *base = VP + P.vp + c;
    const T *xv = x + P.x0 * K;
    const int64_t ld = P.ld;
    if (live) {
        int j = wid;
        for (; j + 4 * (U - 1) < P.rows; j += 4 * U) {
            Vec4<T> a[U];
#pragma unroll
            for (int u = 0; u < U; u++) a[u] = load4(base + (int64_t)(j + 4 * u) * ld);
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const T xj = xv[(int64_t)(j + 4 * u) * K + k];
#pragma unroll
                    for (int e = 0; e < 4; e++) acc[e][k] = fma(a[u].v[e], xj, acc[e][k]);
                }
        }
        for (; j < P.rows; j += 4) {
            const Vec4<T> a = load4(base + (int64_t)j * ld);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const T xj = xv[(int64_t)j * K + k];
#pragma unroll
                for (int e = 0; e < 4; e++) acc[e][k] = fma(a.v[e], xj, acc[e][k]);
            }
        }
    }
    // waves add their partial sums into LDS one after another (fixed order)
    for (int w = 0; w < 4; w++) {
        if (wid == w)
#pragma unroll
            for (int e = 0; e < 4; e++)
#pragma unroll
                for (int k = 0; k < K; k++) red[c + e][k] = (w == 0 ? T(0) : red[c + e][k]) + acc[e][k];
        __syncthreads();
    }
    for (int q = threadIdx.x; q < P.w * K; q += blockDim.x) {
        const int col = q / K, k = q % K;
        if (P.out >= 0) part[(P.out * kProjTile + col) * K + k] = red[col][k];
        else t[tidx[P.tcol + col] * K + k] = red[col][k];
    }
}

template <typename T, int K>
__global__ __launch_bounds__(256) void proj_reduce_kernel(const ProjReduce *__restrict__ items,
                                                          const T *__restrict__ part, T *__restrict__ t,
                                                          const int64_t *__restrict__ tidx) {
    const ProjReduce R = items[blockIdx.x];
    const int col = threadIdx.x;
    if (col >= R.w) return;
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = T(0);
    for (int c = 0; c < R.nchunks; c++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += part[((R.slot0 + c) * kProjTile + col) * K + k];
    const int64_t ti = tidx[R.tcol + col];
#pragma unroll
    for (int k = 0; k < K; k++) t[ti * K + k] = acc[k];
}

template <typename T, int K>
__global__ __launch_bounds__(256) void tcopy_kernel(const TCopy *__restrict__ items, int64_t nitems,
                                                    const T *__restrict__ x, T *__restrict__ t) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int64_t it = (int64_t)blockIdx.x * 4 + wid;
    if (it >= nitems) return;
    const TCopy C = items[it];
    const int64_t tot = (int64_t)C.n * K;
    for (int64_t e = lane; e < tot; e += kWave) t[C.toff * K + e] = x[C.x0 * K + e];
}

template <typename T, int K>
__global__ __launch_bounds__(256) void panel_kernel(const PanelItem *__restrict__ items,
                                                    const Segment *__restrict__ segs,
                                                    const T *__restrict__ NP, const T *__restrict__ UP,
                                                    const T *__restrict__ x, const T *__restrict__ t,
                                                    T *__restrict__ y) {
    constexpr int NW = 4;
    __shared__ T red[NW][kWave][K];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const PanelItem P = items[xcd_remap(blockIdx.x, gridDim.x)];
    const int h = P.h;
    const int G = kWave / h;
    const int o = lane % h, g = lane / h;
    const bool active = g < G;
    const int gg = wid * G + g, GG = NW * G;
    const int64_t ld = P.ld;
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = T(0);
    for (int s = 0; s < P.nseg; s++) {
        const Segment S = segs[P.seg0 + s];
        const bool near = S.arena == 0;
        const T *M = (near ? NP : UP) + S.mat + P.r0 + o;
        const T *v = (near ? x : t) + S.vec * K;
        if (!active) continue;
        const int len = S.len;
        // kU column steps per iteration, predicated, so short segments
        // (near-field blocks of ~leaf-size width) still keep kU loads in flight
        constexpr int kU = 8;
        for (int q = gg; q < len; q += kU * GG) {
            T a[kU], vv[kU][K];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int qu = q + u * GG;
                const bool ok = qu < len;
                a[u] = ok ? load_stream(M + (int64_t)qu * ld) : T(0);
#pragma unroll
                for (int k = 0; k < K; k++) vv[u][k] = ok ? v[(int64_t)qu * K + k] : T(0);
            }
#pragma unroll
            for (int u = 0; u < kU; u++)
#pragma unroll
                for (int k = 0; k < K; k++) acc[k] = fma(a[u], vv[u][k], acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) red[wid][lane][k] = active ? acc[k] : T(0);
    __syncthreads();
    for (int e = threadIdx.x; e < h * K; e += blockDim.x) {
        const int row = e / K, k = e % K;
        T sum = T(0);
        for (int w = 0; w < NW; w++)
            for (int gi = 0; gi < G; gi++) sum += red[w][gi * h + row][k];
        y[(P.y0 + row) * K + k] = sum;
    }
}

template <typename T, int K>
hipError_t launch_matvec_k(const ProjItem *pi, int64_t npi, const ProjReduce *pr, int64_t npr,
                           const TCopy *tc, int64_t ntc, const PanelItem *pa, int64_t npa,
                           const Segment *segs, const T *VP, const T *NP, const T *UP,
                           const T *x, T *t, T *part, const int64_t *tidx, T *y,
                           hipEvent_t *ev, hipStream_t s) {
    if (ev) (void)hipEventRecord(ev[0], s);
    if (npi) proj_kernel<T, K><<<dim3((unsigned)npi), 256, 0, s>>>(pi, npi, VP, x, t, part, tidx);
    if (npr) proj_reduce_kernel<T, K><<<dim3((unsigned)npr), 256, 0, s>>>(pr, part, t, tidx);
    if (ntc) tcopy_kernel<T, K><<<dim3((unsigned)((ntc + 3) / 4)), 256, 0, s>>>(tc, ntc, x, t);
    if (ev) (void)hipEventRecord(ev[1], s);
    if (npa) panel_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, segs, NP, UP, x, t, y);
    if (ev) (void)hipEventRecord(ev[2], s);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_matvec(int nrhs, const ProjItem *pi, int64_t npi, const ProjReduce *pr, int64_t npr,
                         const TCopy *tc, int64_t ntc, const PanelItem *pa, int64_t npa,
                         const Segment *segs, const T *VP, const T *NP, const T *UP, const T *x,
                         T *t, T *part, const int64_t *tidx, T *y, hipEvent_t *ev, hipStream_t s) {
#define HODLR_MV(KK)                                                                                 \
    case KK:                                                                                         \
        return launch_matvec_k<T, KK>(pi, npi, pr, npr, tc, ntc, pa, npa, segs, VP, NP, UP, x, t, \
                                      part, tidx, y, ev, s);
    switch (nrhs) {
        HODLR_MV(1)
        HODLR_MV(2)
        HODLR_MV(4)
        HODLR_MV(8)
        HODLR_MV(16)
    case 32:  // the panel kernel's reduction buffer is 4*64*K*sizeof(T) bytes of LDS
        if constexpr (sizeof(T) == 4)
            return launch_matvec_k<T, 32>(pi, npi, pr, npr, tc, ntc, pa, npa, segs, VP, NP, UP, x, t,
                                          part, tidx, y, ev, s);
        return hipErrorInvalidValue;
    default:
        return hipErrorInvalidValue;
    }
#undef HODLR_MV
}

template hipError_t launch_matvec<double>(int, const ProjItem *, int64_t, const ProjReduce *, int64_t,
                                          const TCopy *, int64_t, const PanelItem *, int64_t,
                                          const Segment *, const double *, const double *, const double *,
                                          const double *, double *, double *, const int64_t *, double *,
                                          hipEvent_t *, hipStream_t);
template hipError_t launch_matvec<float>(int, const ProjItem *, int64_t, const ProjReduce *, int64_t,
                                         const TCopy *, int64_t, const PanelItem *, int64_t,
                                         const Segment *, const float *, const float *, const float *,
                                         const float *, float *, float *, const int64_t *, float *,
                                         hipEvent_t *, hipStream_t);

}  // namespace hodlr
