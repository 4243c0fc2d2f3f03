// skeleton.hip -- on-the-fly product of skeleton blocks (gfx950).
//
// A compressed block whose column node holds more than sys_size points
// (the SYS_SIZE dispatch of LowRank_matrix.hpp:22-36) keeps only its r
// pivot rows Is, its r pivot columns Js and G = C^+, the pseudo-inverse of
// the core C = A(Is, Js) (LowRank_matrix.hpp:685-698).  Every matvec
// evaluates the kernel again (LowRank_matrix.hpp:55-58):
//
//     y_I += A(I, Js) G A(Is, J) x_J
//
// in three phases:
//   proj    t1 = A(Is, J) x_J   per column node J: one lane per pivot row of
//                                every skeleton block (., J), the node's
//                                points streamed through LDS as sources;
//                                nodes longer than kSkelChunk points are cut
//                                into items whose partial sums are added in
//                                a fixed order (deterministic, no atomics)
//   core    t2 = G t1           one wave per block, lanes over G's rows
//   expand  y_I += A(I, Js) t2  per row node I: one lane per row, the pivot
//                                columns of every skeleton block (I, .) as
//                                sources; one launch per tree level, so each
//                                row has one writer per launch
//
// Entries are evaluated in FP64 for both operator precisions (an FP32
// operator reads x and adds into y in FP32): the product through C^+
// cancels large terms, so entries rounded to FP32 would lose the accuracy
// that FP32 storage of ACA factors keeps.  1/r is the hardware reciprocal
// square root refined by Newton steps (two for the FP64 operator: about
// 1 ulp; one for the FP32 operator: ~1e-14); this path does not have to
// round like the CPU restatement.  The phases are compute bound: about 15-20
// FP64 vector instructions per entry, plus one FMA per entry and RHS.
#include "aca_layout.h"
#include "device_common.h"

namespace hodlr {

namespace {

// 1/sqrt(s): the hardware estimate and NEWTON Newton steps (the estimate
// has about 23 good bits: one step gives ~1e-14, two give ~1 ulp)
template <int NEWTON>
__device__ __forceinline__ double rsqrt_nr(double s) {
    double y = __builtin_amdgcn_rsq(s);
#pragma unroll
    for (int it = 0; it < NEWTON; it++) {
        const double e = fma(-(s * y), y, 1.0);  // 1 - s y^2
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// f(r) of the kernel family from s = r^2 (user_kernel.hpp:46-49,101-104,167-170)
template <int FAM, int NEWTON>
__device__ __forceinline__ double kfun(double s) {
    if constexpr (FAM == 1) return rsqrt_nr<NEWTON>(s);          // 1 / r
    else if constexpr (FAM == 0) return 0.5 * log(s);            // log r
    else return exp(-(s * rsqrt_nr<NEWTON>(s)));                 // exp(-r)
}

// FP64 results keep every bit of 1/r; the FP32 operator's results need ~1e-14
template <typename T>
constexpr int kNewton = sizeof(T) == 8 ? 2 : 1;

// Both on-the-fly phases are one sum per lane over wave-uniform sources,
// acc_lane += f(|a_lane - b_s|^2) v_s: the sources are staged 64 at a time
// into LDS (lane l loads source l: coalesced), then every lane reads each
// source by a broadcast LDS read.  Source record: its ND coordinates then
// its NR values, S = ND + NR doubles.  One wave per workgroup, so the
// barriers around a tile only order this wave's LDS traffic.
template <int NR, int ND, int FAM, int NEWTON>
__device__ __forceinline__ void tile_sum(const double *__restrict__ lds, int S, int cnt, const double *a, int nd,
                                         double *acc) {
    constexpr int MD = ND > 0 ? ND : kMaxDimDev;
#pragma unroll 2
    for (int s = 0; s < cnt; s++) {
        const double *src = lds + s * S;
        double d2 = 0.0;
#pragma unroll
        for (int d = 0; d < MD; d++)
            if (d < nd) {
                const double t = a[d] - src[d];
                d2 = fma(t, t, d2);
            }
        const double f = kfun<FAM, NEWTON>(d2);
        const double *v = src + (ND > 0 ? ND : nd);
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] = fma(f, v[c], acc[c]);
    }
}

template <typename T, int NR, int ND, int FAM>
__device__ __forceinline__ void proj_item(const SkelProj &P, const KernelDev &K, const int32_t *__restrict__ SK,
                                          const T *__restrict__ x, double *__restrict__ t1,
                                          double *__restrict__ part, int lane, double *lds) {
    constexpr int MD = ND > 0 ? ND : kMaxDimDev;
    const int nd = ND > 0 ? ND : K.ndim;
    const int S = nd + NR;
    const bool act = lane < P.ntgt;
    const int64_t tp = SK[P.t0 + (act ? lane : 0)];
    const double *xa = K.X + (tp - K.pos0) * nd;
    double a[MD];
#pragma unroll
    for (int d = 0; d < MD; d++) a[d] = d < nd ? xa[d] : 0.0;
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
    for (int j0 = 0; j0 < P.nsrc; j0 += kWave) {
        const int cnt = min(kWave, P.nsrc - j0);
        __syncthreads();
        if (lane < cnt) {
            const int64_t j = P.src0 + j0 + lane;  // source tree position
            const double *xs = K.X + (j - K.pos0) * nd;
            double *dst = lds + lane * S;
#pragma unroll
            for (int d = 0; d < MD; d++)
                if (d < nd) dst[d] = xs[d];
            const double w = K.W ? K.W[j - K.pos0] : 1.0;
#pragma unroll
            for (int c = 0; c < NR; c++) dst[nd + c] = w * (double)x[j * NR + c];
        }
        __syncthreads();
        tile_sum<NR, ND, FAM, kNewton<T>>(lds, S, cnt, a, nd, acc);
    }
    if (!act) return;
    double *o = P.out < 0 ? t1 + (P.t0 + lane) * NR : part + (P.out + lane) * NR;
#pragma unroll
    for (int c = 0; c < NR; c++) o[c] = K.scale * acc[c];
}

template <typename T, int NR, int FAM>
__device__ __forceinline__ void proj_nd(const SkelProj &P, const KernelDev &K, const int32_t *SK, const T *x,
                                        double *t1, double *part, int lane, double *lds) {
    switch (K.ndim) {
    case 2: proj_item<T, NR, 2, FAM>(P, K, SK, x, t1, part, lane, lds); break;
    case 3: proj_item<T, NR, 3, FAM>(P, K, SK, x, t1, part, lane, lds); break;
    case 4: proj_item<T, NR, 4, FAM>(P, K, SK, x, t1, part, lane, lds); break;
    default: proj_item<T, NR, 0, FAM>(P, K, SK, x, t1, part, lane, lds); break;
    }
}

// one item per single-wave workgroup; dynamic LDS: 64 (max ndim + NR) doubles
template <typename T, int NR>
__global__ __launch_bounds__(64) void skel_proj_kernel(const SkelProj *__restrict__ items,
                                                       const KernelDev *__restrict__ Ks,
                                                       const int32_t *__restrict__ SK, const T *__restrict__ x,
                                                       double *__restrict__ t1, double *__restrict__ part) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const SkelProj P = items[blockIdx.x];
    const KernelDev K = Ks[P.prob];
    if (K.family == 1) proj_nd<T, NR, 1>(P, K, SK, x, t1, part, lane, lds);
    else if (K.family == 0) proj_nd<T, NR, 0>(P, K, SK, x, t1, part, lane, lds);
    else proj_nd<T, NR, 2>(P, K, SK, x, t1, part, lane, lds);
}

__device__ __forceinline__ int64_t wave_item() {
    return (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

template <int NR>
__global__ __launch_bounds__(256) void skel_reduce_kernel(const SkelReduce *__restrict__ items, int64_t n,
                                                          const double *__restrict__ part, double *__restrict__ t1) {
    const int64_t it = wave_item();
    if (it >= n) return;
    const int lane = threadIdx.x & (kWave - 1);
    const SkelReduce R = items[it];
    if (lane >= R.ntgt) return;
#pragma unroll
    for (int c = 0; c < NR; c++) {
        double s = 0.0;
        for (int ch = 0; ch < R.nch; ch++) s += part[(R.slot0 + (int64_t)ch * kWave + lane) * NR + c];
        t1[(R.t0 + lane) * NR + c] = s;
    }
}

template <int NR>
__global__ __launch_bounds__(256) void skel_core_kernel(const SkelCore *__restrict__ items, int64_t n,
                                                        const double *__restrict__ G, const double *__restrict__ t1,
                                                        double *__restrict__ t2) {
    const int64_t it = wave_item();
    if (it >= n) return;
    const int lane = threadIdx.x & (kWave - 1);
    const SkelCore C = items[it];
    for (int i0 = 0; i0 < C.r; i0 += kWave) {
        const int i = i0 + lane;
        const bool act = i < C.r;
        double acc[NR];
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] = 0.0;
        for (int k = 0; k < C.r; k++) {
            const double g = act ? G[C.g + (int64_t)k * C.r + i] : 0.0;
#pragma unroll
            for (int c = 0; c < NR; c++) acc[c] = fma(g, t1[(C.t0 + k) * NR + c], acc[c]);
        }
        if (act)
#pragma unroll
            for (int c = 0; c < NR; c++) t2[(C.c0 + i) * NR + c] = acc[c];
    }
}

template <typename T, int NR, int ND, int FAM>
__device__ __forceinline__ void expand_item(const SkelExpand &E, const KernelDev &K,
                                            const int32_t *__restrict__ SKC, const double *__restrict__ t2,
                                            T *__restrict__ y, int lane, double *lds) {
    constexpr int MD = ND > 0 ? ND : kMaxDimDev;
    const int nd = ND > 0 ? ND : K.ndim;
    const int S = nd + NR;
    const bool act = lane < E.h;
    const int64_t ip = E.y0 + (act ? lane : 0);
    const double *xa = K.X + (ip - K.pos0) * nd;
    double a[MD];
#pragma unroll
    for (int d = 0; d < MD; d++) a[d] = d < nd ? xa[d] : 0.0;
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
    for (int q0 = 0; q0 < E.nc; q0 += kWave) {
        const int cnt = min(kWave, E.nc - q0);
        __syncthreads();
        if (lane < cnt) {
            const int64_t q = E.c0 + q0 + lane;  // column slot
            const int64_t cp = SKC[q] - K.pos0;
            const double *xb = K.X + cp * nd;
            double *dst = lds + lane * S;
#pragma unroll
            for (int d = 0; d < MD; d++)
                if (d < nd) dst[d] = xb[d];
            const double w = K.W ? K.W[cp] : 1.0;
#pragma unroll
            for (int c = 0; c < NR; c++) dst[nd + c] = w * t2[q * NR + c];
        }
        __syncthreads();
        tile_sum<NR, ND, FAM, kNewton<T>>(lds, S, cnt, a, nd, acc);
    }
    if (!act) return;
    T *o = y + ip * NR;
#pragma unroll
    for (int c = 0; c < NR; c++) o[c] = (T)((double)o[c] + K.scale * acc[c]);
}

template <typename T, int NR, int FAM>
__device__ __forceinline__ void expand_nd(const SkelExpand &E, const KernelDev &K, const int32_t *SKC,
                                          const double *t2, T *y, int lane, double *lds) {
    switch (K.ndim) {
    case 2: expand_item<T, NR, 2, FAM>(E, K, SKC, t2, y, lane, lds); break;
    case 3: expand_item<T, NR, 3, FAM>(E, K, SKC, t2, y, lane, lds); break;
    case 4: expand_item<T, NR, 4, FAM>(E, K, SKC, t2, y, lane, lds); break;
    default: expand_item<T, NR, 0, FAM>(E, K, SKC, t2, y, lane, lds); break;
    }
}

template <typename T, int NR>
__global__ __launch_bounds__(64) void skel_expand_kernel(const SkelExpand *__restrict__ items,
                                                         const KernelDev *__restrict__ Ks,
                                                         const int32_t *__restrict__ SKC,
                                                         const double *__restrict__ t2, T *__restrict__ y) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const SkelExpand E = items[blockIdx.x];
    const KernelDev K = Ks[E.prob];
    if (K.family == 1) expand_nd<T, NR, 1>(E, K, SKC, t2, y, lane, lds);
    else if (K.family == 0) expand_nd<T, NR, 0>(E, K, SKC, t2, y, lane, lds);
    else expand_nd<T, NR, 2>(E, K, SKC, t2, y, lane, lds);
}

inline dim3 wave_grid(int64_t n) { return dim3((unsigned)((n + 3) / 4)); }

template <typename T, int NR>
hipError_t launch_skeleton_k(const SkeletonPlan &p, const T *x, double *t1, double *part, double *t2, T *y,
                             hipStream_t s) {
    const size_t lds = sizeof(double) * kWave * (size_t)(p.max_ndim + NR);
    if (p.nproj) skel_proj_kernel<T, NR><<<dim3((unsigned)p.nproj), 64, lds, s>>>(p.proj, p.K, p.SK, x, t1, part);
    if (p.nred) skel_reduce_kernel<NR><<<wave_grid(p.nred), 256, 0, s>>>(p.red, p.nred, part, t1);
    if (p.ncore) skel_core_kernel<NR><<<wave_grid(p.ncore), 256, 0, s>>>(p.core, p.ncore, p.G, t1, t2);
    for (int l = 0; l < p.nlvl; l++) {
        const int64_t a = p.ex_lvl[l], b = p.ex_lvl[l + 1];
        if (b > a)
            skel_expand_kernel<T, NR><<<dim3((unsigned)(b - a)), 64, lds, s>>>(p.ex + a, p.K, p.SK + p.nrow_slots,
                                                                               t2, y);
    }
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_skeleton(int nrhs, const SkeletonPlan &p, const T *x, double *t1, double *part, double *t2, T *y,
                           hipStream_t s) {
    switch (nrhs) {
    case 1: return launch_skeleton_k<T, 1>(p, x, t1, part, t2, y, s);
    case 2: return launch_skeleton_k<T, 2>(p, x, t1, part, t2, y, s);
    case 4: return launch_skeleton_k<T, 4>(p, x, t1, part, t2, y, s);
    case 8: return launch_skeleton_k<T, 8>(p, x, t1, part, t2, y, s);
    case 16: return launch_skeleton_k<T, 16>(p, x, t1, part, t2, y, s);
    case 32:
        if constexpr (sizeof(T) == 4) return launch_skeleton_k<T, 32>(p, x, t1, part, t2, y, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_skeleton<double>(int, const SkeletonPlan &, const double *, double *, double *, double *,
                                            double *, hipStream_t);
template hipError_t launch_skeleton<float>(int, const SkeletonPlan &, const float *, double *, double *, double *,
                                           float *, hipStream_t);

}  // namespace hodlr
