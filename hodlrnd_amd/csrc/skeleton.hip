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
//   proj    t1 = A(Is, J) x_J   per column node J: two lanes' worth of pivot
//                                rows of every skeleton block (., J), the node's
//                                points streamed through LDS as sources;
//                                nodes longer than kSkelChunk points are cut
//                                into items whose partial sums are added in
//                                a fixed order (deterministic, no atomics)
//   core    t2 = G t1           one wave per block, lanes over G's rows
//   expand  ys_I += A(I, Js) t2 per row node I: two rows per lane, the pivot
//                                columns of every skeleton block (I, .) as
//                                sources; one launch per tree level, so each
//                                row has one writer per launch
//
// ys is an FP64 vector of its own (zeroed first), added to y once the
// stored blocks' product has written y (launch_skeleton_add).  The phases
// touch y only in that last step, so they can run on a second stream while
// the stored blocks' passes stream their arenas, and the result does not
// depend on whether they did.
//
// Entries are evaluated in FP64 for both operator precisions (an FP32
// operator reads x in FP32 and rounds y + ys to FP32): the product through C^+
// cancels large terms, so entries rounded to FP32 would lose the accuracy
// that FP32 storage of ACA factors keeps.  1/r is the hardware reciprocal
// square root refined by Newton steps (two for the FP64 operator: about
// 1 ulp; one for the FP32 operator: ~1e-14), and the FP32 operator forms
// squared distances from centred coordinates (see tile_sum); this path does
// not have to round like the CPU restatement.  The phases are compute
// bound: about 13-20
// FP64 vector instructions per entry, plus one FMA per entry and RHS.
#include "../../include/hodlrnd/device/aca_layout.h"
#include "../../include/hodlrnd/device/device_common.h"

namespace hodlr {

namespace {

// 1/sqrt(s) from the hardware estimate y (about 23 good bits) and one
// refinement with e = 1 - s y^2: the cubic step y (1 + e/2 + 3 e^2/8)
// (error 5 e^3 / 16, below FP64 rounding) for an FP64 operator, Newton's
// y (1 + e/2) (~1e-14) for the FP32 operator
template <bool CUBIC>
__device__ __forceinline__ double rsqrt_ref(double s) {
    const double y = __builtin_amdgcn_rsq(s);
    const double e = fma(-(s * y), y, 1.0);
    if constexpr (CUBIC) return fma(y * e, fma(e, 0.375, 0.5), y);
    else return fma(0.5 * y, e, y);
}

// f(r) of the kernel family from s = r^2 (user_kernel.hpp:46-49,101-104,167-170)
template <int FAM, bool CUBIC>
__device__ __forceinline__ double kfun(double s) {
    if constexpr (FAM == 1) return rsqrt_ref<CUBIC>(s);          // 1 / r
    else if constexpr (FAM == 0) return 0.5 * log(s);            // log r
    else return exp(-(s * rsqrt_ref<CUBIC>(s)));                 // exp(-r)
}

template <typename T>
constexpr bool kCubic = sizeof(T) == 8;

// Both on-the-fly phases are sums over sources for up to 128 targets per
// wave, two per lane when there are more than 64 (lane l owns targets l and
// l + 64), else one:
//     acc_t += f(|a_t - b_s|^2) v_s.
// The sources are staged 64 at a time into LDS (lane l loads source l:
// coalesced), then every lane reads each source by a broadcast LDS read that
// feeds both of its targets.  One wave per workgroup, so the barriers around
// a tile only order this wave's LDS traffic.
//
// FP64 operator: |a - b|^2 from the differences (ND subtractions and ND
// FMAs per target and source); source record: ND coordinates then NR
// values, S = ND + NR doubles.  FP32 operator: coordinates relative to a
// point c of the item's compact side (the column node of a projection, the
// row node of an expansion), |a'|^2 + |b'|^2 - 2 a'.b' (one addition and
// ND FMAs); source record: -2 b', |b'|^2, then the values, S = ND + 1 + NR.
// The skeleton blocks are admissible, so |a - b| is at least about the
// compact node's size and (|a'|^2 + |b'|^2) / |a - b|^2 stays below ~10:
// the squared distance keeps ~1e-15 relative accuracy, far inside the
// FP32 operator's 1e-14 reciprocal square root.
template <int NR, int ND, int FAM, bool CUBIC, int TPL>
__device__ __forceinline__ void tile_sum(const double *__restrict__ lds, int S, int cnt, const double *a0,
                                         const double *a1, double ta0, double ta1, int nd, double *acc0,
                                         double *acc1) {
    constexpr int MD = ND > 0 ? ND : kMaxDimDev;
    constexpr bool kNorm = !CUBIC;
    // four sources per iteration: their LDS reads issue together (84-90
    // VGPRs, 5-6 waves per SIMD; two sources kept 8 waves but measured 2.5%
    // slower, one 1% slower still)
#pragma unroll 4
    for (int s = 0; s < cnt; s++) {
        const double *src = lds + s * S;
        double d0, d1;
        if constexpr (kNorm) {
            const double tb = src[ND > 0 ? ND : nd];
            d0 = ta0 + tb;
            d1 = ta1 + tb;
#pragma unroll
            for (int d = 0; d < MD; d++)
                if (d < nd) {
                    const double q = src[d];
                    d0 = fma(a0[d], q, d0);
                    if constexpr (TPL == 2) d1 = fma(a1[d], q, d1);
                }
        } else {
            d0 = d1 = 0.0;
#pragma unroll
            for (int d = 0; d < MD; d++)
                if (d < nd) {
                    const double b = src[d];
                    const double t0 = a0[d] - b;
                    d0 = fma(t0, t0, d0);
                    if constexpr (TPL == 2) {
                        const double t1 = a1[d] - b;
                        d1 = fma(t1, t1, d1);
                    }
                }
        }
        const double f0 = kfun<FAM, CUBIC>(d0);
        const double *v = src + (ND > 0 ? ND : nd) + (kNorm ? 1 : 0);
        if constexpr (TPL == 2) {
            const double f1 = kfun<FAM, CUBIC>(d1);
#pragma unroll
            for (int c = 0; c < NR; c++) {
                const double vc = v[c];
                acc0[c] = fma(f0, vc, acc0[c]);
                acc1[c] = fma(f1, vc, acc1[c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NR; c++) acc0[c] = fma(f0, v[c], acc0[c]);
        }
    }
}

// Loads through the global address space: a generic (flat) load also counts
// against the LDS wait counter, so waiting for the tile's LDS reads would wait
// for an in-flight prefetch as well.
template <class T>
__device__ __forceinline__ T gload(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}

// The source loop of one item: tiles of up to 64 sources, staged through
// LDS.  fetch(t, c, v, w) loads this lane's source of tile t -- coordinates
// c, raw values v, weight w -- with plain loads only (no arithmetic, so no
// wait); the record stored to LDS holds c and w v.  With few values per
// source the next tile is fetched into registers while the current one is
// summed, so its load latency overlaps the arithmetic.
template <int NR, int ND, int FAM, bool CUBIC, class V, class Fetch>
__device__ __forceinline__ void source_loop(int nsrc, int nd, int lane, double *lds, const double *a0,
                                            const double *a1, double ta0, double ta1, const double *ctr,
                                            double *acc0, double *acc1, bool two, Fetch fetch) {
    constexpr int MD = ND > 0 ? ND : kMaxDimDev;
    constexpr bool kPrefetch = MD + NR <= 12;
    constexpr bool kNorm = !CUBIC;
    const int S = nd + NR + (kNorm ? 1 : 0);
    const int ntiles = (nsrc + kWave - 1) / kWave;
    double c[MD], w = 1.0;
    V v[NR];
    if (kPrefetch && ntiles > 0 && lane < min(kWave, nsrc)) fetch(0, c, v, w);
    for (int t = 0; t < ntiles; t++) {
        const int cnt = min(kWave, nsrc - t * kWave);
        __syncthreads();  // the previous tile is consumed
        if (lane < cnt) {
            if (!kPrefetch) fetch(t, c, v, w);
            double *dst = lds + lane * S;
            if constexpr (kNorm) {
                double tb = 0.0;
#pragma unroll
                for (int d = 0; d < MD; d++)
                    if (d < nd) {
                        const double bd = c[d] - ctr[d];
                        dst[d] = -2.0 * bd;
                        tb = fma(bd, bd, tb);
                    }
                dst[nd] = tb;
#pragma unroll
                for (int k = 0; k < NR; k++) dst[nd + 1 + k] = w * (double)v[k];
            } else {
#pragma unroll
                for (int d = 0; d < MD; d++)
                    if (d < nd) dst[d] = c[d];
#pragma unroll
                for (int k = 0; k < NR; k++) dst[nd + k] = w * (double)v[k];
            }
        }
        __syncthreads();
        if (kPrefetch && t + 1 < ntiles && lane < min(kWave, nsrc - (t + 1) * kWave)) fetch(t + 1, c, v, w);
        if (two) tile_sum<NR, ND, FAM, CUBIC, 2>(lds, S, cnt, a0, a1, ta0, ta1, nd, acc0, acc1);
        else tile_sum<NR, ND, FAM, CUBIC, 1>(lds, S, cnt, a0, a1, ta0, ta1, nd, acc0, acc1);
    }
}

template <int MD>
__device__ __forceinline__ void load_point(const KernelDev &K, int64_t pos, int nd, double *a) {
    const double *xa = K.X + (pos - K.pos0) * nd;
#pragma unroll
    for (int d = 0; d < MD; d++) a[d] = d < nd ? gload(xa + d) : 0.0;
}

// every load issued so far has landed: the target coordinates are then not
// waited for inside the source loop, where that wait would also wait for
// the in-flight prefetch
__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0); }

// FP32 operator: the targets relative to ctr and their squared norms
template <int MD, bool CUBIC>
__device__ __forceinline__ void center_targets(const double *ctr, int nd, double *a0, double *a1, double &ta0,
                                               double &ta1) {
    ta0 = ta1 = 0.0;
    if constexpr (!CUBIC) {
#pragma unroll
        for (int d = 0; d < MD; d++)
            if (d < nd) {
                a0[d] -= ctr[d];
                a1[d] -= ctr[d];
                ta0 = fma(a0[d], a0[d], ta0);
                ta1 = fma(a1[d], a1[d], ta1);
            }
    }
}

template <typename T, int NR, int ND, int FAM>
__device__ __forceinline__ void proj_item(const SkelProj &P, const KernelDev &K, const int32_t *__restrict__ SK,
                                          const T *__restrict__ x, double *__restrict__ t1,
                                          double *__restrict__ part, int lane, double *lds) {
    constexpr int MD = ND > 0 ? ND : kMaxDimDev;
    const int nd = ND > 0 ? ND : K.ndim;
    const bool act0 = lane < P.ntgt, act1 = lane + kWave < P.ntgt;
    double a0[MD], a1[MD], ctr[MD];
    load_point<MD>(K, SK[P.t0 + (act0 ? lane : 0)], nd, a0);
    load_point<MD>(K, SK[P.t0 + (act1 ? lane + kWave : 0)], nd, a1);
    if constexpr (!kCubic<T>) load_point<MD>(K, P.src0, nd, ctr);  // a point of the column node
    wait_loads();
    double ta0, ta1;
    center_targets<MD, kCubic<T>>(ctr, nd, a0, a1, ta0, ta1);
    double acc0[NR], acc1[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc0[c] = acc1[c] = 0.0;
    source_loop<NR, ND, FAM, kCubic<T>, T>(P.nsrc, nd, lane, lds, a0, a1, ta0, ta1, ctr, acc0, acc1,
                                            P.ntgt > kWave,
                                            [&](int t, double *c, T *v, double &w) {
                                                const int64_t j = P.src0 + (int64_t)t * kWave + lane;
                                                const double *xs = K.X + (j - K.pos0) * nd;
#pragma unroll
                                                for (int d = 0; d < MD; d++)
                                                    if (d < nd) c[d] = gload(xs + d);
                                                if (K.W) w = gload(K.W + (j - K.pos0));
#pragma unroll
                                                for (int k = 0; k < NR; k++) v[k] = gload(x + j * NR + k);
                                            });
    const double sc = K.scale;
    if (act0) {
        double *o = P.out < 0 ? t1 + (P.t0 + lane) * NR : part + (P.out + lane) * NR;
#pragma unroll
        for (int c = 0; c < NR; c++) o[c] = sc * acc0[c];
    }
    if (act1) {
        double *o = P.out < 0 ? t1 + (P.t0 + lane + kWave) * NR : part + (P.out + lane + kWave) * NR;
#pragma unroll
        for (int c = 0; c < NR; c++) o[c] = sc * acc1[c];
    }
}

// one item per single-wave workgroup, one (family, NDIM) per launch (ND = 0:
// any NDIM up to 8); dynamic LDS: 64 (NDIM + NR) doubles, one more per source
// for the FP32 operator
template <typename T, int NR, int ND, int FAM>
__global__ __launch_bounds__(64) void skel_proj_kernel(const SkelProj *__restrict__ items,
                                                       const KernelDev *__restrict__ Ks,
                                                       const int32_t *__restrict__ SK, const T *__restrict__ x,
                                                       double *__restrict__ t1, double *__restrict__ part) {
    extern __shared__ double lds[];
    const SkelProj P = items[blockIdx.x];
    const KernelDev K = Ks[P.prob];
    proj_item<T, NR, ND, FAM>(P, K, SK, x, t1, part, threadIdx.x, lds);
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
    for (int l = lane; l < R.ntgt; l += kWave)
#pragma unroll
        for (int c = 0; c < NR; c++) {
            double s = 0.0;
            for (int ch = 0; ch < R.nch; ch++) s += part[(R.slot0 + (int64_t)ch * kSkelTile + l) * NR + c];
            t1[(R.t0 + l) * NR + c] = s;
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
#pragma unroll 8
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
                                            double *__restrict__ ys, int lane, double *lds) {
    constexpr int MD = ND > 0 ? ND : kMaxDimDev;
    const int nd = ND > 0 ? ND : K.ndim;
    const bool act0 = lane < E.h, act1 = lane + kWave < E.h;
    const int64_t ip0 = E.y0 + (act0 ? lane : 0), ip1 = E.y0 + (act1 ? lane + kWave : 0);
    double a0[MD], a1[MD], ctr[MD];
    load_point<MD>(K, ip0, nd, a0);
    load_point<MD>(K, ip1, nd, a1);
    if constexpr (!kCubic<T>) load_point<MD>(K, E.y0, nd, ctr);  // a point of the row node
    wait_loads();
    double ta0, ta1;
    center_targets<MD, kCubic<T>>(ctr, nd, a0, a1, ta0, ta1);
    double acc0[NR], acc1[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc0[c] = acc1[c] = 0.0;
    source_loop<NR, ND, FAM, kCubic<T>, double>(E.nc, nd, lane, lds, a0, a1, ta0, ta1, ctr, acc0, acc1,
                                                 E.h > kWave,
                                                 [&](int t, double *c, double *v, double &w) {
                                                     const int64_t q = E.c0 + (int64_t)t * kWave + lane;  // slot
                                                     const int64_t cp = gload(SKC + q) - K.pos0;
                                                     const double *xb = K.X + cp * nd;
#pragma unroll
                                                     for (int d = 0; d < MD; d++)
                                                         if (d < nd) c[d] = gload(xb + d);
                                                     if (K.W) w = gload(K.W + cp);
#pragma unroll
                                                     for (int k = 0; k < NR; k++) v[k] = gload(t2 + q * NR + k);
                                                 });
    const double sc = K.scale;
    if (act0) {
        double *o = ys + ip0 * NR;
#pragma unroll
        for (int c = 0; c < NR; c++) o[c] = o[c] + sc * acc0[c];
    }
    if (act1) {
        double *o = ys + ip1 * NR;
#pragma unroll
        for (int c = 0; c < NR; c++) o[c] = o[c] + sc * acc1[c];
    }
}

template <typename T, int NR, int ND, int FAM>
__global__ __launch_bounds__(64) void skel_expand_kernel(const SkelExpand *__restrict__ items,
                                                         const KernelDev *__restrict__ Ks,
                                                         const int32_t *__restrict__ SKC,
                                                         const double *__restrict__ t2, double *__restrict__ ys) {
    extern __shared__ double lds[];
    const SkelExpand E = items[blockIdx.x];
    const KernelDev K = Ks[E.prob];
    expand_item<T, NR, ND, FAM>(E, K, SKC, t2, ys, threadIdx.x, lds);
}

// ys = 0 (the expand launches add into it)
__global__ __launch_bounds__(256) void skel_zero_kernel(double *__restrict__ ys, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) ys[i] = 0.0;
}

// y += ys, rounded once to the operator precision
template <typename T>
__global__ __launch_bounds__(256) void skel_add_kernel(T *__restrict__ y, const double *__restrict__ ys, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        y[i] = (T)((double)y[i] + ys[i]);
}

inline unsigned stream_blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

inline dim3 wave_grid(int64_t n) { return dim3((unsigned)((n + 3) / 4)); }

template <typename T, int NR, int ND, int FAM>
void launch_group(bool proj, const SkeletonPlan &p, const SkelLaunch &L, size_t lds, const T *x, double *t1,
                  double *part, const double *t2, double *ys, hipStream_t s) {
    const dim3 g((unsigned)(L.end - L.begin));
    if (proj)
        skel_proj_kernel<T, NR, ND, FAM><<<g, 64, lds, s>>>(p.proj + L.begin, p.K, p.SK, x, t1, part);
    else
        skel_expand_kernel<T, NR, ND, FAM><<<g, 64, lds, s>>>(p.ex + L.begin, p.K, p.SK + p.nrow_slots, t2, ys);
}

template <typename T, int NR, int FAM>
void launch_nd(bool proj, const SkeletonPlan &p, const SkelLaunch &L, size_t lds, const T *x, double *t1,
               double *part, const double *t2, double *ys, hipStream_t s) {
    switch (L.key % 16) {
    case 2: launch_group<T, NR, 2, FAM>(proj, p, L, lds, x, t1, part, t2, ys, s); break;
    case 3: launch_group<T, NR, 3, FAM>(proj, p, L, lds, x, t1, part, t2, ys, s); break;
    case 4: launch_group<T, NR, 4, FAM>(proj, p, L, lds, x, t1, part, t2, ys, s); break;
    default: launch_group<T, NR, 0, FAM>(proj, p, L, lds, x, t1, part, t2, ys, s); break;
    }
}

template <typename T, int NR>
void launch_key(bool proj, const SkeletonPlan &p, const SkelLaunch &L, const T *x, double *t1, double *part,
                const double *t2, double *ys, hipStream_t s) {
    if (L.end <= L.begin) return;
    // one record per source: NDIM coordinates, its squared norm (FP32
    // operator), NR values
    size_t lds = sizeof(double) * kWave * (size_t)(L.key % 16 + NR + (sizeof(T) == 4 ? 1 : 0));
    if (lds < (size_t)p.lds_min) lds = (size_t)p.lds_min;
    switch (L.key / 16) {
    case 0: launch_nd<T, NR, 0>(proj, p, L, lds, x, t1, part, t2, ys, s); break;
    case 1: launch_nd<T, NR, 1>(proj, p, L, lds, x, t1, part, t2, ys, s); break;
    default: launch_nd<T, NR, 2>(proj, p, L, lds, x, t1, part, t2, ys, s); break;
    }
}

template <typename T, int NR>
hipError_t launch_skeleton_k(const SkeletonPlan &p, const T *x, double *t1, double *part, double *t2, double *ys,
                             int64_t nys, hipStream_t s) {
    skel_zero_kernel<<<stream_blocks(nys), 256, 0, s>>>(ys, nys);
    for (int i = 0; i < p.nproj_l; i++) launch_key<T, NR>(true, p, p.proj_l[i], x, t1, part, t2, ys, s);
    if (p.nred) skel_reduce_kernel<NR><<<wave_grid(p.nred), 256, 0, s>>>(p.red, p.nred, part, t1);
    if (p.ncore) skel_core_kernel<NR><<<wave_grid(p.ncore), 256, 0, s>>>(p.core, p.ncore, p.G, t1, t2);
    for (int i = 0; i < p.nex_l; i++) launch_key<T, NR>(false, p, p.ex_l[i], x, t1, part, t2, ys, s);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_skeleton(int nrhs, const SkeletonPlan &p, const T *x, double *t1, double *part, double *t2,
                           double *ys, int64_t nrows, hipStream_t s) {
    const int64_t n = nrows * nrhs;
    switch (nrhs) {
    case 1: return launch_skeleton_k<T, 1>(p, x, t1, part, t2, ys, n, s);
    case 2: return launch_skeleton_k<T, 2>(p, x, t1, part, t2, ys, n, s);
    case 4: return launch_skeleton_k<T, 4>(p, x, t1, part, t2, ys, n, s);
    case 8: return launch_skeleton_k<T, 8>(p, x, t1, part, t2, ys, n, s);
    case 16: return launch_skeleton_k<T, 16>(p, x, t1, part, t2, ys, n, s);
    case 32:
        if constexpr (sizeof(T) == 4) return launch_skeleton_k<T, 32>(p, x, t1, part, t2, ys, n, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_skeleton_add(T *y, const double *ys, int64_t n, hipStream_t s) {
    skel_add_kernel<T><<<stream_blocks(n), 256, 0, s>>>(y, ys, n);
    return hipGetLastError();
}

template hipError_t launch_skeleton<double>(int, const SkeletonPlan &, const double *, double *, double *, double *,
                                            double *, int64_t, hipStream_t);
template hipError_t launch_skeleton<float>(int, const SkeletonPlan &, const float *, double *, double *, double *,
                                           double *, int64_t, hipStream_t);
template hipError_t launch_skeleton_add<double>(double *, const double *, int64_t, hipStream_t);
template hipError_t launch_skeleton_add<float>(float *, const double *, int64_t, hipStream_t);

}  // namespace hodlr
