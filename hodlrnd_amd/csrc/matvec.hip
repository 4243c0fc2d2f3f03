// matvec.hip -- the fast HODLR matvec (gfx950), FP64 or FP32, 1..32 RHS.
//
// The reference applies every node's blocks one GEMV at a time
// (HDD_tree.hpp:86-107, HDD_node.hpp:223-254, LowRank_matrix.hpp:40-63).
// Here the operator is stored in three streaming arenas and applied by two
// passes, each reading its arena exactly once:
//   proj:   t = V^T x for every compressed block.  The V factors of all
//           blocks that share a column node b form one panel (n_b x T_b,
//           T_b padded to a multiple of 4) stored as 256-column tiles, each
//           a contiguous row-major block; a workgroup streams one tile, each
//           lane four adjacent columns (one vector load per row), and
//           x_b[j] is a broadcast.  Tall tiles are cut
//           into row chunks whose partial sums are added by a second small
//           pass (deterministic order).
//   panel:  y = sum over segments of M v_seg, one workgroup per leaf row
//           tile.  A leaf's segments are its near-field blocks (self and
//           neighbours, times x) and its slice of the U panel of each
//           ancestor, cut into runs of low-rank blocks (times t) and runs of
//           dense-fallback blocks, whose U columns hold K (times x).  All are
//           column-major with leading dimension = leaf size, so G = 64/h
//           column groups of h lanes read G*h contiguous elements per step.
//           Each leaf row is written once (no atomics).
// With 4 or more right-hand sides both passes also have matrix-core forms
// (further down); the operator picks one per pass.
#include "../../include/hodlrnd/device/aca_layout.h"
#include "../../include/hodlrnd/device/device_common.h"

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
                                                   const int32_t *__restrict__ tidx) {
    constexpr int U = 4;  // rows in flight per wave per step
    __shared__ T red[kProjTile][K];
    // uniform wave index: the row loop is scalar, x_b[j] a scalar load
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
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
        else t[(int64_t)tidx[P.tcol + col] * K + k] = red[col][k];
    }
}

// Projection of short items (at most kProjWaveRows rows): one wave per item,
// four items per workgroup.  The wave walks all of its item's rows (lane l:
// columns 4l..4l+3, U rows in flight) and writes its sums from registers --
// no LDS and no workgroup barrier.  In proj_kernel a short item gives each
// of four waves a few rows and then runs four reduction rounds through LDS;
// from 8 RHS on those rounds outweigh the item's loads, while at 1-4 RHS
// the four-wave form, with more loads in flight per item, is faster.
template <typename T, int K>
__global__ __launch_bounds__(256) void proj_wave_kernel(const ProjItem *__restrict__ items, int64_t nitems,
                                                        const T *__restrict__ VP, const T *__restrict__ x,
                                                        T *__restrict__ t, T *__restrict__ part,
                                                        const int32_t *__restrict__ tidx) {
    constexpr int U = 4;
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t it = (int64_t)blockIdx.x * 4 + wid;
    if (it >= nitems) return;
    const ProjItem P = items[it];
    const int c = lane * 4;
    if (c >= P.w) return;
    T acc[4][K];
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[e][k] = T(0);
    const T *base = VP + P.vp + c;
    const T *xv = x + P.x0 * K;
    const int64_t ld = P.ld;
    int j = 0;
    for (; j + U <= P.rows; j += U) {
        Vec4<T> a[U];
#pragma unroll
        for (int u = 0; u < U; u++) a[u] = load4(base + (int64_t)(j + u) * ld);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int k = 0; k < K; k++) {
                const T xj = xv[(int64_t)(j + u) * K + k];
#pragma unroll
                for (int e = 0; e < 4; e++) acc[e][k] = fma(a[u].v[e], xj, acc[e][k]);
            }
    }
    for (; j < P.rows; j++) {
        const Vec4<T> a = load4(base + (int64_t)j * ld);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const T xj = xv[(int64_t)j * K + k];
#pragma unroll
            for (int e = 0; e < 4; e++) acc[e][k] = fma(a.v[e], xj, acc[e][k]);
        }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int col = c + e;
        if (col < P.w) {
            T *dst = P.out >= 0 ? part + (P.out * kProjTile + col) * K : t + (int64_t)tidx[P.tcol + col] * K;
#pragma unroll
            for (int k = 0; k < K; k++) dst[k] = acc[e][k];
        }
    }
}

// Row-packed projection for narrow tiles (1-2 RHS).  A tile of width tw has
// Q = tw/4 column quads per row; a wave reads RW = 64/Q consecutive rows per
// load (lane l: quad l % Q, row offset l / Q), so a 100-column tile keeps 50
// lanes busy instead of 25, and every wave load is one contiguous run of
// RW*tw elements.  Waves take blocks of RW rows in turn, kU loads in flight;
// x_j is a per-lane load.  The partial sums of the RW row offsets and the 4
// waves are added through LDS in a fixed order (one thread per column).
template <typename T, int K>
__global__ __launch_bounds__(256) void proj_pack_kernel(const ProjItem *__restrict__ items, int64_t nitems,
                                                        const T *__restrict__ VP, const T *__restrict__ x,
                                                        T *__restrict__ t, T *__restrict__ part,
                                                        const int32_t *__restrict__ tidx) {
    constexpr int kU = 8;
    __shared__ T red[4 * kWave][4][K];
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const ProjItem P = items[blockIdx.x];
    const int ld = P.ld;
    const int Q = ld >> 2, RW = kWave / Q;
    const int quad = lane % Q, roff = lane / Q;
    const bool live = roff < RW && 4 * quad < P.w;
    T acc[4][K];
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[e][k] = T(0);
    const T *base = VP + P.vp + 4 * quad;
    const T *xv = x + P.x0 * K;
    const int stride = 4 * RW;  // rows between a wave's consecutive loads
    for (int j0 = wid * RW; j0 < P.rows; j0 += kU * stride) {  // wave-uniform
        Vec4<T> a[kU];
        T xj[kU][K];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int j = j0 + u * stride + roff;
            const bool ok = live && j < P.rows;
            if (ok) {
                a[u] = load4(base + (int64_t)j * ld);
#pragma unroll
                for (int k = 0; k < K; k++) xj[u][k] = xv[(int64_t)j * K + k];
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) a[u].v[e] = T(0);
#pragma unroll
                for (int k = 0; k < K; k++) xj[u][k] = T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
#pragma unroll
            for (int k = 0; k < K; k++)
#pragma unroll
                for (int e = 0; e < 4; e++) acc[e][k] = fma(a[u].v[e], xj[u][k], acc[e][k]);
    }
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
        for (int k = 0; k < K; k++) red[threadIdx.x][e][k] = acc[e][k];
    __syncthreads();
    const int col = threadIdx.x;
    if (col < P.w) {
        const int cq = col >> 2, ce = col & 3;
        T s[K];
#pragma unroll
        for (int k = 0; k < K; k++) s[k] = T(0);
        for (int w = 0; w < 4; w++)
            for (int r = 0; r < RW; r++)
#pragma unroll
                for (int k = 0; k < K; k++) s[k] += red[w * kWave + r * Q + cq][ce][k];
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (P.out >= 0) part[(P.out * kProjTile + col) * K + k] = s[k];
            else t[(int64_t)tidx[P.tcol + col] * K + k] = s[k];
        }
    }
}

template <typename T, int K>
__global__ __launch_bounds__(256) void proj_reduce_kernel(const ProjReduce *__restrict__ items,
                                                          const T *__restrict__ part, T *__restrict__ t,
                                                          const int32_t *__restrict__ tidx) {
    const ProjReduce R = items[blockIdx.x];
    const int col = threadIdx.x;
    if (col >= R.w) return;
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = T(0);
    for (int c = 0; c < R.nchunks; c++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += part[((R.slot0 + c) * kProjTile + col) * K + k];
    const int64_t ti = tidx[R.tcol + col];  // widened before the multiply
#pragma unroll
    for (int k = 0; k < K; k++) t[ti * K + k] = acc[k];
}

// Vector-ALU panel product, one segment at a time.  Lane (o, g) of a wave
// owns row o of the tile and column group g (G = 64 / h groups of h lanes);
// the 4 waves interleave column groups inside each segment, kU predicated
// column steps per iteration so short segments still keep kU loads in
// flight.  Best when segments are long (wide U slices, large leaves).
template <typename T, int K>
__global__ __launch_bounds__(256) void panel_seg_kernel(const PanelItem *__restrict__ items,
                                                        const Segment *__restrict__ segs,
                                                        const T *__restrict__ NP, const T *__restrict__ UP,
                                                        const T *__restrict__ x, const T *__restrict__ t,
                                                        T *__restrict__ y) {
    constexpr int NW = 4, kU = 8;
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
        const T *M = (S.arena == 0 ? NP : UP) + S.mat + P.r0 + o;
        const T *v = (S.arena == 1 ? t : x) + S.vec * K;
        if (!active) continue;
        const int len = S.len;
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

// Vector-ALU panel product with 16-byte loads: lane (o, g) owns rows
// R o .. R o + R - 1 of the tile (R = 16 / sizeof(T): two FP64 or four FP32
// rows) in column group g, G = 64 / ceil(h / R) groups, and reads its R
// rows of a column with one dword-aligned 16-byte load.  The last lane of a
// column may read up to R - 1 elements past the tile (the next column, or
// the arena's tail padding) into rows >= h, which the reduction drops.
// Segments one at a time, the 4 waves interleaving column groups, kU
// predicated column steps in flight per lane.
template <typename T, int K>
__global__ __launch_bounds__(256) void panel_vec_kernel(const PanelItem *__restrict__ items,
                                                        const Segment *__restrict__ segs,
                                                        const T *__restrict__ NP, const T *__restrict__ UP,
                                                        const T *__restrict__ x, const T *__restrict__ t,
                                                        T *__restrict__ y) {
    // column steps in flight per lane: fewer for more RHS (registers)
    constexpr int R = 16 / (int)sizeof(T), NW = 4, kU = K <= 2 ? 8 : (K <= 4 ? 4 : 2);
    typedef T vec_t __attribute__((ext_vector_type(R), aligned(sizeof(T))));
    __shared__ T red[NW][kWave][R][K];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const PanelItem P = items[xcd_remap(blockIdx.x, gridDim.x)];
    const int h = P.h;
    const int hl = (h + R - 1) / R;  // lanes per column
    const int G = kWave / hl;
    const int o = lane % hl, g = lane / hl;
    const bool active = g < G;
    const int gg = wid * G + g, GG = NW * G;
    const int64_t ld = P.ld;
    T acc[R][K];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[r][k] = T(0);
    for (int s = 0; s < P.nseg; s++) {
        const Segment S = segs[P.seg0 + s];
        const T *M = (S.arena == 0 ? NP : UP) + S.mat + P.r0 + (int64_t)o * R;
        const T *v = (S.arena == 1 ? t : x) + S.vec * K;
        if (!active) continue;
        const int len = S.len;
        for (int q = gg; q < len; q += kU * GG) {
            vec_t a[kU];
            T vv[kU][K];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int qu = q + u * GG;
                const bool ok = qu < len;
                a[u] = ok ? *reinterpret_cast<const vec_t *>(M + (int64_t)qu * ld) : vec_t(T(0));
                if constexpr (K % R == 0) {  // the column's K values as K/R 16-byte loads
#pragma unroll
                    for (int kk = 0; kk < K / R; kk++) {
                        const vec_t w = ok ? *reinterpret_cast<const vec_t *>(v + (int64_t)qu * K + kk * R)
                                           : vec_t(T(0));
#pragma unroll
                        for (int r = 0; r < R; r++) vv[u][kk * R + r] = w[r];
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < K; k++) vv[u][k] = ok ? v[(int64_t)qu * K + k] : T(0);
                }
            }
#pragma unroll
            for (int u = 0; u < kU; u++)
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int k = 0; k < K; k++) acc[r][k] = fma(a[u][r], vv[u][k], acc[r][k]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int k = 0; k < K; k++) red[wid][lane][r][k] = active ? acc[r][k] : T(0);
    __syncthreads();
    for (int e = threadIdx.x; e < h * K; e += blockDim.x) {
        const int row = e / K, k = e % K;
        const int ol = row / R, r = row % R;
        T sum = T(0);
        for (int w = 0; w < NW; w++)
            for (int gi = 0; gi < G; gi++) sum += red[w][gi * hl + ol][r][k];
        y[(P.y0 + row) * K + k] = sum;
    }
}

// lane s's value (s wave-uniform): a scalar read, no LDS round trip
__device__ __forceinline__ int rdlane(int v, int s) { return __builtin_amdgcn_readlane(v, s); }
__device__ __forceinline__ int64_t rdlane64(int64_t v, int s) {
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)v, s);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), s);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Vector-ALU panel product with 16-byte loads and a flattened walk, for
// leaves whose segment lists hold many short segments (small leaves, dense
// fallback blocks).  Lane (o, g) owns rows R o .. R o + R - 1 (R = 16 /
// sizeof(T)) in column group g of G = 64 / ceil(h / R); the leaf's segments
// are cut into steps of G columns and walked as one list, the 4 waves taking
// groups of kU steps in turn.  Each wave keeps up to 64 segment descriptors
// in registers (lane i: segment i of the current chunk -- its matrix and
// vector base addresses, its length and its first and end step), so a
// step's segment is a ballot, a popcount and six lane reads: no dependent
// load and no division per step (a run-time divide is a long scalar
// sequence, and the scalar unit, one per CU, is what bounds this walk).
// Reads past a column's last row go to rows >= h, which the reduction drops
// (the arenas carry tail padding).
template <typename T, int K>
__global__ __launch_bounds__(256) void panel_vecflat_kernel(const PanelItem *__restrict__ items,
                                                            const Segment *__restrict__ segs,
                                                            const T *__restrict__ NP, const T *__restrict__ UP,
                                                            const T *__restrict__ x, const T *__restrict__ t,
                                                            T *__restrict__ y) {
    // column steps in flight per lane: fewer for more RHS (registers); at 8
    // RHS FP32 keeps 4 (108 VGPRs: 256-problem batch panel 6.43 -> 5.95 ms),
    // FP64 2 (4 took 140 VGPRs and measured slower, 13.6 -> 14.0 ms)
    constexpr int R = 16 / (int)sizeof(T), NW = 4, kU = K <= 2 ? 8 : (K <= 4 || sizeof(T) == 4 ? 4 : 2);
    typedef T vec_t __attribute__((ext_vector_type(R), aligned(sizeof(T))));
    __shared__ T red[NW][kWave][R][K];
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const PanelItem P = items[xcd_remap(blockIdx.x, gridDim.x)];
    const int h = P.h;
    const int hl = (h + R - 1) / R;  // lanes per column
    const int G = kWave / hl;
    const int o = lane % hl, g = lane / hl;
    const bool active = g < G;
    const int64_t ld = P.ld;
    T acc[R][K];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[r][k] = T(0);
    // the current chunk of descriptors: lane i holds segment c0 + i
    int c0 = 0, cbase = 0, cend = 0;  // chunk's first step and one past its last
    int64_t d_m = 0, d_v = 0;         // the segment's matrix (row r0) and vector addresses
    int d_len = 0, d_qbeg = 0, d_qend = 0;
    auto load_chunk = [&]() {
        const int i = c0 + lane;
        int steps = 0;
        if (i < P.nseg) {
            const Segment S = segs[P.seg0 + i];
            d_m = (int64_t)((S.arena == 0 ? NP : UP) + S.mat + P.r0);
            d_v = (int64_t)((S.arena == 1 ? t : x) + S.vec * K);
            d_len = S.len;
            steps = (S.len + G - 1) / G;
        }
        int incl = steps;  // inclusive prefix sum over the wave
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const int v = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += v;
        }
        d_qbeg = cbase + incl - steps;
        d_qend = i < P.nseg ? cbase + incl : INT_MAX;
        cend = cbase + __shfl(incl, kWave - 1, kWave);
    };
    load_chunk();
    int nq = cend;  // steps of all segments (more chunks only past 64 segments)
    for (int cc = kWave; cc < P.nseg; cc += kWave) {
        const int i = cc + lane;
        int steps = i < P.nseg ? (segs[P.seg0 + i].len + G - 1) / G : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) steps += __shfl_xor(steps, off, kWave);
        nq += steps;
    }
    const int64_t lrow = (int64_t)o * R;
    for (int g0 = wid * kU; g0 < nq; g0 += NW * kU) {
        vec_t a[kU];
        T vv[kU][K];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int step = g0 + u;
            while (step >= cend && c0 + kWave < P.nseg) {  // next chunk (uniform)
                c0 += kWave;
                cbase = cend;
                load_chunk();
            }
            const bool in = step < nq;
            // the segment holding this step: lanes whose segment ends at or
            // before it come first
            const int s = in ? __builtin_amdgcn_readfirstlane(__popcll(__ballot(d_qend <= step))) : 0;
            const int len = rdlane(d_len, s);
            const int col = (step - rdlane(d_qbeg, s)) * G + g;
            const bool ok = in && active && col < len;
            // unpredicated global loads: a lane without a column reads the
            // arena's and x's first elements (always allocated); its matrix
            // values are zeroed, so it adds zeros times x's first (finite)
            // values -- a non-finite x reaches every row of the product anyway
            const T *M = ok ? reinterpret_cast<const T *>(rdlane64(d_m, s)) + lrow + (int64_t)col * ld : NP;
            const T *v = ok ? reinterpret_cast<const T *>(rdlane64(d_v, s)) + (int64_t)col * K : x;
            const vec_t av = *(const __attribute__((address_space(1))) vec_t *)M;
            a[u] = ok ? av : vec_t(T(0));
            if constexpr (K % R == 0) {  // the column's K values as K/R 16-byte loads
#pragma unroll
                for (int kk = 0; kk < K / R; kk++) {
                    const vec_t w = ((const __attribute__((address_space(1))) vec_t *)v)[kk];
#pragma unroll
                    for (int r = 0; r < R; r++) vv[u][kk * R + r] = w[r];
                }
            } else {
#pragma unroll
                for (int k = 0; k < K; k++) vv[u][k] = ((const __attribute__((address_space(1))) T *)v)[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int k = 0; k < K; k++) acc[r][k] = fma(a[u][r], vv[u][k], acc[r][k]);
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int k = 0; k < K; k++) red[wid][lane][r][k] = active ? acc[r][k] : T(0);
    __syncthreads();
    for (int e = threadIdx.x; e < h * K; e += blockDim.x) {
        const int row = e / K, k = e % K;
        const int ol = row / R, r = row % R;
        T sum = T(0);
        for (int w = 0; w < NW; w++)
            for (int gi = 0; gi < G; gi++) sum += red[w][gi * hl + ol][r][k];
        y[(P.y0 + row) * K + k] = sum;
    }
}

// Vector-ALU panel product, flattened walk.  Lane (o, g) of a wave owns row o of the tile
// and column group g (G = 64 / h groups of h lanes).  The leaf's segments
// are walked as one list of G-column steps (ceil(len / G) per segment); the
// 4 waves take consecutive groups of kU steps in turn, so kU loads stay in
// flight across segment boundaries.  The walk state is wave-uniform (scalar
// loads).  Per-lane partial sums are added through LDS in a fixed order.
template <typename T, int K>
__global__ __launch_bounds__(256) void panel_kernel(const PanelItem *__restrict__ items,
                                                    const Segment *__restrict__ segs,
                                                    const T *__restrict__ NP, const T *__restrict__ UP,
                                                    const T *__restrict__ x, const T *__restrict__ t,
                                                    T *__restrict__ y) {
    constexpr int NW = 4, kU = 8;
    __shared__ T red[NW][kWave][K];
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const PanelItem P = items[xcd_remap(blockIdx.x, gridDim.x)];
    const int h = P.h;
    const int G = kWave / h;
    const int o = lane % h, g = lane / h;
    const bool active = g < G;
    const int64_t ld = P.ld, nq = P.nqg;
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = T(0);
    int s = -1, len = 0;
    int64_t qbeg = 0, qend = 0;
    const T *Mb = nullptr, *vb = nullptr;
    for (int64_t g0 = (int64_t)wid * kU; g0 < nq; g0 += NW * kU) {
        T a[kU], vv[kU][K];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int64_t step = g0 + u;
            while (step >= qend && s + 1 < P.nseg) {
                s++;
                const Segment S = segs[P.seg0 + s];
                qbeg = qend;
                qend += (S.len + G - 1) / G;
                len = S.len;
                Mb = (S.arena == 0 ? NP : UP) + S.mat + P.r0;
                vb = (S.arena == 1 ? t : x) + S.vec * K;
            }
            const int col = (int)(step - qbeg) * G + g;
            const bool ok = active && step < nq && col < len;
            a[u] = ok ? load_stream(Mb + (int64_t)col * ld + o) : T(0);
#pragma unroll
            for (int k = 0; k < K; k++) vv[u][k] = ok ? vb[(int64_t)col * K + k] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] = fma(a[u], vv[u][k], acc[k]);
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

// ---------------------------------------------------------------------------
// Matrix-core forms for many right-hand sides.  With K >= 4 columns the
// products are small GEMMs: a 16x16x4 MFMA takes a 16-row x 4-column piece
// of the operator (one element per lane) and 4 rows x 16 RHS of the vector
// block, so one A load, one B load and one MFMA replace 4K vector FMAs and
// K broadcast loads per element.  Lane layouts of the 16x16x4 builtins:
// A[i][k] and B[k][j] come from lane i + 16k (resp. j + 16k); accumulator
// register r of lane l holds D[row(l, r)][l % 16].
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};
template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};

// Panel product on matrix cores.  The leaf tile's rows are cut into row
// tiles of 16 (NT <= 4 for h <= 64); the 4 waves are dealt over the row
// tiles (4 / NT waves each) and every wave walks its tile's segments as one
// list of 4-column quads, taking groups of kU consecutive quads in turn with
// the other waves of its tile, so loads stay in flight across the short
// near-field blocks.  Per quad a lane holds one operator element and one
// vector fragment per 16 RHS (vector fragments of a quad are re-read per
// row tile from cache).  Partial tiles are added through LDS in a fixed
// order.
template <typename T, int K>
__global__ __launch_bounds__(256) void panel_mfma_kernel(const PanelItem *__restrict__ items,
                                                         const Segment *__restrict__ segs,
                                                         const T *__restrict__ NP, const T *__restrict__ UP,
                                                         const T *__restrict__ x, const T *__restrict__ t,
                                                         T *__restrict__ y) {
    using MF = Mfma<T>;
    constexpr int NW = 4, NK = (K + 15) / 16, kU = sizeof(T) == 4 ? 16 : 8;
    __shared__ T red[NW][16][K];
    // the wave index is uniform: the segment walk stays in scalar registers
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int i16 = lane & 15, kq = lane >> 4;
    const PanelItem P = items[xcd_remap(blockIdx.x, gridDim.x)];
    const int h = P.h, NT = (h + 15) >> 4;
    const int wpt = NW / NT;                      // waves per row tile
    const int tt = wid / wpt, wsub = wid % wpt;   // this wave's tile and rank in it
    const int64_t ld = P.ld;
    if (tt < NT) {
        const int64_t qtot = P.nq4;
        typename MF::acc_t acc[NK];
#pragma unroll
        for (int n = 0; n < NK; n++) acc[n] = typename MF::acc_t{0, 0, 0, 0};
        const int r = tt * 16 + i16;
        const bool rok = r < h;
        // the wave's position in the segment list
        int s = -1, len = 0;
        int64_t qbeg = 0, qend = 0;
        const T *Mb = nullptr, *vb = nullptr;
        for (int64_t g0 = (int64_t)wsub * kU; g0 < qtot; g0 += (int64_t)wpt * kU) {
            T a[kU], b[kU][NK];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int64_t quad = g0 + u;
                while (quad >= qend && s + 1 < P.nseg) {
                    s++;
                    const Segment S = segs[P.seg0 + s];
                    qbeg = qend;
                    qend += (S.len + 3) >> 2;
                    len = S.len;
                    Mb = (S.arena == 0 ? NP : UP) + S.mat + P.r0;
                    vb = (S.arena == 1 ? t : x) + S.vec * K;
                }
                const int col = (int)(quad - qbeg) * 4 + kq;
                const bool ok = quad < qtot && col < len;
                a[u] = (ok && rok) ? Mb[(int64_t)col * ld + r] : T(0);
#pragma unroll
                for (int n = 0; n < NK; n++) {
                    const int j = n * 16 + i16;
                    b[u][n] = (ok && j < K) ? vb[(int64_t)col * K + j] : T(0);
                }
            }
#pragma unroll
            for (int u = 0; u < kU; u++)
#pragma unroll
                for (int n = 0; n < NK; n++) acc[n] = MF::mma(a[u], b[u][n], acc[n]);
        }
#pragma unroll
        for (int n = 0; n < NK; n++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = n * 16 + i16;
                if (j < K) red[wid][MF::row(lane, q)][j] = acc[n][q];
            }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < h * K; e += blockDim.x) {
        const int row = e / K, k = e % K;
        const int w0 = (row >> 4) * wpt;
        T sum = red[w0][row & 15][k];
        for (int w = 1; w < wpt; w++) sum += red[w0 + w][row & 15][k];
        y[(P.y0 + row) * K + k] = sum;
    }
}

// Geometry of the LDS-staged panel product for waves of RT row tiles of 16.
template <typename T, int K, int RT>
struct WideGeo {
    static constexpr int E = 16 / (int)sizeof(T);     // elements per 16-byte load
    static constexpr int ROWS = 16 * RT;              // rows per wave
    static constexpr int LPQ = ROWS / E;              // lanes per quad: a column's rows in one load
    static constexpr int QG = kWave / LPQ;            // quads per group
    static constexpr int PB = 4 * K / E;              // 16-byte pieces of a quad's vector values
    static constexpr int VB = (PB + LPQ - 1) / LPQ;   // their loads per lane
    static constexpr int AQS = 4 * ROWS + E, BQS = 4 * K + E;  // LDS quad strides (padded against bank conflicts)
    static constexpr int WS = QG * (AQS + BQS);       // a wave's stage
};

// One leaf row tile of the LDS-staged panel product with RT row tiles of 16
// per wave.  Wave w owns WSTR elements of LDS at stage + w WSTR: its stage
// while it walks the quads, then its ROWS x K partial sums (no other wave
// touches them before the workgroup barrier).
template <typename T, int K, int RT, int WSTR>
__device__ __forceinline__ void panel_w_tile(const PanelItem &P, const QuadDesc *__restrict__ quads,
                                             const T *__restrict__ NP, const T *__restrict__ UP,
                                             const T *__restrict__ x, const T *__restrict__ t, T *__restrict__ y,
                                             T *stage, int wid, int lane) {
    using MF = Mfma<T>;
    using G = WideGeo<T, K, RT>;
    constexpr int NW = 4, NK = (K + 15) / 16, E = G::E, ROWS = G::ROWS, LPQ = G::LPQ, QG = G::QG;
    constexpr int PB = G::PB, VB = G::VB, AQS = G::AQS, BQS = G::BQS;
    static_assert(K >= 4 && K % E == 0 && G::WS <= WSTR && ROWS * K <= WSTR, "K");
    typedef T vec_t __attribute__((ext_vector_type(E), aligned(sizeof(T))));
    typedef T lvec_t __attribute__((ext_vector_type(E)));
    const int i16 = lane & 15, kq = lane >> 4;
    const int h = P.h, NT = (h + ROWS - 1) / ROWS;
    const int wpt = NW / NT;
    const int tt = wid / wpt, wsub = wid % wpt;
    const int64_t ld = P.ld;
    const int uq = lane / LPQ, part = lane % LPQ;  // this lane's quad of a group, and its share of it
    if (tt < NT) {
        typename MF::acc_t acc[RT][NK];
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int n = 0; n < NK; n++) acc[r][n] = typename MF::acc_t{0, 0, 0, 0};
        const int64_t qtot = P.nq4;
        const QuadDesc *qd = quads + P.q0;
        T *const As = stage + wid * WSTR;
        T *const Bs = As + QG * AQS;
        T *as = As + uq * AQS + part * E;
        T *bs = Bs + uq * BQS;
        for (int64_t g0 = (int64_t)wsub * QG; g0 < qtot; g0 += (int64_t)wpt * QG) {
            const int64_t quad = g0 + uq;
            const bool qok = quad < qtot;
            const QuadDesc Q = qd[qok ? quad : 0];
            const int kind = Q.info & 3, ncol = qok ? Q.info >> 8 : 0;
            const T *Mc = (kind == 0 ? NP : UP) + Q.mat + P.r0 + tt * ROWS + part * E;
            const T *vc = (kind == 1 ? t : x) + (int64_t)Q.vec * K;
            vec_t av[4], bv[VB];
#pragma unroll
            for (int m = 0; m < 4; m++) av[m] = m < ncol ? *reinterpret_cast<const vec_t *>(Mc + m * ld) : vec_t(T(0));
#pragma unroll
            for (int m = 0; m < VB; m++) {
                const int pc = m * LPQ + part;  // piece of the quad's 4 x K values
                bv[m] = pc * E < ncol * K ? *reinterpret_cast<const vec_t *>(vc + pc * E) : vec_t(T(0));
            }
#pragma unroll
            for (int m = 0; m < 4; m++) *reinterpret_cast<lvec_t *>(as + m * ROWS) = av[m];
#pragma unroll
            for (int m = 0; m < VB; m++) {
                const int pc = m * LPQ + part;
                if (PB % LPQ == 0 || pc < PB) *reinterpret_cast<lvec_t *>(bs + pc * E) = bv[m];
            }
            // the wave's LDS operations run in order: a compiler fence is all
            // the stage needs between its lanes' writes and reads
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < QG; u++) {
                T b[NK];
#pragma unroll
                for (int n = 0; n < NK; n++) {
                    const int j = n * 16 + i16;
                    b[n] = j < K ? Bs[u * BQS + kq * K + j] : T(0);
                }
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    const T a = As[u * AQS + kq * ROWS + r * 16 + i16];
#pragma unroll
                    for (int n = 0; n < NK; n++) acc[r][n] = MF::mma(a, b[n], acc[r][n]);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int n = 0; n < NK; n++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int j = n * 16 + i16;
                    if (j < K) As[(r * 16 + MF::row(lane, q)) * K + j] = acc[r][n][q];
                }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < h * K; e += blockDim.x) {
        const int row = e / K, k = e % K;
        const int w0 = (row / ROWS) * wpt, rk = (row % ROWS) * K + k;
        T sum = stage[w0 * WSTR + rk];
        for (int w = 1; w < wpt; w++) sum += stage[(w0 + w) * WSTR + rk];
        y[(P.y0 + row) * K + k] = sum;
    }
}

// Matrix-core panel product with 16-byte operand loads.  In
// panel_mfma_kernel every A and B fragment is its own 4-byte-per-lane load:
// 32 memory instructions per 16 quads, and the kernel streams at ~3.4 TB/s
// on any operator (cutting the vector-ALU form's instruction count the
// same way made it 1.4x faster).  Here a wave stages a group of quads at a
// time into its own LDS area and reads the MFMA fragments from there.  The
// LPQ lanes of a quad load its descriptor, then, with one 16-byte load per
// lane and instruction, column m's rows of the wave (one contiguous run)
// for m = 0..3, and its 4 x K vector values (also contiguous): 1 + 4 +
// ceil(4K / (16 LPQ)) loads per lane and group instead of 32.  An FP32 wave
// takes 32 rows (two MFMA row tiles: whole 128-byte column runs, and each
// quad's vector values loaded once for both) when the rows it reads past
// the leaf tile stay within the arena's 16-element tail padding, else 16;
// an FP64 wave 16 (128 bytes).  Rows past the tile are read (the next
// column or the padding) and land only in discarded accumulator rows.
// A wave's partial sums replace its stage after its walk (the smaller LDS
// footprint measured faster than a separate buffer).  (Issuing the next
// group's loads before the current group's products measured slower.)
template <typename T, int K>
__global__ __launch_bounds__(256) void panel_mfma_w_kernel(const PanelItem *__restrict__ items,
                                                           const QuadDesc *__restrict__ quads,
                                                           const T *__restrict__ NP, const T *__restrict__ UP,
                                                           const T *__restrict__ x, const T *__restrict__ t,
                                                           T *__restrict__ y) {
    constexpr bool kPair = sizeof(T) == 4;
    constexpr int WS1 = WideGeo<T, K, 1>::WS, WS2 = WideGeo<T, K, 2>::WS;
    constexpr int WS = kPair && WS2 > WS1 ? WS2 : WS1;  // per wave
    static_assert(kArenaPad >= 15, "row overread");
    __shared__ __align__(16) T stage[4 * WS];
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const PanelItem P = items[xcd_remap(blockIdx.x, gridDim.x)];
    const int h = P.h;
    if constexpr (kPair) {
        if (h > 16 && (h + 31) / 32 * 32 - h <= 15) {
            panel_w_tile<T, K, 2, WS>(P, quads, NP, UP, x, t, y, stage, wid, lane);
            return;
        }
    }
    panel_w_tile<T, K, 1, WS>(P, quads, NP, UP, x, t, y, stage, wid, lane);
}

// V^T x on matrix cores.  A (tile, row chunk) item is split over 4
// workgroups of 64 columns; wave w of quarter q owns the 16 columns
// 64q + 16w .. + 15 over all rows of the chunk, so no cross-wave reduction
// is needed and short items still put 4 x 4 waves on the tile.
template <typename T, int K>
__global__ __launch_bounds__(256) void proj_mfma_kernel(const ProjItem *__restrict__ items, int64_t nitems,
                                                        const T *__restrict__ VP, const T *__restrict__ x,
                                                        T *__restrict__ t, T *__restrict__ part,
                                                        const int32_t *__restrict__ tidx) {
    using MF = Mfma<T>;
    constexpr int NK = (K + 15) / 16, kU = 8;
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int i16 = lane & 15, kq = lane >> 4;
    const ProjItem P = items[blockIdx.x >> 2];
    const int c0 = (blockIdx.x & 3) * 64 + wid * 16;
    if (c0 >= P.w) return;
    const T *base = VP + P.vp;
    const T *xv = x + P.x0 * K;
    const int64_t ld = P.ld;
    const int col = c0 + i16;
    const bool cok = col < P.w;
    typename MF::acc_t acc[NK];
#pragma unroll
    for (int n = 0; n < NK; n++) acc[n] = typename MF::acc_t{0, 0, 0, 0};
    for (int j0 = 0; j0 < P.rows; j0 += 4 * kU) {
        T a[kU], b[kU][NK];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int row = j0 + 4 * u + kq;
            const bool ok = row < P.rows;
            a[u] = (ok && cok) ? base[(int64_t)row * ld + col] : T(0);
#pragma unroll
            for (int n = 0; n < NK; n++) {
                const int j = n * 16 + i16;
                b[u][n] = (ok && j < K) ? xv[(int64_t)row * K + j] : T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
#pragma unroll
            for (int n = 0; n < NK; n++) acc[n] = MF::mma(a[u], b[u][n], acc[n]);
    }
#pragma unroll
    for (int n = 0; n < NK; n++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int oc = c0 + MF::row(lane, q), j = n * 16 + i16;
            if (oc < P.w && j < K) {
                if (P.out >= 0) part[(P.out * kProjTile + oc) * K + j] = acc[n][q];
                else t[(int64_t)tidx[P.tcol + oc] * K + j] = acc[n][q];
            }
        }
}

template <typename T, int K>
hipError_t launch_matvec_k(const ProjItem *pi, int64_t npi, const ProjItem *pw, int64_t npw_big,
                           const ProjReduce *pr, int64_t npr,
                           const PanelItem *pa, int64_t npa,
                           const Segment *segs, const QuadDesc *quads, const T *VP, const T *NP, const T *UP,
                           const T *x, T *t, T *part, const int32_t *tidx, T *y,
                           int variant, hipEvent_t *ev, hipStream_t s) {
    constexpr bool kMfma = K >= 4;
    if (ev) (void)hipEventRecord(ev[0], s);
    if constexpr (sizeof(T) == 8 && K > 16) {
        // FP64 with 32 RHS: matrix-core forms only (the vector-ALU forms'
        // accumulators and reduction buffers do not fit)
        if (npi) proj_mfma_kernel<T, K><<<dim3((unsigned)(4 * npi)), 256, 0, s>>>(pi, npi, VP, x, t, part, tidx);
        if (npr) proj_reduce_kernel<T, K><<<dim3((unsigned)npr), 256, 0, s>>>(pr, part, t, tidx);
        if (ev) (void)hipEventRecord(ev[1], s);
        if (npa) {
            if (variant & kMatvecPanelWide)
                panel_mfma_w_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, quads, NP, UP, x, t, y);
            else
                panel_mfma_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, segs, NP, UP, x, t, y);
        }
        if (ev) (void)hipEventRecord(ev[2], s);
        return hipGetLastError();
    } else {
        if (npi) {
            bool done = false;
            if constexpr (kMfma)
                if (variant & kMatvecProjMfma) {
                    proj_mfma_kernel<T, K><<<dim3((unsigned)(4 * npi)), 256, 0, s>>>(pi, npi, VP, x, t, part, tidx);
                    done = true;
                }
            if (!done && (variant & kMatvecProjWave)) {
                // the same items, the long ones first: one workgroup each,
                // then the short ones four per workgroup
                if (npw_big)
                    proj_kernel<T, K><<<dim3((unsigned)npw_big), 256, 0, s>>>(pw, npw_big, VP, x, t, part, tidx);
                if (npi > npw_big)
                    proj_wave_kernel<T, K><<<dim3((unsigned)((npi - npw_big + 3) / 4)), 256, 0, s>>>(
                        pw + npw_big, npi - npw_big, VP, x, t, part, tidx);
                done = true;
            }
            if constexpr (K <= 2)
                if (!done && (variant & kMatvecProjPack)) {
                    proj_pack_kernel<T, K><<<dim3((unsigned)npi), 256, 0, s>>>(pi, npi, VP, x, t, part, tidx);
                    done = true;
                }
            if (!done) proj_kernel<T, K><<<dim3((unsigned)npi), 256, 0, s>>>(pi, npi, VP, x, t, part, tidx);
        }
        if (npr) proj_reduce_kernel<T, K><<<dim3((unsigned)npr), 256, 0, s>>>(pr, part, t, tidx);
        if (ev) (void)hipEventRecord(ev[1], s);
        if (npa) {
            bool done = false;
            if constexpr (kMfma)
                if (variant & kMatvecPanelMfma) {
                    if (variant & kMatvecPanelWide)
                        panel_mfma_w_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, quads, NP, UP, x, t, y);
                    else
                        panel_mfma_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, segs, NP, UP, x, t, y);
                    done = true;
                }
            if constexpr (K <= 8)
                if (!done && (variant & kMatvecPanelVec)) {
                    if (variant & kMatvecPanelFlat)
                        panel_vecflat_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, segs, NP, UP, x, t, y);
                    else
                        panel_vec_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, segs, NP, UP, x, t, y);
                    done = true;
                }
            if (!done) {
                if (variant & kMatvecPanelFlat)
                    panel_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, segs, NP, UP, x, t, y);
                else
                    panel_seg_kernel<T, K><<<dim3((unsigned)npa), 256, 0, s>>>(pa, segs, NP, UP, x, t, y);
            }
        }
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_matvec(int nrhs, const ProjItem *pi, int64_t npi, const ProjItem *pw, int64_t npw_big,
                         const ProjReduce *pr, int64_t npr,
                         const PanelItem *pa, int64_t npa,
                         const Segment *segs, const QuadDesc *quads, const T *VP, const T *NP, const T *UP,
                         const T *x, T *t, T *part, const int32_t *tidx, T *y, int variant, hipEvent_t *ev,
                         hipStream_t s) {
#define HODLR_MV(KK)                                                                                 \
    case KK:                                                                                         \
        return launch_matvec_k<T, KK>(pi, npi, pw, npw_big, pr, npr, pa, npa, segs, quads, VP, NP, UP, x, t,  \
                                      part, tidx, y, variant, ev, s);
    switch (nrhs) {
        HODLR_MV(1)
        HODLR_MV(2)
        HODLR_MV(4)
        HODLR_MV(8)
        HODLR_MV(16)
    case 32:  // FP64: matrix-core forms only
        return launch_matvec_k<T, 32>(pi, npi, pw, npw_big, pr, npr, pa, npa, segs, quads, VP, NP, UP, x, t, part,
                                      tidx, y,
                                      variant, ev, s);
    default:
        return hipErrorInvalidValue;
    }
#undef HODLR_MV
}

template hipError_t launch_matvec<double>(int, const ProjItem *, int64_t, const ProjItem *, int64_t,
                                          const ProjReduce *, int64_t,
                                          const PanelItem *, int64_t,
                                          const Segment *, const QuadDesc *, const double *, const double *,
                                          const double *,
                                          const double *, double *, double *, const int32_t *, double *,
                                          int, hipEvent_t *, hipStream_t);
template hipError_t launch_matvec<float>(int, const ProjItem *, int64_t, const ProjItem *, int64_t,
                                         const ProjReduce *, int64_t,
                                         const PanelItem *, int64_t,
                                         const Segment *, const QuadDesc *, const float *, const float *,
                                         const float *,
                                         const float *, float *, float *, const int32_t *, float *,
                                         int, hipEvent_t *, hipStream_t);

}  // namespace hodlr
