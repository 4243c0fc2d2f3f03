// entry_kernels.hpp -- the kernels that evaluate matrix entries during
// assembly (gfx950), as templates over the entry function.
//
// Fn supplies f(r) through `static double f(const KernelDev &, double r)`;
// an entry is A_ij = diag (i == j), scale * f(|x_i - x_j|) [* w_j] otherwise
// (user_kernel.hpp:50-58,105-113,171-177).  The library instantiates them for
// its built-in families (FamilyFn); a caller compiled with hipcc instantiates
// them for its own Kernel_Fun and hands the launchers to
// hodlr_op_assemble_custom as an EntryLaunchers table.  Code including this
// header must be compiled with -ffp-contract=off (every fused multiply-add
// here is explicit, so entries and residuals round like the FP64 CPU
// restatement).
//
// ACA: one workgroup runs the partially pivoted ACA of one admissible block
// with the reference's control flow (LowRank_matrix.hpp:64-377): first pivot
// row 0, absolute pivot floor eps_ACA with up to ten retries (ends of the
// index range while rank < 3, then alternately a random remaining index and
// the midpoint of the widest gap of used indices), the squared-norm stop
// test, and the dense fallback when the rank reaches min(m, n) - 1.  Every
// thread runs the same control flow; block reductions broadcast pivots and
// norms.  The kernel runs twice per block: a rank pass that only records the
// rank (so the operator's exact footprint is known before it is allocated)
// and a fill pass that repeats the identical computation and writes the
// factors into their final panels.  Residual updates round the product,
// then the difference (v - u v, no fused multiply-add), as the reference's
// Eigen expressions do, so the GPU and the CPU pick the same pivots.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>

#include "aca_layout.h"
#include "device_common.h"

namespace hodlr {

namespace entry_detail {

template <int BS>
struct Shared {
    double v[BS / kWave], a[BS / kWave];
    int i[BS / kWave];
    double du[BS / kWave][kWave], dv[BS / kWave][kWave];
    int64_t cnt[BS];
    int bcast;
    int64_t bcast64;
};

__device__ __forceinline__ bool better(double aa, int ai, double ba, int bi) {
    return aa > ba || (aa == ba && ai < bi);
}

// argmax of |v| over the valid entries, lowest index on ties
// (kernel_function.hpp:55-68 scans ascending with a strict >)
template <int BS>
__device__ void block_argmax(Shared<BS> &sh, double &val, double &av, int &idx) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double ov = __shfl_xor(val, o, kWave);
        double oa = __shfl_xor(av, o, kWave);
        int oi = __shfl_xor(idx, o, kWave);
        if (better(oa, oi, av, idx)) { val = ov; av = oa; idx = oi; }
    }
    if constexpr (BS > kWave) {
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        if (lane == 0) { sh.v[wid] = val; sh.a[wid] = av; sh.i[wid] = idx; }
        __syncthreads();
        val = sh.v[0]; av = sh.a[0]; idx = sh.i[0];
        for (int w = 1; w < BS / kWave; w++)
            if (better(sh.a[w], sh.i[w], av, idx)) { val = sh.v[w]; av = sh.a[w]; idx = sh.i[w]; }
        __syncthreads();
    }
}

template <int BS>
__device__ double block_sum(Shared<BS> &sh, double v) {
    v = wave_sum(v);
    if constexpr (BS > kWave) {
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        if (lane == 0) sh.v[wid] = v;
        __syncthreads();
        v = sh.v[0];
        for (int w = 1; w < BS / kWave; w++) v += sh.v[w];
        __syncthreads();
    }
    return v;
}

template <int BS>
__device__ int block_min_int(Shared<BS> &sh, int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    if constexpr (BS > kWave) {
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        if (lane == 0) sh.i[wid] = v;
        __syncthreads();
        v = sh.i[0];
        for (int w = 1; w < BS / kWave; w++) v = min(v, sh.i[w]);
        __syncthreads();
    }
    return v;
}

template <int BS>
__device__ int block_max_int(Shared<BS> &sh, int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    if constexpr (BS > kWave) {
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        if (lane == 0) sh.i[wid] = v;
        __syncthreads();
        v = sh.i[0];
        for (int w = 1; w < BS / kWave; w++) v = max(v, sh.i[w]);
        __syncthreads();
    }
    return v;
}

template <int BS>
__device__ int first_unused(Shared<BS> &sh, const unsigned char *used, int len) {
    int best = INT_MAX;
    for (int i = threadIdx.x; i < len; i += BS)
        if (!used[i]) { best = i; break; }
    best = block_min_int(sh, best);
    return best == INT_MAX ? 0 : best;
}

template <int BS>
__device__ int last_unused(Shared<BS> &sh, const unsigned char *used, int len) {
    int best = -1;
    for (int i = len - 1 - (int)threadIdx.x; i >= 0; i -= BS)
        if (!used[i]) { best = i; break; }
    best = block_max_int(sh, best);
    return best < 0 ? 0 : best;
}

// k-th (0-based) unused index in ascending order (std::advance over a set)
template <int BS>
__device__ int kth_unused(Shared<BS> &sh, const unsigned char *used, int len, int64_t k) {
    const int tid = threadIdx.x;
    const int chunk = (len + BS - 1) / BS;
    const int b = tid * chunk, e = min(len, b + chunk);
    int64_t c = 0;
    for (int i = b; i < e; i++) c += used[i] ? 0 : 1;
    sh.cnt[tid] = c;
    __syncthreads();
    if (tid == 0) {
        int64_t run = 0;
        int found = 0;
        for (int t = 0; t < BS; t++) {
            int64_t ct = sh.cnt[t];
            if (run + ct > k) {
                int64_t want = k - run;
                int bb = t * chunk, ee = min(len, bb + chunk);
                for (int i = bb; i < ee; i++) {
                    if (used[i]) continue;
                    if (want == 0) { found = i; break; }
                    want--;
                }
                break;
            }
            run += ct;
        }
        sh.bcast = found;
    }
    __syncthreads();
    int r = sh.bcast;
    __syncthreads();
    return r;
}

// midpoint of the widest gap of the used indices (LowRank_matrix.hpp:181-198)
template <int BS>
__device__ int mid_gap(Shared<BS> &sh, const int *ind, int cnt, int *tmp) {
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < cnt; i++) {
            int v = ind[i], p = i;
            while (p > 0 && tmp[p - 1] > v) { tmp[p] = tmp[p - 1]; p--; }
            tmp[p] = v;
        }
        unsigned int mx = 0;
        int at = 0;
        for (int i = 0; i + 1 < cnt; i++) {
            unsigned int d = (unsigned int)(tmp[i + 1] - tmp[i]);
            if (d > mx) { at = i; mx = d; }
        }
        sh.bcast = tmp[at] + (int)(mx / 2);
    }
    __syncthreads();
    int r = sh.bcast;
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint64_t splitmix64_next(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// entry() from coordinates already at hand (the wave form stages them in LDS)
template <class Fn>
__device__ __forceinline__ double entry_xy(const KernelDev &k, const double *a, const double *b, int64_t pi,
                                           int64_t pj, double wj) {
    if (pi == pj) return k.diag;
    double s = 0.0;
    for (int d = 0; d < k.ndim; d++) {
        double t = a[d] - b[d];
        double t2 = t * t;
        s = s + t2;
    }
    double r = sqrt(s);
    double f = Fn::f(k, r);
    double v = k.scale * f;
    if (k.W) v = v * wj;
    return v;
}

// wave argmax of |v| over candidate lanes, lowest index on ties
__device__ __forceinline__ void wave_argmax(double &val, double &av, int &idx) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double ov = __shfl_xor(val, o, kWave);
        double oa = __shfl_xor(av, o, kWave);
        int oi = __shfl_xor(idx, o, kWave);
        if (better(oa, oi, av, idx)) { val = ov; av = oa; idx = oi; }
    }
}

__device__ __forceinline__ uint64_t low_bits(int len) { return len >= 64 ? ~0ull : ((1ull << len) - 1ull); }

// k-th (0-based) set bit of mask, 0 when there is none
__device__ __forceinline__ int kth_bit(uint64_t mask, int64_t k) {
    for (int64_t t = 0; t < k && mask; t++) mask &= mask - 1;
    return mask ? __builtin_ctzll(mask) : 0;
}

// midpoint of the widest gap between the distinct indices in mask
// (LowRank_matrix.hpp:181-198: sorted list, first strictly widest gap)
__device__ __forceinline__ int mask_mid_gap(uint64_t mask) {
    if (!mask) return 0;
    int prev = __builtin_ctzll(mask), at = prev;
    unsigned int mx = 0;
    mask &= mask - 1;
    while (mask) {
        const int cur = __builtin_ctzll(mask);
        const unsigned int d = (unsigned int)(cur - prev);
        if (d > mx) { mx = d; at = prev; }
        prev = cur;
        mask &= mask - 1;
    }
    return at + (int)(mx / 2);
}

// v - a_0 b_0 - a_1 b_1 - ... - a_{r-1} b_{r-1}, q ascending, each product
// rounded before its difference (Eigen's row - u * v, LowRank_matrix.hpp:
// 130,240), a_q = A[q sa], b_q = B[q sb].  The loads of eight terms are
// issued before their eight updates, so a residual entry waits for r / 8
// memory round trips instead of r.
__device__ __forceinline__ double aca_residual(double v, const double *A, int64_t sa, const double *B, int64_t sb,
                                               int r) {
    int q = 0;
    for (; q + 8 <= r; q += 8) {
        double a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a[k] = A[(int64_t)(q + k) * sa];
            b[k] = B[(int64_t)(q + k) * sb];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) v = v - a[k] * b[k];
    }
    for (; q < r; q++) v = v - A[(int64_t)q * sa] * B[(int64_t)q * sb];
    return v;
}

template <class Fn, int BS>
__global__ __launch_bounds__(BS) void aca_kernel(const AcaBlock *__restrict__ blocks,
                                                 const KernelDev *__restrict__ Ks,
                                                 double tol, double *__restrict__ pool,
                                                 int32_t *__restrict__ out_rank,
                                                 int32_t *__restrict__ out_flags, int fill,
                                                 const FillDest *__restrict__ dest,
                                                 const int64_t *__restrict__ leaf_start,
                                                 const int32_t *__restrict__ leaf_cnt,
                                                 double *__restrict__ UP, double *__restrict__ VP,
                                                 int32_t *__restrict__ SK, AcaStage S) {
    __shared__ Shared<BS> sh;
    extern __shared__ double pts_lds[];
    const AcaBlock B = blocks[blockIdx.x];
    const KernelDev K = Ks[B.prob];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    const int m = B.m, n = B.n, cap = B.cap;
    const int mn = m < n ? m : n;
    // the block's points in LDS when they fit in S.pts_lds bytes: row then
    // column coordinates, column weights, row then column point ids, so a
    // residual row or column reads them without the order-map and
    // coordinate gathers (the same values: the entries are bit-identical)
    const int nd = K.ndim;
    const bool staged = (int64_t)(m + n) * (nd + 1) * 8 + (K.W ? (int64_t)n * 8 : 0) <= (int64_t)S.pts_lds;
    double *sxr = pts_lds, *sxc = sxr + (size_t)m * nd, *swc = sxc + (size_t)n * nd;
    int64_t *sgr = reinterpret_cast<int64_t *>(swc + (K.W ? n : 0)), *sgc = sgr + m;
    if (staged) {
        for (int i = tid; i < m; i += BS) {
            const int64_t p = B.ord[B.row0l + i];
            sgr[i] = p;
            for (int d = 0; d < nd; d++) sxr[(size_t)i * nd + d] = K.X[(p - K.pos0) * nd + d];
        }
        for (int j = tid; j < n; j += BS) {
            const int64_t p = B.ord[B.col0l + j];
            sgc[j] = p;
            for (int d = 0; d < nd; d++) sxc[(size_t)j * nd + d] = K.X[(p - K.pos0) * nd + d];
            if (K.W) swc[j] = K.W[p - K.pos0];
        }
    }
    double *U = pool + B.scratch;
    double *V = U + (size_t)m * cap;
    double *row = V + (size_t)n * cap;
    double *col = row + n;
    unsigned char *rused = reinterpret_cast<unsigned char *>(col + m);
    unsigned char *cused = rused + m;
    int *rind = reinterpret_cast<int *>(col + m + (m + n + 7) / 8);
    int *cind = rind + aca_icap(cap);
    int *tmp = cind + aca_icap(cap);

    for (int i = tid; i < m; i += BS) rused[i] = 0;
    for (int j = tid; j < n; j += BS) cused[j] = 0;
    __syncthreads();
    if (tid == 0) { rused[0] = 1; rind[0] = 0; }
    __syncthreads();
    int rrem = m - 1, crem = n, nri = 1, nci = 0;
    int r = 0, pivot = 0, flags = 0;
    int64_t nrow_ev = 0, ncol_ev = 0;  // rows and columns evaluated
    double maxv = 0.0, mnorm = 0.0, rnorm = 0.0, cnorm = 0.0, gamma = 0.0;
    uint64_t rng = B.seed;

    // residual row ri of the block (LowRank_matrix.hpp:126-133); argmax over
    // the columns not yet used as pivots
    auto eval_row = [&](int ri) {
        nrow_ev++;
        const int64_t pi = staged ? sgr[ri] : B.ord[B.row0l + ri];
        const double *xa = sxr + (size_t)ri * nd;
        double bv = 0.0, ba = -1.0;
        int bi = INT_MAX;
        for (int j = tid; j < n; j += BS) {
            double v = staged ? entry_xy<Fn>(K, xa, sxc + (size_t)j * nd, pi, sgc[j], K.W ? swc[j] : 0.0)
                              : entry_t<Fn>(K, pi, B.ord[B.col0l + j]);
            v = aca_residual(v, U + ri, m, V + j, n, r);
            row[j] = v;
            if (!cused[j]) {
                double a = fabs(v);
                if (better(a, j, ba, bi)) { ba = a; bv = v; bi = j; }
            }
        }
        block_argmax(sh, bv, ba, bi);
        if (bi == INT_MAX) {  // no column left: index 0, as *begin() of an empty set
            __syncthreads();
            bi = 0;
            bv = row[0];
        }
        maxv = bv;
        pivot = bi;
    };
    // residual column cj (LowRank_matrix.hpp:237-243)
    auto eval_col = [&](int cj) {
        ncol_ev++;
        const int64_t pj = staged ? sgc[cj] : B.ord[B.col0l + cj];
        const double *xb = sxc + (size_t)cj * nd;
        const double wb = staged && K.W ? swc[cj] : 0.0;
        double bv = 0.0, ba = -1.0;
        int bi = INT_MAX;
        for (int i = tid; i < m; i += BS) {
            double v = staged ? entry_xy<Fn>(K, sxr + (size_t)i * nd, xb, sgr[i], pj, wb)
                              : entry_t<Fn>(K, B.ord[B.row0l + i], pj);
            v = aca_residual(v, V + cj, n, U + i, m, r);
            col[i] = v;
            if (!rused[i]) {
                double a = fabs(v);
                if (better(a, i, ba, bi)) { ba = a; bv = v; bi = i; }
            }
        }
        block_argmax(sh, bv, ba, bi);
        if (bi == INT_MAX) {
            __syncthreads();
            bi = 0;
            bv = col[0];
        }
        maxv = bv;
        pivot = bi;
    };

    do {
        __syncthreads();
        eval_row(rind[nri - 1]);
        int count = 0;
        bool at_end = false, use_rand = true;
        while (fabs(maxv) < tol && count < kMaxTries && crem > 0 && rrem > 0) {  // :144-216
            nri--;
            int nr;
            if (r < 3) {
                nr = at_end ? last_unused(sh, rused, m) : first_unused(sh, rused, m);
                at_end = !at_end;
            } else {
                if (use_rand) nr = kth_unused(sh, rused, m, (int64_t)(splitmix64_next(rng) % (uint64_t)rrem));
                else nr = mid_gap(sh, rind, nri, tmp);
                use_rand = !use_rand;
            }
            __syncthreads();
            const bool was_used = rused[nr] != 0;
            __syncthreads();
            if (tid == 0) {
                rind[nri] = nr;
                rused[nr] = 1;
            }
            nri++;
            if (!was_used) rrem--;
            __syncthreads();
            eval_row(nr);
            count++;
        }
        if (count == kMaxTries || crem == 0 || rrem == 0) break;  // :220-225
        count = 0;
        __syncthreads();
        if (tid == 0) {
            cind[nci] = pivot;
            cused[pivot] = 1;
        }
        nci++;
        crem--;
        gamma = 1.0 / maxv;  // :233
        __syncthreads();
        eval_col(pivot);
        while (fabs(maxv) < tol && count < kMaxTries && crem > 0 && rrem > 0) {  // :245-316
            nci--;
            int nc;
            if (nci < 3) {
                // :256-258: *remaining_col_ind.end() reads the set's size in libstdc++
                nc = at_end ? crem : first_unused(sh, cused, n);
                at_end = !at_end;
            } else {
                if (use_rand) nc = kth_unused(sh, cused, n, (int64_t)(splitmix64_next(rng) % (uint64_t)crem));
                else nc = mid_gap(sh, cind, nci, tmp);
                use_rand = !use_rand;
            }
            __syncthreads();
            const bool was_used = cused[nc] != 0;
            __syncthreads();
            if (tid == 0) {
                cind[nci] = nc;
                cused[nc] = 1;
            }
            nci++;
            if (!was_used) crem--;
            __syncthreads();
            eval_col(nc);
            count++;
        }
        __syncthreads();
        {  // :318-319
            const bool was_used = rused[pivot] != 0;
            __syncthreads();
            if (tid == 0) {
                rind[nri] = pivot;
                rused[pivot] = 1;
            }
            nri++;
            if (!was_used) rrem--;
        }
        if (r >= cap) {
            if (cap < mn - 1) flags |= 2;  // rank budget exhausted
            break;
        }
        // u = gamma col, v = row and the norm estimate (:322-338)
        double rs = 0.0, cs = 0.0;
        for (int j = tid; j < n; j += BS) {
            double v = row[j];
            V[(size_t)r * n + j] = v;
            rs = rs + v * v;
        }
        for (int i = tid; i < m; i += BS) {
            double c = col[i];
            U[(size_t)r * m + i] = gamma * c;
            cs = cs + c * c;
        }
        rs = block_sum(sh, rs);
        cs = block_sum(sh, cs);
        rnorm = sqrt(rs);
        cnorm = sqrt(cs);
        mnorm += fabs(gamma * gamma * rs * cs);
        for (int q0 = 0; q0 < r; q0 += kWave) {
            const int qn = min(kWave, r - q0);
            const double *ur = U + (size_t)r * m, *vr = V + (size_t)r * n;
            int qq = 0;
            // eight q at a time: the same per-thread sums (i ascending from
            // tid) and wave trees as one at a time, their loads and
            // shuffles interleaved
            for (; qq + 8 <= qn; qq += 8) {
                double d[8];
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = 0.0;
                for (int i = tid; i < m; i += BS) {
                    const double x = ur[i];
#pragma unroll
                    for (int k = 0; k < 8; k++) d[k] = fma(U[(size_t)(q0 + qq + k) * m + i], x, d[k]);
                }
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = wave_sum(d[k]);
                if (lane == 0)
#pragma unroll
                    for (int k = 0; k < 8; k++) sh.du[wid][qq + k] = d[k];
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = 0.0;
                for (int j = tid; j < n; j += BS) {
                    const double x = vr[j];
#pragma unroll
                    for (int k = 0; k < 8; k++) d[k] = fma(V[(size_t)(q0 + qq + k) * n + j], x, d[k]);
                }
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = wave_sum(d[k]);
                if (lane == 0)
#pragma unroll
                    for (int k = 0; k < 8; k++) sh.dv[wid][qq + k] = d[k];
            }
            for (; qq < qn; qq++) {
                const double *uq = U + (size_t)(q0 + qq) * m;
                const double *vq = V + (size_t)(q0 + qq) * n;
                double du = 0.0, dv = 0.0;
                for (int i = tid; i < m; i += BS) du = fma(uq[i], ur[i], du);
                for (int j = tid; j < n; j += BS) dv = fma(vq[j], vr[j], dv);
                du = wave_sum(du);
                dv = wave_sum(dv);
                if (lane == 0) { sh.du[wid][qq] = du; sh.dv[wid][qq] = dv; }
            }
            __syncthreads();
            double contrib = 0.0;
            if (tid < qn) {
                double a = 0.0, b = 0.0;
                for (int w = 0; w < BS / kWave; w++) { a += sh.du[w][tid]; b += sh.dv[w][tid]; }
                contrib = 2.0 * fabs(a) * fabs(b);
            }
            __syncthreads();
            mnorm += block_sum(sh, contrib);
        }
        r++;
    } while ((double)(r * (m + n)) * rnorm * cnorm > fabs(maxv) * tol * mnorm && r < mn);  // :341-344

    const bool dense = r >= mn - 1;  // :348
    if (tid == 0) {
        out_rank[B.id] = r;
        out_flags[B.id] = flags | (dense ? 1 : 0);
        if (S.evals) S.evals[B.id] = nrow_ev * n + ncol_ev * m;
    }
    if (!fill && S.buf && B.stage && !dense && !(flags & 2)) {  // keep the factors or pivots
        const int64_t len = B.stage == 2 ? (int64_t)r : (int64_t)r * (m + n);
        if (tid == 0) {
            const int64_t o = (int64_t)atomicAdd(S.cursor, (unsigned long long)len);
            const int64_t off = o + len <= S.cap ? o : -1;
            S.off[B.id] = off;
            sh.bcast64 = off;
        }
        __syncthreads();
        const int64_t off = sh.bcast64;
        if (off >= 0 && B.stage == 2) {  // pivot positions, as the fill pass would write them
            int32_t *pv = reinterpret_cast<int32_t *>(S.buf + off);
            for (int q = tid; q < r; q += BS) {
                pv[q] = B.ord[B.row0l + rind[q]];
                pv[r + q] = B.ord[B.col0l + cind[q]];
            }
        } else if (off >= 0) {  // U and V columns 0..r-1 are contiguous in the scratch
            for (int64_t e = tid; e < (int64_t)r * m; e += BS) S.buf[off + e] = U[e];
            for (int64_t e = tid; e < (int64_t)r * n; e += BS) S.buf[off + (int64_t)r * m + e] = V[e];
        }
        return;
    }
    if (!fill || dense || (flags & 2)) return;
    __syncthreads();
    const FillDest D = dest[B.id];
    if (D.sk_row >= 0) {  // skeleton block: its pivot rows and columns, in step order
        for (int q = tid; q < r; q += BS) {
            SK[D.sk_row + q] = B.ord[B.row0l + rind[q]];
            SK[D.sk_col + q] = B.ord[B.col0l + cind[q]];
        }
        return;
    }
    // U columns into the row node's leaf-blocked panel
    for (int i = tid; i < m; i += BS) {
        const int64_t p = B.ord[B.row0l + i];
        const int64_t ls = leaf_start[p];
        const int64_t nc = leaf_cnt[p];
        const int64_t base = D.up_base + (int64_t)D.R * (ls - B.row0) + (p - ls);
        for (int q = 0; q < r; q++) UP[base + (int64_t)(D.coff + q) * nc] = U[(size_t)q * m + i];
    }
    // V columns into the column node's tile-major panel: tile i holds panel
    // columns [256 i, 256 i + tw) as an n x tw row-major block
    for (int j = tid; j < n; j += BS) {
        const int64_t row = B.ord[B.col0l + j] - B.col0;
        for (int q = 0; q < r; q++) {
            const int col = D.voff + q;
            const int tile = col / kProjTile;
            const int tw = min(kProjTile, D.T - tile * kProjTile);
            VP[D.vp_base + (int64_t)tile * kProjTile * n + row * tw + (col - tile * kProjTile)] =
                V[(size_t)q * n + j];
        }
    }
}

// ---------------------------------------------------------------------------
// Wave form of the same ACA for small blocks (max(m, n) <= 64 and
// min(m, n) <= CAP <= 32): one 64-lane wave per block, lane l owning row l
// and column l.  The factors live in registers (lane i holds U[0..r)[i],
// lane j holds V[0..r)[j]; a pivot's values are broadcast by shuffles), the
// block's point coordinates in LDS, and the used-index sets are 64-bit
// masks, so an ACA step has no global-memory round trip and no barrier.
// The control flow, pivot rules and every reduction (wave_sum trees, the
// same product-then-difference residual updates) are those of aca_kernel; the norm
// estimate adds the 2|<u_q,u_r>||<v_q,v_r>| terms one q at a time, in the
// order of LowRank_matrix.hpp:335-338.
template <class Fn, int CAP>
__global__ __launch_bounds__(64) void aca_wave_kernel(const AcaBlock *__restrict__ blocks,
                                                      const KernelDev *__restrict__ Ks, double tol,
                                                      int32_t *__restrict__ out_rank,
                                                      int32_t *__restrict__ out_flags, int fill,
                                                      const FillDest *__restrict__ dest,
                                                      const int64_t *__restrict__ leaf_start,
                                                      const int32_t *__restrict__ leaf_cnt,
                                                      double *__restrict__ UP, double *__restrict__ VP,
                                                      AcaStage S) {
    __shared__ double xr[kWave * kMaxDimDev], xc[kWave * kMaxDimDev];
    __shared__ int64_t prow[kWave];
    const AcaBlock B = blocks[blockIdx.x];
    const KernelDev K = Ks[B.prob];
    const int l = threadIdx.x;
    const int m = B.m, n = B.n, cap = B.cap;
    const int mn = m < n ? m : n;
    const int nd = K.ndim;
    // stage the block's points: row l and column l of lane l
    const int64_t pi = l < m ? (int64_t)B.ord[B.row0l + l] : -1;
    const int64_t pj = l < n ? (int64_t)B.ord[B.col0l + l] : -2;
    for (int d = 0; d < nd; d++) {
        xr[l * kMaxDimDev + d] = l < m ? K.X[(pi - K.pos0) * nd + d] : 0.0;
        xc[l * kMaxDimDev + d] = l < n ? K.X[(pj - K.pos0) * nd + d] : 0.0;
    }
    prow[l] = pi;
    const double wj = (K.W && l < n) ? K.W[pj - K.pos0] : 0.0;
    __syncthreads();

    double Ur[CAP], Vr[CAP];
#pragma unroll
    for (int q = 0; q < CAP; q++) { Ur[q] = 0.0; Vr[q] = 0.0; }
    uint64_t rused = 1ull, cused = 0ull;           // row 0 is the first pivot row
    const uint64_t rvalid = low_bits(m), cvalid = low_bits(n);
    // the pivot index lists rind / cind as multisets: lane k counts the
    // entries equal to k (the gap rule only needs the distinct values, but
    // a retry pops the last entry, which may also occur earlier)
    int rcnt = l == 0 ? 1 : 0, ccnt = 0;
    int rlast = 0;                                 // rind[nri - 1]
    int rrem = m - 1, crem = n, nci = 0;
    int r = 0, pivot = 0, flags = 0;
    double maxv = 0.0, mnorm = 0.0, rnorm = 0.0, cnorm = 0.0, gamma = 0.0;
    double rowv = 0.0, colv = 0.0;                 // residual row (lane j) and column (lane i)
    int64_t nrow_ev = 0, ncol_ev = 0;              // rows and columns evaluated
    uint64_t rng = B.seed;

    // residual row ri (LowRank_matrix.hpp:126-133)
    auto eval_row = [&](int ri) {
        nrow_ev++;
        const double *a = xr + ri * kMaxDimDev;
        double v = l < n ? entry_xy<Fn>(K, a, xc + l * kMaxDimDev, prow[ri], pj, wj) : 0.0;
#pragma unroll
        for (int q = 0; q < CAP; q++)
            if (q < r) {
                const double uq = __shfl(Ur[q], ri, kWave);
                v = v - uq * Vr[q];
            }
        rowv = v;
        double bv = 0.0, ba = -1.0;
        int bi = INT_MAX;
        if (l < n && !((cused >> l) & 1ull)) { bv = v; ba = fabs(v); bi = l; }
        wave_argmax(bv, ba, bi);
        if (bi == INT_MAX) {  // no column left: index 0, as *begin() of an empty set
            bi = 0;
            bv = __shfl(v, 0, kWave);
        }
        maxv = bv;
        pivot = bi;
    };
    // residual column cj (LowRank_matrix.hpp:237-243)
    auto eval_col = [&](int cj) {
        ncol_ev++;
        const double *b = xc + cj * kMaxDimDev;
        const int64_t pjc = __shfl(pj, cj, kWave);
        const double wc = __shfl(wj, cj, kWave);
        double v = l < m ? entry_xy<Fn>(K, xr + l * kMaxDimDev, b, pi, pjc, wc) : 0.0;
#pragma unroll
        for (int q = 0; q < CAP; q++)
            if (q < r) {
                const double vq = __shfl(Vr[q], cj, kWave);
                v = v - vq * Ur[q];
            }
        colv = v;
        double bv = 0.0, ba = -1.0;
        int bi = INT_MAX;
        if (l < m && !((rused >> l) & 1ull)) { bv = v; ba = fabs(v); bi = l; }
        wave_argmax(bv, ba, bi);
        if (bi == INT_MAX) {
            bi = 0;
            bv = __shfl(v, 0, kWave);
        }
        maxv = bv;
        pivot = bi;
    };

    do {
        eval_row(rlast);
        int count = 0;
        bool at_end = false, use_rand = true;
        while (fabs(maxv) < tol && count < kMaxTries && crem > 0 && rrem > 0) {  // :144-216
            int nr;
            const uint64_t freer = ~rused & rvalid;
            if (r < 3) {
                nr = at_end ? (freer ? 63 - __builtin_clzll(freer) : 0) : (freer ? __builtin_ctzll(freer) : 0);
                at_end = !at_end;
            } else {
                // the retried row was popped from the list: the gap search
                // sees the list without it
                if (use_rand) nr = kth_bit(freer, (int64_t)(splitmix64_next(rng) % (uint64_t)rrem));
                else nr = mask_mid_gap(__ballot(rcnt - (l == rlast ? 1 : 0) > 0));
                use_rand = !use_rand;
            }
            const bool was_used = (rused >> nr) & 1ull;
            rcnt += (l == nr ? 1 : 0) - (l == rlast ? 1 : 0);  // pop the failed row, push nr
            rlast = nr;
            rused |= 1ull << nr;
            if (!was_used) rrem--;
            eval_row(nr);
            count++;
        }
        if (count == kMaxTries || crem == 0 || rrem == 0) break;  // :220-225
        count = 0;
        ccnt += l == pivot ? 1 : 0;
        int clast = pivot;
        cused |= 1ull << pivot;
        nci++;
        crem--;
        gamma = 1.0 / maxv;  // :233
        eval_col(pivot);
        while (fabs(maxv) < tol && count < kMaxTries && crem > 0 && rrem > 0) {  // :245-316
            nci--;
            int nc;
            const uint64_t freec = ~cused & cvalid;
            if (nci < 3) {
                // :256-258: *remaining_col_ind.end() reads the set's size in libstdc++
                nc = at_end ? crem : (freec ? __builtin_ctzll(freec) : 0);
                at_end = !at_end;
            } else {
                if (use_rand) nc = kth_bit(freec, (int64_t)(splitmix64_next(rng) % (uint64_t)crem));
                else nc = mask_mid_gap(__ballot(ccnt - (l == clast ? 1 : 0) > 0));
                use_rand = !use_rand;
            }
            const bool was_used = (cused >> nc) & 1ull;
            ccnt += (l == nc ? 1 : 0) - (l == clast ? 1 : 0);  // pop the failed column, push nc
            clast = nc;
            cused |= 1ull << nc;
            nci++;
            if (!was_used) crem--;
            eval_col(nc);
            count++;
        }
        {  // :318-319
            const bool was_used = (rused >> pivot) & 1ull;
            rcnt += l == pivot ? 1 : 0;
            rlast = pivot;
            rused |= 1ull << pivot;
            if (!was_used) rrem--;
        }
        if (r >= cap) {
            if (cap < mn - 1) flags |= 2;  // rank budget exhausted
            break;
        }
        // u = gamma col, v = row and the norm estimate (:322-338)
        const double rs = wave_sum(l < n ? rowv * rowv : 0.0);
        const double cs = wave_sum(l < m ? colv * colv : 0.0);
        const double unew = l < m ? gamma * colv : 0.0;
        const double vnew = l < n ? rowv : 0.0;
#pragma unroll
        for (int q = 0; q < CAP; q++)
            if (q == r) { Ur[q] = unew; Vr[q] = vnew; }
        rnorm = sqrt(rs);
        cnorm = sqrt(cs);
        mnorm += fabs(gamma * gamma * rs * cs);
#pragma unroll
        for (int q = 0; q < CAP; q++)
            if (q < r) {
                const double du = wave_sum(Ur[q] * unew);
                const double dv = wave_sum(Vr[q] * vnew);
                mnorm += 2.0 * fabs(du) * fabs(dv);
            }
        r++;
    } while ((double)(r * (m + n)) * rnorm * cnorm > fabs(maxv) * tol * mnorm && r < mn);  // :341-344

    const bool dense = r >= mn - 1;  // :348
    if (l == 0) {
        out_rank[B.id] = r;
        out_flags[B.id] = flags | (dense ? 1 : 0);
        if (S.evals) S.evals[B.id] = nrow_ev * n + ncol_ev * m;
    }
    if (!fill && S.buf && B.stage && !dense && !(flags & 2)) {  // keep the factors
        const int64_t len = (int64_t)r * (m + n);
        unsigned long long o = 0;
        if (l == 0) o = atomicAdd(S.cursor, (unsigned long long)len);
        const int64_t ob = (int64_t)__shfl(o, 0, kWave);
        const int64_t off = ob + len <= S.cap ? ob : -1;
        if (l == 0) S.off[B.id] = off;
        if (off >= 0) {
#pragma unroll
            for (int q = 0; q < CAP; q++)
                if (q < r) {
                    if (l < m) S.buf[off + (int64_t)q * m + l] = Ur[q];
                    if (l < n) S.buf[off + (int64_t)r * m + (int64_t)q * n + l] = Vr[q];
                }
        }
        return;
    }
    if (!fill || dense || (flags & 2)) return;
    const FillDest D = dest[B.id];
    // skeleton blocks are routed to aca_kernel, which writes their pivots;
    // the host finds any pivot slot left unwritten
    if (D.sk_row >= 0) return;
    if (l < m) {  // U columns into the row node's leaf-blocked panel
        const int64_t ls = leaf_start[pi];
        const int64_t nc = leaf_cnt[pi];
        const int64_t base = D.up_base + (int64_t)D.R * (ls - B.row0) + (pi - ls);
#pragma unroll
        for (int q = 0; q < CAP; q++)
            if (q < r) UP[base + (int64_t)(D.coff + q) * nc] = Ur[q];
    }
    if (l < n) {  // V columns into the column node's tile-major panel
        const int64_t row = pj - B.col0;
#pragma unroll
        for (int q = 0; q < CAP; q++)
            if (q < r) {
                const int col = D.voff + q;
                const int tile = col / kProjTile;
                const int tw = min(kProjTile, D.T - tile * kProjTile);
                VP[D.vp_base + (int64_t)tile * kProjTile * n + row * tw + (col - tile * kProjTile)] = Vr[q];
            }
    }
}

// dense-fallback blocks written into the row node's U panel
template <class Fn>
__global__ __launch_bounds__(256) void dense_fill_kernel(const DenseBlock *__restrict__ blocks,
                                                         const KernelDev *__restrict__ Ks,
                                                         const int64_t *__restrict__ leaf_start,
                                                         const int32_t *__restrict__ leaf_cnt,
                                                         double *__restrict__ UP) {
    const DenseBlock B = blocks[blockIdx.x];
    const KernelDev K = Ks[B.prob];
    const int64_t tot = (int64_t)B.m * B.n;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        const int64_t i = e % B.m, q = e / B.m;
        const int64_t p = B.row0 + i;
        const int64_t ls = leaf_start[p], nc = leaf_cnt[p];
        UP[B.up_base + (int64_t)B.R * (ls - B.row0) + (int64_t)(B.coff + q) * nc + (p - ls)] =
            entry_t<Fn>(K, p, B.col0 + q);
    }
}

// near-field panel of one leaf: self block then neighbours (HDD_node.hpp:136-144)
template <class Fn>
__global__ __launch_bounds__(256) void near_fill_kernel(const NearLeaf *__restrict__ leaves,
                                                        const int64_t *__restrict__ colsrc,
                                                        const KernelDev *__restrict__ Ks,
                                                        double *__restrict__ NP) {
    const NearLeaf L = leaves[blockIdx.x];
    const KernelDev K = Ks[L.prob];
    const int64_t tot = (int64_t)L.nc * L.ncols;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        const int64_t i = e % L.nc, q = e / L.nc;
        NP[L.np_base + e] = entry_t<Fn>(K, L.row0 + i, colsrc[L.colsrc + q]);
    }
}

// Direct-sum rows of the dense operator: y[r] = sum_j A(rows[r], j) x[j]
// over the row's problem, FP64, entries evaluated on the fly (the check of
// main.cpp:42-57 on chosen rows).  One workgroup per row; fixed-order sums.
template <class Fn>
__global__ __launch_bounds__(256) void dense_rows_kernel(const int64_t *__restrict__ rows,
                                                         const int32_t *__restrict__ rprob,
                                                         const KernelDev *__restrict__ Ks,
                                                         const int64_t *__restrict__ pos_base,
                                                         const double *__restrict__ x, double *__restrict__ y) {
    __shared__ double red[4];
    const int64_t i = rows[blockIdx.x];
    const int p = rprob[blockIdx.x];
    const KernelDev K = Ks[p];
    const int64_t j0 = pos_base[p], j1 = pos_base[p + 1];
    double acc = 0.0;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) acc = fma(entry_t<Fn>(K, i, j), x[j], acc);
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) y[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace entry_detail

// ---- host launchers ----

template <class Fn>
hipError_t launch_aca_t(int bs, const AcaBlock *blocks, int64_t nblocks, const KernelDev *K, double tol,
                        double *pool, int32_t *rank, int32_t *flags, int fill, const FillDest *dest,
                        const int64_t *leaf_start, const int32_t *leaf_cnt, double *UP, double *VP, int32_t *SK,
                        const AcaStage &S, hipStream_t s) {
    using namespace entry_detail;
    if (nblocks == 0) return hipSuccess;
    dim3 g((unsigned)nblocks);
    const size_t lds = S.pts_lds > 0 ? (size_t)S.pts_lds : 0;
    switch (bs) {
    case 8:
        aca_wave_kernel<Fn, 8><<<g, 64, 0, s>>>(blocks, K, tol, rank, flags, fill, dest, leaf_start, leaf_cnt, UP, VP, S);
        break;
    case 16:
        aca_wave_kernel<Fn, 16><<<g, 64, 0, s>>>(blocks, K, tol, rank, flags, fill, dest, leaf_start, leaf_cnt, UP, VP,
                                                 S);
        break;
    case 32:
        aca_wave_kernel<Fn, 32><<<g, 64, 0, s>>>(blocks, K, tol, rank, flags, fill, dest, leaf_start, leaf_cnt, UP, VP,
                                                 S);
        break;
    case 64:
        aca_kernel<Fn, 64><<<g, 64, lds, s>>>(blocks, K, tol, pool, rank, flags, fill, dest, leaf_start, leaf_cnt, UP,
                                              VP, SK, S);
        break;
    case 256:
        aca_kernel<Fn, 256><<<g, 256, lds, s>>>(blocks, K, tol, pool, rank, flags, fill, dest, leaf_start, leaf_cnt, UP,
                                                VP, SK, S);
        break;
    default:
        aca_kernel<Fn, 1024><<<g, 1024, lds, s>>>(blocks, K, tol, pool, rank, flags, fill, dest, leaf_start, leaf_cnt,
                                                  UP, VP, SK, S);
        break;
    }
    return hipGetLastError();
}

template <class Fn>
hipError_t launch_dense_fill_t(const DenseBlock *b, int64_t nb, const KernelDev *K, const int64_t *leaf_start,
                               const int32_t *leaf_cnt, double *UP, hipStream_t s) {
    if (nb == 0) return hipSuccess;
    entry_detail::dense_fill_kernel<Fn><<<dim3((unsigned)nb), 256, 0, s>>>(b, K, leaf_start, leaf_cnt, UP);
    return hipGetLastError();
}

template <class Fn>
hipError_t launch_near_fill_t(const NearLeaf *l, int64_t nl, const int64_t *colsrc, const KernelDev *K, double *NP,
                              hipStream_t s) {
    if (nl == 0) return hipSuccess;
    entry_detail::near_fill_kernel<Fn><<<dim3((unsigned)nl), 256, 0, s>>>(l, colsrc, K, NP);
    return hipGetLastError();
}

template <class Fn>
hipError_t launch_dense_rows_t(const int64_t *rows, const int32_t *rprob, int64_t nrows, const KernelDev *K,
                               const int64_t *pos_base, const double *x, double *y, hipStream_t s) {
    if (nrows == 0) return hipSuccess;
    entry_detail::dense_rows_kernel<Fn><<<dim3((unsigned)nrows), 256, 0, s>>>(rows, rprob, K, pos_base, x, y);
    return hipGetLastError();
}

// the launcher table of Fn (one per program)
template <class Fn>
const EntryLaunchers *entry_launchers() {
    static const EntryLaunchers t{kEntryAbiVersion, entry_abi_layout(), &launch_aca_t<Fn>, &launch_dense_fill_t<Fn>,
                                  &launch_near_fill_t<Fn>, &launch_dense_rows_t<Fn>};
    return &t;
}

}  // namespace hodlr
