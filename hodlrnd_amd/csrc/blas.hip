// blas.hip -- FP64 vector kernels for the device GMRES (GMRES.hpp:160-303).
//
// Arnoldi uses classical Gram-Schmidt with one re-orthogonalisation (CGS2):
// two multi-dot passes h = V^T w and two updates w -= V h per iteration
// instead of the reference's i+1 dependent dot/axpy pairs (modified
// Gram-Schmidt, GMRES.hpp:236-240).  Dot products reduce in a fixed order:
// per-chunk workgroup sums, then one pass over the chunks.
#include "../../include/hodlrnd/device/aca_layout.h"
#include "../../include/hodlrnd/device/device_common.h"

namespace hodlr {

constexpr int kMaxBasis = 128;

// partial[p * nv + k] = sum over chunk p of V[k * ldv + i] * w[i]
__global__ __launch_bounds__(256) void mdot_kernel(const double *__restrict__ V, int64_t ldv, int nv,
                                                   const double *__restrict__ w, int64_t n, int64_t chunk,
                                                   double *__restrict__ partial) {
    __shared__ double red[4][kMaxBasis];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t b = (int64_t)blockIdx.x * chunk, e = min(n, b + chunk);
    for (int k = 0; k < nv; k++) {
        const double *vk = V + (int64_t)k * ldv;
        double s = 0.0;
        for (int64_t i = b + tid; i < e; i += 256) s = fma(vk[i], w[i], s);
        s = wave_sum(s);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
    for (int k = tid; k < nv; k += 256)
        partial[(int64_t)blockIdx.x * nv + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// one workgroup per output: out[k] = sum_p partial[p * nv + k], fixed order
__global__ __launch_bounds__(256) void mdot_finish_kernel(const double *__restrict__ partial, int np, int nv,
                                                          double *__restrict__ out, int accumulate) {
    __shared__ double red[4];
    const int k = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int p = tid; p < np; p += 256) s += partial[(int64_t)p * nv + k];
    s = wave_sum(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        const double v = (red[0] + red[1]) + (red[2] + red[3]);
        out[k] = accumulate ? out[k] + v : v;
    }
}

// w[i] += sign * sum_k h[k] V[k * ldv + i]
__global__ __launch_bounds__(256) void gemv_kernel(const double *__restrict__ V, int64_t ldv, int nv,
                                                   const double *__restrict__ h, double sign,
                                                   double *__restrict__ w, int64_t n) {
    __shared__ double hs[kMaxBasis];
    for (int k = threadIdx.x; k < nv; k += blockDim.x) hs[k] = sign * h[k];
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double acc = w[i];
        for (int k = 0; k < nv; k++) acc = fma(hs[k], V[(int64_t)k * ldv + i], acc);
        w[i] = acc;
    }
}

__global__ void scale_kernel(const double *__restrict__ src, double alpha, double *__restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i] * alpha;
}

__global__ void sub_kernel(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c,
                           int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c[i] = a[i] - b[i];
}

__global__ void add_kernel(double *__restrict__ a, const double *__restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = a[i] + b[i];
}

__global__ void to_f32_kernel(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

__global__ void to_f64_kernel(const float *__restrict__ a, double *__restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

// out = (double) y + diag * x: the FP32 operator's result plus its diagonal
// applied in FP64 (n rows of k interleaved columns)
__global__ void f32_result_kernel(const float *__restrict__ y, const double *__restrict__ x,
                                  const double *__restrict__ diag, double *__restrict__ out, int64_t n, int k) {
    const int64_t tot = n * k;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)y[i] + diag[i / k] * x[i];
}

// y += diag * x in FP32 (the public FP32 matvec)
__global__ void add_diag_f32_kernel(float *__restrict__ y, const float *__restrict__ x,
                                    const double *__restrict__ diag, int64_t n, int k) {
    const int64_t tot = n * k;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = fmaf((float)diag[i / k], x[i], y[i]);
}

// ---- launchers ----

static inline unsigned grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return (unsigned)g;
}

int64_t mdot_chunks(int64_t n) {
    int64_t p = (n + 4095) / 4096;
    if (p > 1024) p = 1024;
    if (p < 1) p = 1;
    return p;
}

hipError_t launch_mdot(const double *V, int64_t ldv, int nv, const double *w, int64_t n, double *partial,
                       double *out, int accumulate, hipStream_t s) {
    if (nv > kMaxBasis) return hipErrorInvalidValue;
    const int64_t np = mdot_chunks(n);
    const int64_t chunk = (n + np - 1) / np;
    mdot_kernel<<<dim3((unsigned)np), 256, 0, s>>>(V, ldv, nv, w, n, chunk, partial);
    mdot_finish_kernel<<<dim3((unsigned)nv), 256, 0, s>>>(partial, (int)np, nv, out, accumulate);
    return hipGetLastError();
}

hipError_t launch_gemv(const double *V, int64_t ldv, int nv, const double *h, double sign, double *w, int64_t n,
                       hipStream_t s) {
    if (nv > kMaxBasis) return hipErrorInvalidValue;
    gemv_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(V, ldv, nv, h, sign, w, n);
    return hipGetLastError();
}

hipError_t launch_scale(const double *src, double alpha, double *dst, int64_t n, hipStream_t s) {
    scale_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(src, alpha, dst, n);
    return hipGetLastError();
}

hipError_t launch_sub(const double *a, const double *b, double *c, int64_t n, hipStream_t s) {
    sub_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(a, b, c, n);
    return hipGetLastError();
}

hipError_t launch_add(double *a, const double *b, int64_t n, hipStream_t s) {
    add_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

hipError_t launch_to_f32(const double *a, float *b, int64_t n, hipStream_t s) {
    to_f32_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

hipError_t launch_to_f64(const float *a, double *b, int64_t n, hipStream_t s) {
    to_f64_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

}  // namespace hodlr

namespace hodlr {
hipError_t launch_f32_result(const float *y, const double *x, const double *diag, double *out, int64_t n, int k,
                             hipStream_t s) {
    f32_result_kernel<<<dim3(grid_for(n * k)), 256, 0, s>>>(y, x, diag, out, n, k);
    return hipGetLastError();
}
hipError_t launch_add_diag_f32(float *y, const float *x, const double *diag, int64_t n, int k, hipStream_t s) {
    add_diag_f32_kernel<<<dim3(grid_for(n * k)), 256, 0, s>>>(y, x, diag, n, k);
    return hipGetLastError();
}
}  // namespace hodlr

namespace hodlr {
// plain streaming copy, 16 bytes per lane per step: the bandwidth ceiling
// the matvec kernels are compared with, and the known-size kernel used to
// scale hardware byte counters
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4 *__restrict__ a, float4 *__restrict__ b,
                                                          int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {  // four loads in flight per lane
        const float4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        b[i] = v0;
        b[i + stride] = v1;
        b[i + 2 * stride] = v2;
        b[i + 3 * stride] = v3;
    }
    for (; i < n4; i += stride) b[i] = a[i];
}
hipError_t launch_stream_copy(const void *a, void *b, int64_t bytes, hipStream_t s) {
    stream_copy_kernel<<<dim3(4096), 256, 0, s>>>(static_cast<const float4 *>(a), static_cast<float4 *>(b), bytes / 16);
    return hipGetLastError();
}

// plain streaming read (the matvec kernels' other ceiling: they read far
// more than they write): each lane sums 16-byte pieces, eight in flight;
// the sum is stored only if it is NaN, which keeps the loads live
__global__ __launch_bounds__(256) void stream_read_kernel(const float4 *__restrict__ a, int64_t n4,
                                                          float *__restrict__ out) {
    constexpr int U = 8;
    const int64_t stride = (int64_t)gridDim.x * 256 * U;
    float s = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * 256 * U + threadIdx.x; i + (U - 1) * 256 < n4; i += stride) {
        float4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = a[i + u * 256];
#pragma unroll
        for (int u = 0; u < U; u++) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    if (s != s) out[0] = s;
}
hipError_t launch_stream_read(const void *a, int64_t bytes, float *out, hipStream_t s) {
    // 65536 workgroups measured fastest (6.39 TB/s on 16 GiB, against 6.21-6.30 with 2048-16384)
    stream_read_kernel<<<dim3(65536), 256, 0, s>>>(static_cast<const float4 *>(a), bytes / 16, out);
    return hipGetLastError();
}
}  // namespace hodlr

// ---- segmented (per problem, per column) kernels for batched solves ----
// Vectors hold n rows of K interleaved columns; rows are split into chunks
// that never cross a problem boundary.  Every reduction sums in a fixed
// order that depends only on the problem's own rows, so an instance gives
// the same bits alone or inside any batch.
namespace hodlr {

// partial[(chunk * nv + v) * K + c] = sum over the chunk of V_v[row, c] * w[row, c]
// The batched kernels below read n x K interleaved blocks (element (row, c)
// at row * K + c) with lanes mapped to columns: thread t owns column t % K
// and row offset t / K of each step of 256 / K rows, so one wave load reads
// 64 consecutive elements (512 bytes).  A column's arithmetic never depends
// on the other columns.

// sum of x over the lanes of a wave that share lane % K (K a power of two,
// at most 32), in a fixed butterfly order; every such lane gets the sum
template <int K>
__device__ __forceinline__ double col_sum(double x) {
#pragma unroll
    for (int m = 32; m >= K; m >>= 1) x += __shfl_xor(x, m);
    return x;
}

// partial[(chunk * nv + v) * K + c] = sum over the chunk's rows of V_v[row, c] * w[row, c].
// Basis vectors are taken G at a time, so w is read once per G of them; a
// thread takes R of its rows per step (in row order), so (G + 1) R loads are
// in flight even when an engine's few chunks put few waves on the chip.
template <int K>
__global__ __launch_bounds__(256) void bmdot_kernel(const double *const *__restrict__ V, int nv,
                                                    const double *__restrict__ w, const Chunk *__restrict__ chunks,
                                                    double *__restrict__ partial) {
    constexpr int RP = 256 / K, G = 8, R = 4;
    __shared__ double red[G][4][K];
    const Chunk C = chunks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, c = tid % K, ro = tid / K;
    for (int v0 = 0; v0 < nv; v0 += G) {
        const int gn = min(G, nv - v0);
        const double *vp[G];
#pragma unroll
        for (int g = 0; g < G; g++) vp[g] = V[v0 + min(g, gn - 1)];
        double acc[G];
#pragma unroll
        for (int g = 0; g < G; g++) acc[g] = 0.0;
        int r = ro;
        for (; r + (R - 1) * RP < C.rows; r += R * RP) {
            int64_t e[R];
            double wv[R];
#pragma unroll
            for (int q = 0; q < R; q++) {
                e[q] = (C.row0 + r + q * RP) * K + c;
                wv[q] = w[e[q]];
            }
#pragma unroll
            for (int g = 0; g < G; g++)
                if (g < gn) {
                    double a[R];
#pragma unroll
                    for (int q = 0; q < R; q++) a[q] = vp[g][e[q]];
#pragma unroll
                    for (int q = 0; q < R; q++) acc[g] = fma(a[q], wv[q], acc[g]);
                }
        }
        for (; r < C.rows; r += RP) {
            const int64_t e = (C.row0 + r) * K + c;
            const double wv = w[e];
#pragma unroll
            for (int g = 0; g < G; g++)
                if (g < gn) acc[g] = fma(vp[g][e], wv, acc[g]);
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const double sw = col_sum<K>(acc[g]);
            if (g < gn && lane < K) red[g][wid][lane] = sw;
        }
        __syncthreads();
        for (int q = tid; q < gn * K; q += 256) {
            const int g = q / K, cc = q % K;
            partial[((int64_t)blockIdx.x * nv + v0 + g) * K + cc] =
                (red[g][0][cc] + red[g][1][cc]) + (red[g][2][cc] + red[g][3][cc]);
        }
        __syncthreads();
    }
}

// out[(p * nv + v) * K + c] = sum of partials over problem p's chunks
__global__ __launch_bounds__(256) void bfinish_kernel(const double *__restrict__ partial,
                                                      const int32_t *__restrict__ pchunk, int nv, int K,
                                                      double *__restrict__ out) {
    const int p = blockIdx.x;
    const int c0 = pchunk[p], c1 = pchunk[p + 1];
    for (int e = threadIdx.x; e < nv * K; e += blockDim.x) {
        double s = 0.0;
        for (int ch = c0; ch < c1; ch++) s += partial[(int64_t)ch * nv * K + e];
        out[(int64_t)p * nv * K + e] = s;
    }
}

// w[row, c] += sign * sum_v coef[(prob * nv + v) * K + c] * V_v[row, c]
// (per element: w first, then the basis vectors in order).  The chunk's
// nv x K coefficients sit in dynamic LDS; a thread updates R of its rows at
// once, so R loads of each basis vector are in flight.
template <int K>
__global__ __launch_bounds__(256) void bgemv_kernel(const double *const *__restrict__ V, int nv,
                                                    const double *__restrict__ coef, double sign,
                                                    double *__restrict__ w, const Chunk *__restrict__ chunks) {
    constexpr int RP = 256 / K, R = 8;
    extern __shared__ double cs[];
    const Chunk C = chunks[blockIdx.x];
    for (int e = threadIdx.x; e < nv * K; e += blockDim.x) cs[e] = sign * coef[(int64_t)C.prob * nv * K + e];
    __syncthreads();
    const int c = threadIdx.x % K, ro = threadIdx.x / K;
    for (int r0 = ro; r0 < C.rows; r0 += R * RP) {
        int64_t e[R];
        double acc[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            const int r = r0 + q * RP < C.rows ? r0 + q * RP : r0;  // past the chunk: r0 again, not stored
            e[q] = (C.row0 + r) * K + c;
            acc[q] = w[e[q]];
        }
        for (int v = 0; v < nv; v++) {
            const double cf = cs[v * K + c];
            const double *vv = V[v];
#pragma unroll
            for (int q = 0; q < R; q++) acc[q] = fma(cf, vv[e[q]], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < R; q++)
            if (r0 + q * RP < C.rows) w[e[q]] = acc[q];
    }
}

// dst[row, c] = src[row, c] * coef[prob * K + c]
template <int K>
__global__ __launch_bounds__(256) void bscale_kernel(const double *__restrict__ src, const double *__restrict__ coef,
                                                     double *__restrict__ dst, const Chunk *__restrict__ chunks) {
    constexpr int RP = 256 / K;
    const Chunk C = chunks[blockIdx.x];
    const int c = threadIdx.x % K, ro = threadIdx.x / K;
    const double cf = coef[(int64_t)C.prob * K + c];
    for (int r = ro; r < C.rows; r += RP) {
        const int64_t e = (C.row0 + r) * K + c;
        dst[e] = src[e] * cf;
    }
}

#define HODLR_BATCH_K(MACRO) \
    MACRO(1)                 \
    MACRO(2)                 \
    MACRO(4)                 \
    MACRO(8)                 \
    MACRO(16)                \
    MACRO(32)

hipError_t launch_bmdot(const double *const *V, int nv, const double *w, const Chunk *chunks, int nchunk,
                        double *partial, const int32_t *pchunk, int nprob, int K, double *out, hipStream_t s) {
    if (nv < 0 || nv > 128) return hipErrorInvalidValue;
    switch (K) {
#define CASE(KK) \
    case KK: bmdot_kernel<KK><<<dim3((unsigned)nchunk), 256, 0, s>>>(V, nv, w, chunks, partial); break;
        HODLR_BATCH_K(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    bfinish_kernel<<<dim3((unsigned)nprob), 256, 0, s>>>(partial, pchunk, nv, K, out);
    return hipGetLastError();
}

hipError_t launch_bgemv(const double *const *V, int nv, const double *coef, double sign, double *w,
                        const Chunk *chunks, int nchunk, int K, hipStream_t s) {
    if (nv < 0 || nv > 128) return hipErrorInvalidValue;
    if (nv == 0) return hipSuccess;  // w unchanged
    switch (K) {
#define CASE(KK) \
    case KK:                                                                                               \
        bgemv_kernel<KK><<<dim3((unsigned)nchunk), 256, sizeof(double) * (size_t)nv * KK, s>>>(V, nv, coef, sign, w, \
                                                                                               chunks);              \
        break;
        HODLR_BATCH_K(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bscale(const double *src, const double *coef, double *dst, const Chunk *chunks, int nchunk, int K,
                         hipStream_t s) {
    switch (K) {
#define CASE(KK) \
    case KK: bscale_kernel<KK><<<dim3((unsigned)nchunk), 256, 0, s>>>(src, coef, dst, chunks); break;
        HODLR_BATCH_K(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


// ---- per-column kernels of the serving engine ----
// Every column of an n x K interleaved block is an independent solve at its
// own Krylov step; these move single columns between blocks.

// out[r, c] = src[c][r, c] (zero where src[c] is null)
__global__ __launch_bounds__(256) void gather_cols_kernel(const double *const *__restrict__ src,
                                                          double *__restrict__ out, int64_t n, int K) {
    const int64_t tot = n * K;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % K);
        const double *p = src[c];
        out[e] = p ? p[e] : 0.0;
    }
}

// dst[c][r, c] = w[r, c] * scale[c] where dst[c] is not null
__global__ __launch_bounds__(256) void scatter_cols_kernel(const double *__restrict__ w,
                                                           const double *__restrict__ scale,
                                                           double *const *__restrict__ dst, int64_t n, int K) {
    const int64_t tot = n * K;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % K);
        double *p = dst[c];
        if (p) p[e] = w[e] * scale[c];
    }
}

// cf[v * K + c] = 0 for v >= lim[c]
__global__ __launch_bounds__(256) void mask_coef_kernel(double *__restrict__ cf, int nv, int K,
                                                        const int32_t *__restrict__ lim) {
    for (int e = threadIdx.x; e < nv * K; e += blockDim.x)
        if (e / K >= lim[e % K]) cf[e] = 0.0;
}

// column c of an interleaved block to / from a contiguous vector
__global__ __launch_bounds__(256) void put_col_kernel(double *__restrict__ dst, const double *__restrict__ src,
                                                      int64_t n, int K, int c) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        dst[r * K + c] = src ? src[r] : 0.0;
}
__global__ __launch_bounds__(256) void get_col_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                      int64_t n, int K, int c) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        dst[r] = src[r * K + c];
}

hipError_t launch_gather_cols(const double *const *src, double *out, int64_t n, int K, hipStream_t s) {
    gather_cols_kernel<<<dim3(grid_for(n * K)), 256, 0, s>>>(src, out, n, K);
    return hipGetLastError();
}
hipError_t launch_scatter_cols(const double *w, const double *scale, double *const *dst, int64_t n, int K,
                               hipStream_t s) {
    scatter_cols_kernel<<<dim3(grid_for(n * K)), 256, 0, s>>>(w, scale, dst, n, K);
    return hipGetLastError();
}
hipError_t launch_mask_coef(double *cf, int nv, int K, const int32_t *lim, hipStream_t s) {
    mask_coef_kernel<<<dim3(1), 256, 0, s>>>(cf, nv, K, lim);
    return hipGetLastError();
}
hipError_t launch_put_col(double *dst, const double *src, int64_t n, int K, int c, hipStream_t s) {
    put_col_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(dst, src, n, K, c);
    return hipGetLastError();
}
hipError_t launch_get_col(const double *src, double *dst, int64_t n, int K, int c, hipStream_t s) {
    get_col_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(src, dst, n, K, c);
    return hipGetLastError();
}

}  // namespace hodlr

// ---- Givens rotations and the convergence test on the device ----
// GMRES.hpp:83-146,244-268 for one solve (gmres_*_kernel) and for every
// instance of a batch (bgmres_*_kernel, one thread per instance).  The
// Hessenberg matrix, rotations and rotated right-hand side stay in device
// memory; the host reads only each step's small state record.
namespace hodlr {

namespace {

// GMRES.hpp:83-110 (real rotations)
__device__ __forceinline__ void gen_rot_d(double dx, double dy, double &cs, double &sn) {
    if (dy == 0.0) {
        cs = 1.0;
        sn = 0.0;
    } else if (fabs(dy) > fabs(dx)) {
        const double t = dx / dy;
        sn = 1.0 / sqrt(1.0 + t * t);
        cs = t * sn;
    } else {
        const double t = dy / dx;
        cs = 1.0 / sqrt(1.0 + t * t);
        sn = t * cs;
    }
}

// GMRES.hpp:112-129
__device__ __forceinline__ void app_rot_d(double &dx, double &dy, double cs, double sn) {
    const double t = cs * dx + sn * dy;
    dy = (-sn) * dx + cs * dy;
    dx = t;
}

// one Arnoldi step's column: H(0..i, i) = h[k], H(i+1, i) = hn; rotations;
// the rotated right-hand side; the residual test.  Returns the scale of the
// next basis vector (0 when the solve ends here or hn is 0).
__device__ double givens_step(double *H, double *cs, double *sn, double *sv, int ld, int i, double hn,
                              GmresDevState &st) {
    double *col = H + (int64_t)i * ld;
    col[i + 1] = hn;
    for (int k = 0; k < i; k++) app_rot_d(col[k], col[k + 1], cs[k], sn[k]);
    gen_rot_d(col[i], col[i + 1], cs[i], sn[i]);
    app_rot_d(col[i], col[i + 1], cs[i], sn[i]);
    app_rot_d(sv[i], sv[i + 1], cs[i], sn[i]);
    st.resid = fabs(sv[i + 1]) / st.normb;
    st.steps++;
    if (!(hn == hn) || hn == INFINITY) {  // non-finite Arnoldi norm: the solve breaks down here
        st.active = 0;
        st.conv = -2;
        return 0.0;
    }
    if (st.resid < st.tol || st.steps >= st.budget) {  // GMRES.hpp:253-254, or the iteration budget
        st.active = 0;
        st.conv = i;
        return 0.0;
    }
    return hn != 0.0 ? 1.0 / hn : 0.0;
}

// y = H(0..k, 0..k)^-1 sv(0..k) by back substitution (GMRES.hpp:132-146)
__device__ void back_solve(const double *H, const double *sv, int ld, int k, double *y) {
    for (int a = 0; a <= k; a++) y[a] = sv[a];
    for (int a = k; a >= 0; a--) {
        y[a] /= H[(int64_t)a * ld + a];
        for (int q = a - 1; q >= 0; q--) y[q] -= H[(int64_t)a * ld + q] * y[a];
    }
}

}  // namespace

// one solve: h1 at hd[0..i], h2 at hd[128..128+i], |w|^2 at hd[256]
__global__ void gmres_step_kernel(double *__restrict__ H, double *__restrict__ cs, double *__restrict__ sn,
                                  double *__restrict__ sv, const double *__restrict__ hd, int i, int ld,
                                  GmresDevState *__restrict__ stp) {
    if (threadIdx.x != 0) return;
    GmresDevState st = *stp;
    if (!st.active) {  // ended at an earlier step: this one was launched ahead and changes nothing
        stp->scale = 0.0;
        return;
    }
    double *col = H + (int64_t)i * ld;
    for (int k = 0; k <= i; k++) col[k] = hd[k] + hd[128 + k];
    st.scale = givens_step(H, cs, sn, sv, ld, i, sqrt(hd[256]), st);
    *stp = st;
}

// start of a cycle: beta = ||r|| (r2 = ||r||^2), the residual test
// (GMRES.hpp:198 `<=` before the first cycle, :284 `<` at a restart), the
// rotated right-hand side beta e_1 and the scale of v_0 = r / beta
__global__ void gmres_cycle_kernel(const double *__restrict__ r2, double *__restrict__ sv, int ld, int first,
                                   GmresDevState *__restrict__ stp) {
    if (threadIdx.x != 0) return;
    GmresDevState st = *stp;
    const double beta = sqrt(*r2);
    st.resid = beta / st.normb;
    st.conv = -1;
    const bool done = first ? st.resid <= st.tol : st.resid < st.tol;
    st.active = (!done && st.steps < st.budget) ? 1 : 0;
    if (done) st.conv = -3;
    for (int a = 0; a < ld; a++) sv[a] = 0.0;
    sv[0] = beta;
    st.scale = beta != 0.0 ? 1.0 / beta : 0.0;
    *stp = st;
}

// ||b|| (b2 = ||b||^2; a zero b counts as 1, GMRES.hpp:192-193) and the
// fresh state of a solve
__global__ void gmres_init_kernel(const double *__restrict__ b2, double tol, int budget,
                                  GmresDevState *__restrict__ stp) {
    if (threadIdx.x != 0) return;
    GmresDevState st;
    const double nb = sqrt(*b2);
    st.normb = nb == 0.0 ? 1.0 : nb;
    st.tol = tol;
    st.resid = 0.0;
    st.scale = 0.0;
    st.active = 1;
    st.conv = -1;
    st.steps = 0;
    st.budget = budget;
    st.iters = 0;
    st.converged = 0;
    *stp = st;
}

// y for the update x += V y over columns 0..k
__global__ void gmres_backsolve_kernel(const double *__restrict__ H, const double *__restrict__ sv, int ld, int k,
                                       double *__restrict__ y) {
    if (threadIdx.x == 0) back_solve(H, sv, ld, k, y);
}

// dst = src * (*alpha): the next basis vector scaled by a value the device computed
__global__ void scale_by_kernel(const double *__restrict__ src, const double *__restrict__ alpha,
                                double *__restrict__ dst, int64_t n) {
    const double a = *alpha;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i] * a;
}

hipError_t launch_gmres_init(const double *b2, double tol, int budget, GmresDevState *st, hipStream_t s) {
    gmres_init_kernel<<<1, 64, 0, s>>>(b2, tol, budget, st);
    return hipGetLastError();
}
hipError_t launch_gmres_cycle(const double *r2, double *sv, int ld, int first, GmresDevState *st, hipStream_t s) {
    gmres_cycle_kernel<<<1, 64, 0, s>>>(r2, sv, ld, first, st);
    return hipGetLastError();
}
hipError_t launch_gmres_step(double *H, double *cs, double *sn, double *sv, const double *hd, int i, int ld,
                             GmresDevState *st, hipStream_t s) {
    if (i + 1 >= ld) return hipErrorInvalidValue;
    gmres_step_kernel<<<1, 64, 0, s>>>(H, cs, sn, sv, hd, i, ld, st);
    return hipGetLastError();
}
hipError_t launch_gmres_backsolve(const double *H, const double *sv, int ld, int k, double *y, hipStream_t s) {
    gmres_backsolve_kernel<<<1, 64, 0, s>>>(H, sv, ld, k, y);
    return hipGetLastError();
}
hipError_t launch_scale_by(const double *src, const double *alpha, double *dst, int64_t n, hipStream_t s) {
    scale_by_kernel<<<dim3(grid_for(n)), 256, 0, s>>>(src, alpha, dst, n);
    return hipGetLastError();
}

}  // namespace hodlr

// ---- the same for every instance of a batch (instance a = p * K + c) ----
namespace hodlr {

namespace {
// active instances of this workgroup -> out[0], breakdowns -> out[1]
__device__ void block_count(int act, int bad, int32_t *out) {
    __shared__ int sa[16], sb[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        act += __shfl_down(act, o, 64);
        bad += __shfl_down(bad, o, 64);
    }
    if (lane == 0) {
        sa[wid] = act;
        sb[wid] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tb = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
            ta += sa[k];
            tb += sb[k];
        }
        out[0] = ta;
        out[1] = tb;
    }
}
}  // namespace

__global__ __launch_bounds__(1024) void bgmres_init_kernel(int NI, const double *__restrict__ b2,
                                                           const double *__restrict__ tol,
                                                           const int32_t *__restrict__ budget,
                                                           GmresDevState *__restrict__ st) {
    for (int a = threadIdx.x; a < NI; a += blockDim.x) {
        GmresDevState s;
        const double nb = sqrt(b2[a]);
        s.normb = nb == 0.0 ? 1.0 : nb;
        s.tol = tol[a];
        s.resid = 0.0;
        s.scale = 0.0;
        s.active = 1;
        s.conv = -1;
        s.steps = 0;
        s.budget = budget[a];
        s.iters = 0;
        s.converged = 0;
        st[a] = s;
    }
}

// cycle start of every instance still running: beta, the residual test
// (`<=` on the first cycle, `<` at a restart), sv = beta e_1, the scale of
// v_0; instances that ended earlier only reset their cycle marker
__global__ __launch_bounds__(1024) void bgmres_cycle_kernel(int NI, const double *__restrict__ r2, int ld, int first,
                                                            double *__restrict__ sv, GmresDevState *__restrict__ st,
                                                            double *__restrict__ scale, int32_t *__restrict__ count) {
    int act = 0;
    for (int a = threadIdx.x; a < NI; a += blockDim.x) {
        GmresDevState s = st[a];
        s.conv = -1;
        if (s.active) {
            const double beta = sqrt(r2[a]);
            s.resid = beta / s.normb;
            const bool done = first ? s.resid <= s.tol : s.resid < s.tol;
            if (done) {  // GMRES.hpp:198-213 / 284-295: j is one past the last step at a restart
                s.active = 0;
                s.conv = -3;
                s.iters = first ? 0 : s.steps + 1;
                s.converged = 1;
            } else if (s.steps >= s.budget) {
                s.active = 0;
                s.iters = s.steps;
                s.converged = 0;
            }
            double *sa = sv + (int64_t)a * ld;
            for (int k = 0; k < ld; k++) sa[k] = 0.0;
            sa[0] = beta;
            s.scale = (s.active && beta != 0.0) ? 1.0 / beta : 0.0;
        } else {
            s.scale = 0.0;
        }
        scale[a] = s.scale;
        st[a] = s;
        act += s.active;
    }
    block_count(act, 0, count);
}

// Arnoldi step i of every running instance: h1, h2 laid out (p, v, c) with
// nv = i + 1, |w|^2 per instance in n2
__global__ __launch_bounds__(1024) void bgmres_step_kernel(int NI, int K, const double *__restrict__ h1,
                                                           const double *__restrict__ h2,
                                                           const double *__restrict__ n2, int i, int ld,
                                                           double *__restrict__ H, double *__restrict__ cs,
                                                           double *__restrict__ sn, double *__restrict__ sv,
                                                           GmresDevState *__restrict__ st,
                                                           double *__restrict__ scale, int32_t *__restrict__ count) {
    const int nv = i + 1;
    int act = 0, bad = 0;
    for (int a = threadIdx.x; a < NI; a += blockDim.x) {
        GmresDevState s = st[a];
        if (!s.active) {
            scale[a] = 0.0;
            continue;
        }
        const int p = a / K, c = a % K;
        double *Ha = H + (int64_t)a * ld * (ld - 1);
        double *col = Ha + (int64_t)i * ld;
        for (int k = 0; k < nv; k++) {
            const int64_t e = ((int64_t)p * nv + k) * K + c;
            col[k] = h1[e] + h2[e];
        }
        s.scale = givens_step(Ha, cs + (int64_t)a * ld, sn + (int64_t)a * ld, sv + (int64_t)a * ld, ld, i,
                              sqrt(n2[a]), s);
        if (!s.active) {
            s.iters = s.steps;
            s.converged = s.conv >= 0 && s.resid < s.tol;
            bad += s.conv == -2;
        }
        scale[a] = s.scale;
        st[a] = s;
        act += s.active;
    }
    block_count(act, bad, count);
}

// end of a cycle whose last step had index kmax: y of every instance that
// ran in it (to kmax, or to the step it ended at) into the update
// coefficients laid out (p, v, c) with nv = kmax + 1; zero elsewhere
__global__ void bgmres_backsolve_kernel(int NI, int K, int kmax, int ld, const double *__restrict__ H,
                                        const double *__restrict__ sv, const GmresDevState *__restrict__ st,
                                        double *__restrict__ y, double *__restrict__ coef) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= NI) return;
    const GmresDevState s = st[a];
    const int k = s.active ? kmax : s.conv;
    double *ya = y + (int64_t)a * ld;
    if (k >= 0) back_solve(H + (int64_t)a * ld * (ld - 1), sv + (int64_t)a * ld, ld, k, ya);
    const int p = a / K, c = a % K, nv = kmax + 1;
    for (int v = 0; v < nv; v++) coef[((int64_t)p * nv + v) * K + c] = v <= k ? ya[v] : 0.0;
}

hipError_t launch_bgmres_init(int NI, const double *b2, const double *tol, const int32_t *budget, GmresDevState *st,
                              hipStream_t s) {
    bgmres_init_kernel<<<1, 1024, 0, s>>>(NI, b2, tol, budget, st);
    return hipGetLastError();
}
hipError_t launch_bgmres_cycle(int NI, const double *r2, int ld, int first, double *sv, GmresDevState *st,
                               double *scale, int32_t *count, hipStream_t s) {
    bgmres_cycle_kernel<<<1, 1024, 0, s>>>(NI, r2, ld, first, sv, st, scale, count);
    return hipGetLastError();
}
hipError_t launch_bgmres_step(int NI, int K, const double *h1, const double *h2, const double *n2, int i, int ld,
                              double *H, double *cs, double *sn, double *sv, GmresDevState *st, double *scale,
                              int32_t *count, hipStream_t s) {
    if (i + 1 >= ld) return hipErrorInvalidValue;
    bgmres_step_kernel<<<1, 1024, 0, s>>>(NI, K, h1, h2, n2, i, ld, H, cs, sn, sv, st, scale, count);
    return hipGetLastError();
}
hipError_t launch_bgmres_backsolve(int NI, int K, int kmax, int ld, const double *H, const double *sv,
                                   const GmresDevState *st, double *y, double *coef, hipStream_t s) {
    if (kmax < 0 || kmax >= ld - 1) return hipErrorInvalidValue;
    bgmres_backsolve_kernel<<<dim3((unsigned)((NI + 255) / 256)), 256, 0, s>>>(NI, K, kmax, ld, H, sv, st, y, coef);
    return hipGetLastError();
}

}  // namespace hodlr
