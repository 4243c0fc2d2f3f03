// capi.cpp -- extern "C" entry points of libhodlrnd (include/hodlrnd.h).
#include <cstring>
#include <exception>
#include <new>

#include <algorithm>
#include <cmath>
#include <iterator>
#include <map>
#include <mutex>

#include "op.h"
#include "solver.h"

using namespace hodlr;

namespace {

template <class F>
int guard(hodlr_ctx *ctx, F &&f) {
    try {
        f();
        return HODLR_OK;
    } catch (const HipError &e) {
        if (ctx) ctx->last_error = std::string(hipGetErrorString(e.code)) + " in " + e.where;
        return HODLR_E_DEVICE;
    } catch (const ArgError &e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->last_error = "host allocation failed";
        return HODLR_E_NOMEM;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        return HODLR_E_STATE;
    }
}

int padded_rhs(int nrhs) {
    static const int ks[] = {1, 2, 4, 8, 16, 32};
    for (int k : ks)
        if (k >= nrhs) return k;
    return -1;
}

// host vectors: problems concatenated, each in its point order, column-major
// N x k; device: tree order, K interleaved columns (zero-padded past k)
void to_tree(const hodlr_op *op, int k, int K, const double *src, std::vector<double> &dst) {
    const int64_t N = op->N;
    dst.assign((size_t)N * K, 0.0);
    for (size_t p = 0; p < op->trees.size(); p++) {
        const int64_t pb = op->pos_base[p];
        const auto &perm = op->trees[p]->perm;
        for (int64_t q = 0; q < (int64_t)perm.size(); q++)
            for (int c = 0; c < k; c++) dst[(pb + q) * K + c] = src[(int64_t)c * N + pb + perm[q]];
    }
}

void from_tree(const hodlr_op *op, int k, int K, const double *src, double *dst) {
    const int64_t N = op->N;
    for (size_t p = 0; p < op->trees.size(); p++) {
        const int64_t pb = op->pos_base[p];
        const auto &perm = op->trees[p]->perm;
        for (int64_t q = 0; q < (int64_t)perm.size(); q++)
            for (int c = 0; c < k; c++) dst[(int64_t)c * N + pb + perm[q]] = src[(pb + q) * K + c];
    }
}

}  // namespace

namespace hodlr {
thread_local DevSlab *tl_slab = nullptr;
thread_local bool tl_slab_synced = false;

struct DevSlab::Impl {
    std::mutex mu;
    std::map<size_t, size_t> free_run;  // offset -> length, disjoint, never adjacent
    std::map<size_t, size_t> used;      // offset -> length
    size_t in_use = 0;
};

DevSlab::DevSlab(size_t bytes) : impl(new Impl) {
    bytes &= ~(size_t)255;
    if (bytes == 0 || hipMalloc(reinterpret_cast<void **>(&base), bytes) != hipSuccess) {
        (void)hipGetLastError();
        base = nullptr;
        delete impl;
        throw ArgError{HODLR_E_NOMEM, "device slab of " + std::to_string(bytes) + " bytes"};
    }
    size = bytes;
    impl->free_run[0] = bytes;
}

DevSlab::~DevSlab() {
    if (base) (void)hipFree(base);
    delete impl;
}

void *DevSlab::alloc(size_t nb) {
    nb = (nb + 255) & ~(size_t)255;
    std::lock_guard<std::mutex> lk(impl->mu);
    for (auto it = impl->free_run.begin(); it != impl->free_run.end(); ++it) {
        if (it->second < nb) continue;
        const size_t off = it->first, len = it->second;
        impl->free_run.erase(it);
        if (len > nb) impl->free_run[off + nb] = len - nb;
        impl->used[off] = nb;
        impl->in_use += nb;
        return base + off;
    }
    return nullptr;
}

void DevSlab::free(void *p) {
    const size_t off = (size_t)(static_cast<char *>(p) - base);
    std::lock_guard<std::mutex> lk(impl->mu);
    auto u = impl->used.find(off);
    if (u == impl->used.end()) return;  // not from this slab: nothing to do
    size_t start = off, len = u->second;
    impl->in_use -= len;
    impl->used.erase(u);
    auto nx = impl->free_run.lower_bound(off);
    if (nx != impl->free_run.end() && nx->first == start + len) {  // merge with the next run
        len += nx->second;
        nx = impl->free_run.erase(nx);
    }
    if (nx != impl->free_run.begin()) {  // merge with the previous run
        auto pv = std::prev(nx);
        if (pv->first + pv->second == start) {
            start = pv->first;
            len += pv->second;
            impl->free_run.erase(pv);
        }
    }
    impl->free_run[start] = len;
}

size_t DevSlab::largest_free() {
    std::lock_guard<std::mutex> lk(impl->mu);
    size_t m = 0;
    for (auto &kv : impl->free_run) m = std::max(m, kv.second);
    return m;
}

size_t DevSlab::in_use() {
    std::lock_guard<std::mutex> lk(impl->mu);
    return impl->in_use;
}

size_t device_free_bytes() {
    if (tl_slab) return tl_slab->largest_free();
    size_t fr = 0, tot = 0;
    HIP_CHECK(hipMemGetInfo(&fr, &tot));
    return fr;
}

// Stream-ordered temporaries (assembly scratch, work lists) come from the
// device's default memory pool; keeping up to 8 GiB of freed pool memory
// mapped spares every later assembly the cost of mapping and clearing it
// again.  Best effort: without the attribute the pool still works.
void keep_pool_memory(int device) {
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) != hipSuccess || !pool) {
        (void)hipGetLastError();
        return;
    }
    uint64_t keep = (uint64_t)8 << 30;
    if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) (void)hipGetLastError();
}
}  // namespace hodlr

namespace {

void set_device(hodlr_ctx *ctx) {
    if (ctx->device < 0) throw ArgError{HODLR_E_STATE, "host-only context (device -1) cannot run device work"};
    HIP_CHECK(hipSetDevice(ctx->device));
}

}  // namespace

extern "C" {

const char *hodlr_version(void) { return "hodlrnd-mi355x 0.1 (gfx950)"; }

int hodlr_ctx_set_assembly_staging(hodlr_ctx *ctx, int64_t max_bytes) {
    if (!ctx || max_bytes < -1) return HODLR_E_ARG;
    ctx->stage_bytes = max_bytes;
    return HODLR_OK;
}

int hodlr_ctx_create(int device, hodlr_ctx **out) {
    if (!out) return HODLR_E_ARG;
    *out = nullptr;
    auto *ctx = new (std::nothrow) hodlr_ctx;
    if (!ctx) return HODLR_E_NOMEM;
    ctx->device = device;
    int rc = guard(ctx, [&] {
        if (device == -1) return;  // host-only context: trees and orderings, no device work
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw ArgError{HODLR_E_ARG, "no such device"};
        HIP_CHECK(hipSetDevice(device));
        HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        keep_pool_memory(device);
        noalloc_scratch_setup(device);  // this device's copy of the allocation-free matvec's array
    });
    if (rc != HODLR_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return HODLR_OK;
}

void hodlr_ctx_destroy(hodlr_ctx *ctx) {
    if (!ctx) return;
    if (ctx->stream && ctx->device >= 0) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *hodlr_last_error(hodlr_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

int hodlr_tree_build(hodlr_ctx *ctx, int ndim, int64_t n, const double *pts, const double *lo,
                     const double *hi, int nmax, int dprime, hodlr_tree **out) {
    if (!ctx || !out) return HODLR_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        if (ndim < 1 || ndim > kMaxDim) throw ArgError{HODLR_E_ARG, "ndim must be in 1..8"};
        if (n < 0 || (n > 0 && !pts) || !lo || !hi) throw ArgError{HODLR_E_ARG, "null points or box"};
        if (nmax < 1) throw ArgError{HODLR_E_ARG, "nmax must be positive"};
        if (n > INT32_MAX) throw ArgError{HODLR_E_ARG, "at most 2^31-1 points"};
        for (int k = 0; k < ndim; k++)
            if (!(std::isfinite(lo[k]) && std::isfinite(hi[k]) && lo[k] < hi[k]))
                throw ArgError{HODLR_E_ARG, "bounding box needs finite lo < hi in every dimension"};
        // a point outside the box is never separated from the box's far
        // corner by the midpoint splits, so the tree would deepen forever
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < ndim; k++) {
                const double x = pts[i * ndim + k];
                if (!(x >= lo[k] && x <= hi[k]))
                    throw ArgError{HODLR_E_ARG, "point " + std::to_string(i) + " lies outside the bounding box"};
            }
        auto *t = new hodlr_tree;
        t->ctx = ctx;
        t->ndim = ndim;
        t->n = n;
        t->nmax = nmax;
        t->dprime = dprime;
        t->pts.assign(pts, pts + n * ndim);
        t->lo.assign(lo, lo + ndim);
        t->hi.assign(hi, hi + ndim);
        try {
            build_tree(*t);
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

int hodlr_tree_get_info(hodlr_tree *t, hodlr_tree_info *info) {
    if (!t || !info) return HODLR_E_ARG;
    std::memset(info, 0, sizeof(*info));
    info->nlevels = t->nlevels;
    info->ndim = t->ndim;
    info->npoints = t->n;
    info->nnodes = t->level_off.back();
    const int L = t->nlevels - 1;
    info->nleaves = t->level_size(L);
    for (int64_t j = 0; j < info->nleaves; j++) {
        const int64_t g = t->node(L, j);
        if (t->count[g] > info->max_leaf) info->max_leaf = t->count[g];
        if (!t->count[g]) continue;
        for (int64_t a = t->nbr_ptr[g]; a < t->nbr_ptr[g + 1]; a++)
            if (t->count[t->node(L, t->nbr[a])]) info->n_near++;
    }
    for (int l = 0; l <= L; l++)
        for (int64_t j = 0; j < t->level_size(l); j++) {
            const int64_t g = t->node(l, j);
            if (!t->count[g]) continue;
            for (int64_t a = t->il_ptr[g]; a < t->il_ptr[g + 1]; a++)
                if (t->count[t->node(l, t->il[a])]) info->n_admissible++;
        }
    return HODLR_OK;
}

int hodlr_tree_perm(hodlr_tree *t, int64_t *perm) {
    if (!t || !perm) return HODLR_E_ARG;
    std::memcpy(perm, t->perm.data(), sizeof(int64_t) * t->n);
    return HODLR_OK;
}

int hodlr_tree_nodes(hodlr_tree *t, int64_t *level_offset, int64_t *start, int64_t *count) {
    if (!t) return HODLR_E_ARG;
    if (level_offset) std::memcpy(level_offset, t->level_off.data(), sizeof(int64_t) * t->level_off.size());
    if (start) std::memcpy(start, t->start.data(), sizeof(int64_t) * t->start.size());
    if (count) std::memcpy(count, t->count.data(), sizeof(int64_t) * t->count.size());
    return HODLR_OK;
}

int hodlr_tree_lists(hodlr_tree *t, int which, int64_t *ptr, int64_t *idx) {
    if (!t || !ptr) return HODLR_E_ARG;
    const auto &p = which ? t->il_ptr : t->nbr_ptr;
    const auto &v = which ? t->il : t->nbr;
    std::memcpy(ptr, p.data(), sizeof(int64_t) * p.size());
    if (idx && !v.empty()) std::memcpy(idx, v.data(), sizeof(int64_t) * v.size());
    return HODLR_OK;
}

void hodlr_tree_destroy(hodlr_tree *t) { delete t; }

int hodlr_op_assemble(hodlr_ctx *ctx, hodlr_tree *t, const hodlr_kernel_desc *k, double eps_aca, int prec_flags,
                      hodlr_op **out) {
    if (!ctx || !t || !k || !out) return HODLR_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        if (k->ndim != t->ndim) throw ArgError{HODLR_E_ARG, "kernel ndim differs from the tree's"};
        if (k->family < 0 || k->family > 2) throw ArgError{HODLR_E_ARG, "unknown kernel family"};
        if (!(eps_aca > 0.0 && eps_aca < 1.0)) throw ArgError{HODLR_E_ARG, "eps_aca must be in (0, 1)"};
        set_device(ctx);
        auto *op = new hodlr_op;
        op->ctx = ctx;
        op->trees = {t};
        op->kds = {*k};
        op->eps = eps_aca;
        op->prec_flags = prec_flags | HODLR_PREC_F64;
        try {
            assemble(*op);
        } catch (...) {
            delete op;
            throw;
        }
        for (auto &kd : op->kds) kd.weight = nullptr;  // caller arrays are not referenced after assembly
        op->kd.weight = nullptr;
        *out = op;
    });
}

int hodlr_op_assemble_custom(hodlr_ctx *ctx, hodlr_tree *t, const hodlr_kernel_desc *k, const void *kernel_obj,
                             int64_t kernel_obj_bytes, const void *entries, double eps_aca, int prec_flags,
                             hodlr_op **out) {
    if (!ctx || !t || !k || !entries || !out || kernel_obj_bytes < 0 || (kernel_obj_bytes > 0 && !kernel_obj))
        return HODLR_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        if (k->ndim != t->ndim) throw ArgError{HODLR_E_ARG, "kernel ndim differs from the tree's"};
        if (k->family != HODLR_KERNEL_CUSTOM)
            throw ArgError{HODLR_E_ARG, "hodlr_op_assemble_custom takes the HODLR_KERNEL_CUSTOM family"};
        const auto *L = static_cast<const EntryLaunchers *>(entries);
        if (L->version != kEntryAbiVersion || L->layout != entry_abi_layout())
            throw ArgError{HODLR_E_ARG, "entry launcher table built against another layout of the entry kernels"};
        if (!L->aca || !L->dense_fill || !L->near_fill || !L->dense_rows)
            throw ArgError{HODLR_E_ARG, "entry launcher table has a null launcher"};
        if (kernel_obj_bytes > (int64_t)1 << 20) throw ArgError{HODLR_E_ARG, "kernel object larger than 1 MiB"};
        if (!(eps_aca > 0.0 && eps_aca < 1.0)) throw ArgError{HODLR_E_ARG, "eps_aca must be in (0, 1)"};
        set_device(ctx);
        auto *op = new hodlr_op;
        op->ctx = ctx;
        op->trees = {t};
        op->kds = {*k};
        op->eps = eps_aca;
        op->prec_flags = prec_flags | HODLR_PREC_F64;
        op->entries = L;
        try {
            if (kernel_obj_bytes > 0) {
                op->kobj.alloc((size_t)kernel_obj_bytes);
                HIP_CHECK(hipMemcpy(op->kobj.p, kernel_obj, (size_t)kernel_obj_bytes, hipMemcpyHostToDevice));
            }
            assemble(*op);
        } catch (...) {
            delete op;
            throw;
        }
        for (auto &kd : op->kds) kd.weight = nullptr;
        op->kd.weight = nullptr;
        *out = op;
    });
}

int hodlr_op_assemble_batch(hodlr_ctx *ctx, int nprob, hodlr_tree *const *trees, const hodlr_kernel_desc *k,
                            double eps_aca, int prec_flags, hodlr_op **out) {
    return hodlr_op_assemble_skeleton(ctx, nprob, trees, k, eps_aca, -1, prec_flags, out);
}

int hodlr_op_assemble_skeleton(hodlr_ctx *ctx, int nprob, hodlr_tree *const *trees, const hodlr_kernel_desc *k,
                               double eps_aca, int sys_size, int prec_flags, hodlr_op **out) {
    if (!ctx || nprob < 1 || !trees || !k || !out) return HODLR_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        int64_t total = 0;
        for (int p = 0; p < nprob; p++) {
            if (!trees[p]) throw ArgError{HODLR_E_ARG, "null tree in batch"};
            if (k[p].ndim != trees[p]->ndim) throw ArgError{HODLR_E_ARG, "kernel ndim differs from the tree's"};
            if (k[p].family < 0 || k[p].family > 2) throw ArgError{HODLR_E_ARG, "unknown kernel family"};
            total += trees[p]->n;
        }
        if (total > INT32_MAX) throw ArgError{HODLR_E_ARG, "a batch holds at most 2^31-1 points"};
        if (!(eps_aca > 0.0 && eps_aca < 1.0)) throw ArgError{HODLR_E_ARG, "eps_aca must be in (0, 1)"};
        set_device(ctx);
        auto *op = new hodlr_op;
        op->ctx = ctx;
        op->trees.assign(trees, trees + nprob);
        op->kds.assign(k, k + nprob);
        op->eps = eps_aca;
        op->sys_size = sys_size < 0 ? -1 : sys_size;
        op->prec_flags = prec_flags | HODLR_PREC_F64;
        try {
            assemble(*op);
        } catch (...) {
            delete op;
            throw;
        }
        for (auto &kd : op->kds) kd.weight = nullptr;
        op->kd.weight = nullptr;
        *out = op;
    });
}

int hodlr_op_problem_offsets(hodlr_op *op, int64_t *offsets, int cap) {
    if (!op) return HODLR_E_ARG;
    const int P = (int)op->trees.size();
    if (!offsets) return P;
    if (cap < P + 1) return HODLR_E_ARG;
    for (int p = 0; p <= P; p++) offsets[p] = op->pos_base[p];
    return P;
}

int hodlr_op_get_stats(hodlr_op *op, hodlr_op_stats *s) {
    if (!op || !s) return HODLR_E_ARG;
    *s = op->stats;
    return HODLR_OK;
}

int hodlr_op_block_ranks(hodlr_op *op, int32_t *rank, int32_t *dense, int64_t cap) {
    if (!op) return HODLR_E_ARG;
    if (cap < op->nblk) return (int)HODLR_E_ARG;
    for (int64_t b = 0; b < op->nblk; b++) {
        const bool dn = op->flags[b] & 1;
        if (rank) rank[b] = dn ? std::min(op->b_m[b], op->b_n[b]) : op->rank[b];
        if (dense) dense[b] = dn;
    }
    return HODLR_OK;
}

void hodlr_op_destroy(hodlr_op *op) {
    if (!op) return;
    if (op->ctx->device >= 0) (void)hipSetDevice(op->ctx->device);
    delete op;
}

int hodlr_to_tree_order(hodlr_tree *t, int nrhs, const double *src, double *dst) {  // one problem
    if (!t || !src || !dst || nrhs < 1) return HODLR_E_ARG;
    for (int64_t p = 0; p < t->n; p++)
        for (int c = 0; c < nrhs; c++) dst[p * nrhs + c] = src[(int64_t)c * t->n + t->perm[p]];
    return HODLR_OK;
}

int hodlr_from_tree_order(hodlr_tree *t, int nrhs, const double *src, double *dst) {
    if (!t || !src || !dst || nrhs < 1) return HODLR_E_ARG;
    for (int64_t p = 0; p < t->n; p++)
        for (int c = 0; c < nrhs; c++) dst[(int64_t)c * t->n + t->perm[p]] = src[p * nrhs + c];
    return HODLR_OK;
}

int hodlr_matvec(hodlr_op *op, int prec, int nrhs, const double *x, double *y) {
    if (!op || !x || !y || nrhs < 1) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        const int K = padded_rhs(nrhs);
        if (K < 0 || (prec == HODLR_PREC_F64 && K > 16 && op->n_skel > 0))
            throw ArgError{HODLR_E_ARG, "too many right-hand sides"};
        if (prec != HODLR_PREC_F64 && prec != HODLR_PREC_F32) throw ArgError{HODLR_E_ARG, "bad precision"};
        set_device(op->ctx);
        const int64_t n = op->N;
        hipStream_t s = op->ctx->stream;
        std::vector<double> xt, yt((size_t)n * K);
        to_tree(op, nrhs, K, x, xt);
        DevBuf dx, dy;
        if (prec == HODLR_PREC_F64) {
            upload(dx, xt);
            dy.alloc(sizeof(double) * n * K);
            op->matvec_dev(prec, K, dx.p, dy.p, s);
            HIP_CHECK(hipStreamSynchronize(s));
            HIP_CHECK(hipMemcpy(yt.data(), dy.p, sizeof(double) * n * K, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> xf(xt.begin(), xt.end()), yf((size_t)n * K);
            upload(dx, xf);
            dy.alloc(sizeof(float) * n * K);
            op->matvec_dev(prec, K, dx.p, dy.p, s);
            HIP_CHECK(hipStreamSynchronize(s));
            HIP_CHECK(hipMemcpy(yf.data(), dy.p, sizeof(float) * n * K, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < yf.size(); i++) yt[i] = yf[i];
        }
        from_tree(op, nrhs, K, yt.data(), y);
    });
}

int hodlr_matvec_device(hodlr_op *op, int prec, int nrhs, const void *x, void *y, void *stream) {
    if (!op || !x || !y) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        if (padded_rhs(nrhs) != nrhs) throw ArgError{HODLR_E_ARG, "nrhs must be 1, 2, 4, 8, 16 or 32"};
        if (prec == HODLR_PREC_F64 && nrhs > 16 && op->n_skel > 0)
            throw ArgError{HODLR_E_ARG, "FP64 skeleton operators support at most 16 RHS"};
        if (prec != HODLR_PREC_F64 && prec != HODLR_PREC_F32) throw ArgError{HODLR_E_ARG, "bad precision"};
        set_device(op->ctx);
        op->matvec_dev(prec, nrhs, x, y, stream ? static_cast<hipStream_t>(stream) : op->ctx->stream);
    });
}

int hodlr_dense_rows(hodlr_op *op, int64_t nrows, const int64_t *rows, const double *x, double *y) {
    if (!op || nrows < 0 || (nrows && (!rows || !x || !y))) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        set_device(op->ctx);
        hipStream_t s = op->ctx->stream;
        const int64_t n = op->N;
        std::vector<int64_t> tr(nrows);
        std::vector<int32_t> pr(nrows);
        std::vector<std::vector<int64_t>> inv(op->trees.size());  // point id -> tree position
        for (int64_t r = 0; r < nrows; r++) {  // point ids -> global tree positions and problems
            const int64_t g = rows[r];
            if (g < 0 || g >= n) throw ArgError{HODLR_E_ARG, "row index out of range"};
            const int p = (int)(std::upper_bound(op->pos_base.begin(), op->pos_base.end(), g) -
                                op->pos_base.begin()) - 1;
            if (inv[p].empty()) {
                const auto &perm = op->trees[p]->perm;
                inv[p].resize(perm.size());
                for (size_t q = 0; q < perm.size(); q++) inv[p][perm[q]] = (int64_t)q;
            }
            tr[r] = op->pos_base[p] + inv[p][g - op->pos_base[p]];
            pr[r] = p;
        }
        std::vector<double> xt;
        to_tree(op, 1, 1, x, xt);
        DevBuf drows, dprob, dx, dy, dpb;
        upload(drows, tr, s);
        upload(dprob, pr, s);
        upload(dx, xt, s);
        upload(dpb, op->pos_base, s);
        dy.alloc(sizeof(double) * std::max<int64_t>(nrows, 1));
        HIP_CHECK(op->entries->dense_rows(drows.as<int64_t>(), dprob.as<int32_t>(), nrows, op->kdev.as<KernelDev>(),
                                    dpb.as<int64_t>(), dx.as<double>(), dy.as<double>(), s));
        HIP_CHECK(hipMemcpyAsync(y, dy.p, sizeof(double) * nrows, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    });
}

int hodlr_gmres_device(hodlr_op *op, const double *b, double *x, const hodlr_gmres_params *params,
                       hodlr_gmres_info *info) {
    if (!op || !b || !x || !params || !info) return HODLR_E_ARG;
    int rc = guard(op->ctx, [&] {
        set_device(op->ctx);
        solve(*op, b, x, *params, *info);
    });
    if (rc == HODLR_OK && !info->converged) return HODLR_E_NOCONV;
    return rc;
}

int hodlr_gmres(hodlr_op *op, const double *b, double *x, const hodlr_gmres_params *params,
                hodlr_gmres_info *info) {
    if (!op || !b || !x || !params || !info) return HODLR_E_ARG;
    int rc = guard(op->ctx, [&] {
        set_device(op->ctx);
        const int64_t n = op->N;
        std::vector<double> bt, xt;
        to_tree(op, 1, 1, b, bt);
        to_tree(op, 1, 1, x, xt);
        DevBuf db, dx;
        upload(db, bt);
        upload(dx, xt);
        solve(*op, db.as<double>(), dx.as<double>(), *params, *info);
        HIP_CHECK(hipMemcpy(xt.data(), dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        from_tree(op, 1, 1, xt.data(), x);
    });
    if (rc == HODLR_OK && !info->converged) return HODLR_E_NOCONV;
    return rc;
}

int hodlr_gmres_batch_device(hodlr_op *op, int nrhs, const double *b, double *x, const hodlr_gmres_params *params,
                             hodlr_gmres_info *infos) {
    if (!op || !b || !x || !params || !infos || params->ortho != HODLR_ORTHO_CGS2) return HODLR_E_ARG;
    int rc = guard(op->ctx, [&] {
        if (nrhs != 1 && nrhs != 2 && nrhs != 4 && nrhs != 8 && nrhs != 16)
            throw ArgError{HODLR_E_ARG, "nrhs must be 1, 2, 4, 8 or 16"};
        set_device(op->ctx);
        solve_batch(*op, nrhs, b, x, *params, infos);
    });
    if (rc != HODLR_OK) return rc;
    for (size_t a = 0; a < op->trees.size() * (size_t)nrhs; a++)
        if (!infos[a].converged) return HODLR_E_NOCONV;
    return HODLR_OK;
}

int hodlr_gmres_batch(hodlr_op *op, int nrhs, const double *b, double *x, const hodlr_gmres_params *params,
                      hodlr_gmres_info *infos) {
    if (!op || !b || !x || !params || !infos || nrhs < 1 || params->ortho != HODLR_ORTHO_CGS2) return HODLR_E_ARG;
    int rc = guard(op->ctx, [&] {
        const int K = padded_rhs(nrhs);
        if (K < 0 || K > 16) throw ArgError{HODLR_E_ARG, "at most 16 right-hand sides per problem"};
        set_device(op->ctx);
        const int64_t n = op->N;
        std::vector<double> bt, xt;
        to_tree(op, nrhs, K, b, bt);
        to_tree(op, nrhs, K, x, xt);
        DevBuf db, dx;
        upload(db, bt);
        upload(dx, xt);
        const int P = (int)op->trees.size();
        std::vector<hodlr_gmres_info> all((size_t)P * K);
        solve_batch(*op, K, db.as<double>(), dx.as<double>(), *params, all.data());
        HIP_CHECK(hipMemcpy(xt.data(), dx.p, sizeof(double) * n * K, hipMemcpyDeviceToHost));
        from_tree(op, nrhs, K, xt.data(), x);
        for (int p = 0; p < P; p++)
            for (int c = 0; c < nrhs; c++) infos[(size_t)p * nrhs + c] = all[(size_t)p * K + c];
    });
    if (rc != HODLR_OK) return rc;
    for (size_t a = 0; a < op->trees.size() * (size_t)nrhs; a++)
        if (!infos[a].converged) return HODLR_E_NOCONV;
    return HODLR_OK;
}

int hodlr_matvec_kernel_bytes(hodlr_op *op, int prec, int nrhs, double *bytes) {
    if (!op || !bytes) return HODLR_E_ARG;
    const double s = prec == HODLR_PREC_F32 ? 4.0 : 8.0, k = nrhs, n = (double)op->N;
    bytes[0] = s * (op->e_near + op->e_far) + 2.0 * k * s * n;
    bytes[1] = s * op->e_v + k * s * (n + (double)op->n_t);
    bytes[2] = s * (double)(op->n_np + op->n_up) + k * s * (n + (double)op->n_t) + k * s * n;
    return HODLR_OK;
}

int hodlr_op_save(hodlr_op *op, const char *path) {
    if (!op || !path) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        for (const auto &kd : op->kds)
            if (kd.family == HODLR_KERNEL_CUSTOM)
                throw ArgError{HODLR_E_STATE, "an operator with a caller-compiled kernel cannot be saved"};
        set_device(op->ctx);
        save_op(*op, path);
    });
}

int hodlr_op_load(hodlr_ctx *ctx, const char *path, int prec_flags, hodlr_op **out) {
    if (!ctx || !path || !out) return HODLR_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] { *out = load_op(ctx, path, prec_flags); });
}

int hodlr_op_get_tree(hodlr_op *op, int problem, hodlr_tree **tree) {
    if (!op || !tree || problem < 0 || problem >= (int)op->trees.size()) return HODLR_E_ARG;
    *tree = op->trees[problem];
    return HODLR_OK;
}

int hodlr_op_tune_matvec(hodlr_op *op, int prec_mask, int reps) {
    if (!op || reps < 1) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        set_device(op->ctx);
        op->tune_matvec(prec_mask, reps);
    });
}

namespace {
int variant_slot(int prec, int nrhs) {
    if (prec != HODLR_PREC_F64 && prec != HODLR_PREC_F32) return -1;
    if (padded_rhs(nrhs) != nrhs) return -1;
    int lg = 0;
    while ((1 << lg) < nrhs) lg++;
    return lg;
}
}  // namespace

int hodlr_op_set_solver_graphs(hodlr_op *op, int enable) {
    if (!op) return HODLR_E_ARG;
    op->solver_graphs = enable != 0;
    return HODLR_OK;
}

int hodlr_op_set_matvec_variant(hodlr_op *op, int prec, int nrhs, int variant) {
    if (!op || variant < 0 || variant > 127) return HODLR_E_ARG;
    const int lg = variant_slot(prec, nrhs);
    if (lg < 0) return HODLR_E_ARG;
    op->mv_variant[prec == HODLR_PREC_F64 ? 0 : 1][lg] = variant;
    op->work_gen++;
    return HODLR_OK;
}

int hodlr_op_get_matvec_variant(hodlr_op *op, int prec, int nrhs, int *variant) {
    if (!op || !variant) return HODLR_E_ARG;
    const int lg = variant_slot(prec, nrhs);
    if (lg < 0) return HODLR_E_ARG;
    *variant = op->mv_variant[prec == HODLR_PREC_F64 ? 0 : 1][lg];
    return HODLR_OK;
}

int hodlr_op_set_skeleton_schedule(hodlr_op *op, int mode) {
    if (!op || mode < HODLR_SKEL_AUTO || mode > HODLR_SKEL_CAPPED) return HODLR_E_ARG;
    op->skel_mode = mode;
    return HODLR_OK;
}

int hodlr_op_get_skeleton_schedule(hodlr_op *op, int prec, int nrhs, int *mode) {
    if (!op || !mode) return HODLR_E_ARG;
    if (variant_slot(prec, nrhs) < 0) return HODLR_E_ARG;
    *mode = op->n_skcore ? op->skel_schedule(prec, nrhs) : HODLR_SKEL_SERIAL;
    return HODLR_OK;
}

int hodlr_matvec_timing(hodlr_op *op, int prec, int nrhs, int reps, double *ms) {
    if (!op || !ms || reps < 1) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        if (padded_rhs(nrhs) != nrhs) throw ArgError{HODLR_E_ARG, "nrhs must be 1, 2, 4, 8, 16 or 32"};
        set_device(op->ctx);
        hipStream_t s = op->ctx->stream;
        const int64_t n = op->N;
        const size_t es = prec == HODLR_PREC_F32 ? sizeof(float) : sizeof(double);
        DevBuf dx, dy;
        dx.alloc(es * n * nrhs);
        dy.alloc(es * n * nrhs);
        HIP_CHECK(hipMemset(dx.p, 0, es * n * nrhs));
        const size_t E = kMatvecEvents;
        std::vector<hipEvent_t> ev(E * reps);
        for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
        op->matvec_dev(prec, nrhs, dx.p, dy.p, s);  // warm-up
        for (int r = 0; r < reps; r++) op->matvec_dev(prec, nrhs, dx.p, dy.p, s, &ev[E * r]);
        HIP_CHECK(hipStreamSynchronize(s));
        double a = 0, b = 0, c = 0, t = 0;
        for (int r = 0; r < reps; r++) {
            float t01 = 0, t12 = 0, t03 = 0, t45 = 0;
            HIP_CHECK(hipEventElapsedTime(&t01, ev[E * r], ev[E * r + 1]));
            HIP_CHECK(hipEventElapsedTime(&t12, ev[E * r + 1], ev[E * r + 2]));
            HIP_CHECK(hipEventElapsedTime(&t03, ev[E * r], ev[E * r + 3]));
            if (op->n_skcore) HIP_CHECK(hipEventElapsedTime(&t45, ev[E * r + 4], ev[E * r + 5]));
            a += t01;
            b += t12;
            c += t45;
            t += t03;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        ms[0] = t / reps;
        ms[1] = a / reps;
        ms[2] = b / reps;
        ms[3] = c / reps;
    });
}

int hodlr_op_profile(hodlr_op *op, int enable) {
    if (!op) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        op->profile_collect();
        op->profiling = enable != 0;
    });
}

int hodlr_op_profile_read(hodlr_op *op, double *ms, int64_t *counts, int reset) {
    if (!op) return HODLR_E_ARG;
    return guard(op->ctx, [&] {
        op->profile_collect();
        if (ms) {
            ms[0] = op->prof_ms[0][0];
            ms[1] = op->prof_ms[0][1];
            ms[2] = op->prof_ms[1][0];
            ms[3] = op->prof_ms[1][1];
            ms[4] = op->prof_ms[0][2];
            ms[5] = op->prof_ms[1][2];
        }
        if (counts) {
            counts[0] = op->prof_count[0];
            counts[1] = op->prof_count[1];
        }
        if (reset) {
            std::memset(op->prof_ms, 0, sizeof(op->prof_ms));
            op->prof_count[0] = op->prof_count[1] = 0;
        }
    });
}

int hodlr_stream_copy(hodlr_ctx *ctx, int64_t bytes, int reps, double *gbps) {
    if (!ctx || bytes < 16 || reps < 1 || !gbps) return HODLR_E_ARG;
    return guard(ctx, [&] {
        set_device(ctx);
        hipStream_t s = ctx->stream;
        bytes = bytes / 16 * 16;
        DevBuf a, b;
        a.alloc((size_t)bytes);
        b.alloc((size_t)bytes);
        HIP_CHECK(hipMemsetAsync(a.p, 1, (size_t)bytes, s));
        HIP_CHECK(launch_stream_copy(a.p, b.p, bytes, s));  // warm-up
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) HIP_CHECK(launch_stream_copy(a.p, b.p, bytes, s));
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) * 1e-9;
    });
}

int hodlr_stream_read(hodlr_ctx *ctx, int64_t bytes, int reps, double *gbps) {
    if (!ctx || bytes < 16 || reps < 1 || !gbps) return HODLR_E_ARG;
    return guard(ctx, [&] {
        set_device(ctx);
        hipStream_t s = ctx->stream;
        bytes = bytes / 16 * 16;
        DevBuf a, o;
        a.alloc((size_t)bytes);
        o.alloc(sizeof(float));
        HIP_CHECK(hipMemsetAsync(a.p, 0, (size_t)bytes, s));  // zeros: the kernel stores nothing
        HIP_CHECK(launch_stream_read(a.p, bytes, o.as<float>(), s));  // warm-up
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) HIP_CHECK(launch_stream_read(a.p, bytes, o.as<float>(), s));
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = (double)bytes * reps / (ms * 1e-3) * 1e-9;
    });
}

int hodlr_device_alloc(hodlr_ctx *ctx, int64_t bytes, void **ptr) {
    if (!ctx || !ptr || bytes < 0) return HODLR_E_ARG;
    return guard(ctx, [&] {
        set_device(ctx);
        if (hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 8) != hipSuccess)
            throw ArgError{HODLR_E_NOMEM, "hipMalloc failed"};
    });
}

int hodlr_device_free(hodlr_ctx *ctx, void *ptr) {
    if (!ctx) return HODLR_E_ARG;
    return guard(ctx, [&] { HIP_CHECK(hipFree(ptr)); });
}

int hodlr_copy_to_device(hodlr_ctx *ctx, void *dst, const void *src, int64_t bytes) {
    if (!ctx || !dst || !src || bytes < 0) return HODLR_E_ARG;
    return guard(ctx, [&] {  // after the work queued on the context's stream
        set_device(ctx);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        HIP_CHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    });
}

int hodlr_copy_to_host(hodlr_ctx *ctx, void *dst, const void *src, int64_t bytes) {
    if (!ctx || !dst || !src || bytes < 0) return HODLR_E_ARG;
    return guard(ctx, [&] {
        set_device(ctx);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        HIP_CHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
