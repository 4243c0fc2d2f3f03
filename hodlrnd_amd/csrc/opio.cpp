// opio.cpp -- operator files (the cereal vector files are in vecio.cpp).
//
// The reference keeps operators in memory only.  An operator
// file lets a compressed operator outlive the process (a restarted server
// reloads its resident operators without repeating ACA).  Layout, all
// little-endian:
//   "HODLROP1"  u32 version (2)  u32 0x01020304
//   i32 P  f64 eps_aca  i32 sys_size (-1: every factor stored; 0 in version 1)
//   per problem: i32 ndim  i32 nmax  i32 dprime  i32 family  i64 n
//                f64 scale  f64 diag  i32 has_weights  i32 0
//                f64 lo[ndim]  f64 hi[ndim]  f64 points[n * ndim] (point order)
//                f64 weights[n] (point order, if has_weights)
//   i64 nblocks  i32 rank[nblocks]  i32 flags[nblocks]
//   i64 n_np  i64 n_up  i64 n_vp  f64 NP[n_np]  f64 UP[n_up]  f64 VP[n_vp]
//   if sys_size >= 0 (skeleton blocks):
//     i64 n_sk  i64 n_g  i32 SK[2 n_sk] (pivot tree positions)  f64 G[n_g] (cores)
//   u64 FNV-1a hash (8-byte words, per write) of everything above
// The trees and block lists are rebuilt from the points (the tree and list
// construction is deterministic), the stored ranks replace the ACA rank
// pass and the stored FP64 arenas replace the fill pass; the FP32 copy is
// converted on load as after assembly.
#include <cstdio>
#include <cstring>
#include <fstream>

#include "op.h"

namespace hodlr {
namespace {

constexpr char kMagic[8] = {'H', 'O', 'D', 'L', 'R', 'O', 'P', '1'};
constexpr uint32_t kVersion = 2, kEndian = 0x01020304u;
constexpr size_t kChunk = (size_t)1 << 26;  // bytes per device <-> file transfer

// FNV-1a over 8-byte little-endian words (a zero-padded tail word), chained
// across calls
uint64_t fnv_words(uint64_t h, const void *p, size_t nb) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    size_t i = 0;
    for (; i + 8 <= nb; i += 8) {
        uint64_t w;
        std::memcpy(&w, c + i, 8);
        h = (h ^ w) * 1099511628211ull;
    }
    if (i < nb) {
        uint64_t w = 0;
        std::memcpy(&w, c + i, nb - i);
        h = (h ^ w) * 1099511628211ull;
    }
    return h;
}

struct Writer {
    std::ofstream f;
    uint64_t h = 1469598103934665603ull;
    explicit Writer(const char *path) : f(path, std::ios::binary | std::ios::trunc) {
        if (!f) throw ArgError{HODLR_E_ARG, std::string("cannot open ") + path + " for writing"};
    }
    void raw(const void *p, size_t nb) {
        h = fnv_words(h, p, nb);
        f.write(static_cast<const char *>(p), (std::streamsize)nb);
        if (!f) throw ArgError{HODLR_E_ARG, "write failed"};
    }
    template <class T> void put(T v) { raw(&v, sizeof(T)); }
    void device(const void *dp, size_t nb, std::vector<char> &buf) {
        for (size_t o = 0; o < nb; o += kChunk) {
            const size_t c = std::min(kChunk, nb - o);
            buf.resize(c);
            HIP_CHECK(hipMemcpy(buf.data(), static_cast<const char *>(dp) + o, c, hipMemcpyDeviceToHost));
            raw(buf.data(), c);
        }
    }
};

struct Reader {
    std::ifstream f;
    uint64_t h = 1469598103934665603ull;
    uint64_t size = 0;
    explicit Reader(const char *path) : f(path, std::ios::binary | std::ios::ate) {
        if (!f) throw ArgError{HODLR_E_ARG, std::string("cannot open ") + path};
        size = (uint64_t)f.tellg();
        f.seekg(0);
    }
    // bytes not read yet: bounds every count taken from the file before it
    // sizes a host allocation
    uint64_t left() { return size - (uint64_t)f.tellg(); }
    void raw(void *p, size_t nb) {
        f.read(static_cast<char *>(p), (std::streamsize)nb);
        if ((size_t)f.gcount() != nb) throw ArgError{HODLR_E_ARG, "operator file is truncated"};
        h = fnv_words(h, p, nb);
    }
    template <class T> T get() {
        T v;
        raw(&v, sizeof(T));
        return v;
    }
    void device(void *dp, size_t nb, std::vector<char> &buf, hipStream_t s) {
        for (size_t o = 0; o < nb; o += kChunk) {
            const size_t c = std::min(kChunk, nb - o);
            buf.resize(c);
            raw(buf.data(), c);
            HIP_CHECK(hipMemcpyAsync(static_cast<char *>(dp) + o, buf.data(), c, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
    }
};

}  // namespace

void save_op(hodlr_op &op, const char *path) {
    Writer w(path);
    w.raw(kMagic, 8);
    w.put<uint32_t>(kVersion);
    w.put<uint32_t>(kEndian);
    const int P = (int)op.trees.size();
    w.put<int32_t>(P);
    w.put<double>(op.eps);
    w.put<int32_t>(op.sys_size);
    std::vector<double> wt;
    for (int p = 0; p < P; p++) {
        const hodlr_tree &t = *op.trees[p];
        const hodlr_kernel_desc &k = op.kds[p];
        w.put<int32_t>(t.ndim);
        w.put<int32_t>(t.nmax);
        w.put<int32_t>(t.dprime);
        w.put<int32_t>(k.family);
        w.put<int64_t>(t.n);
        w.put<double>(k.scale);
        w.put<double>(k.diag);
        const bool hw = op.w_off.size() > (size_t)p && op.w_off[p] >= 0;
        w.put<int32_t>(hw ? 1 : 0);
        w.put<int32_t>(0);
        w.raw(t.lo.data(), sizeof(double) * t.ndim);  // root box
        w.raw(t.hi.data(), sizeof(double) * t.ndim);
        w.raw(t.pts.data(), sizeof(double) * t.pts.size());
        if (hw) {  // device weights are in tree order
            std::vector<double> wtree(t.n);
            HIP_CHECK(hipMemcpy(wtree.data(), op.W.as<double>() + op.w_off[p], sizeof(double) * t.n,
                                hipMemcpyDeviceToHost));
            wt.assign(t.n, 0.0);
            for (int64_t q = 0; q < t.n; q++) wt[t.perm[q]] = wtree[q];
            w.raw(wt.data(), sizeof(double) * t.n);
        }
    }
    w.put<int64_t>(op.nblk);
    w.raw(op.rank.data(), sizeof(int32_t) * op.rank.size());
    w.raw(op.flags.data(), sizeof(int32_t) * op.flags.size());
    w.put<int64_t>(op.n_np);
    w.put<int64_t>(op.n_up);
    w.put<int64_t>(op.n_vp);
    std::vector<char> buf;
    w.device(op.NP64.p, sizeof(double) * op.n_np, buf);
    w.device(op.UP64.p, sizeof(double) * op.n_up, buf);
    w.device(op.VP64.p, sizeof(double) * op.n_vp, buf);
    if (op.sys_size >= 0) {
        w.put<int64_t>(op.n_sk);
        w.put<int64_t>(op.n_g);
        w.device(op.SK.p, sizeof(int32_t) * 2 * op.n_sk, buf);
        w.device(op.GA.p, sizeof(double) * op.n_g, buf);
    }
    const uint64_t h = w.h;
    w.put<uint64_t>(h);
    w.f.close();
    if (!w.f) throw ArgError{HODLR_E_ARG, "write failed"};
}

hodlr_op *load_op(hodlr_ctx *ctx, const char *path, int prec_flags) {
    Reader r(path);
    char magic[8];
    r.raw(magic, 8);
    if (std::memcmp(magic, kMagic, 8) != 0) throw ArgError{HODLR_E_ARG, "not an operator file"};
    const uint32_t version = r.get<uint32_t>();
    if (version != 1 && version != kVersion) throw ArgError{HODLR_E_ARG, "unsupported operator file version"};
    if (r.get<uint32_t>() != kEndian) throw ArgError{HODLR_E_ARG, "operator file has the wrong byte order"};
    const int P = r.get<int32_t>();
    const double eps = r.get<double>();
    const int32_t sys_field = r.get<int32_t>();
    const int sys_size = version == 1 ? -1 : (sys_field < 0 ? -1 : sys_field);
    if (P < 1 || P > (1 << 24) || !(eps > 0 && eps < 1)) throw ArgError{HODLR_E_ARG, "corrupt operator header"};
    if (ctx->device < 0) throw ArgError{HODLR_E_STATE, "host-only context"};
    auto op = std::make_unique<hodlr_op>();
    op->ctx = ctx;
    op->eps = eps;
    op->sys_size = sys_size;
    op->prec_flags = prec_flags | HODLR_PREC_F64;
    std::vector<std::vector<double>> weights(P);
    for (int p = 0; p < P; p++) {
        auto t = std::make_unique<hodlr_tree>();
        t->ctx = ctx;
        t->ndim = r.get<int32_t>();
        t->nmax = r.get<int32_t>();
        t->dprime = r.get<int32_t>();
        const int family = r.get<int32_t>();
        t->n = r.get<int64_t>();
        const double scale = r.get<double>(), diag = r.get<double>();
        const int hw = r.get<int32_t>();
        (void)r.get<int32_t>();
        if (t->ndim < 1 || t->ndim > kMaxDim || t->n < 0 || t->n > INT32_MAX || t->nmax < 1 || family < 0 ||
            family > 2 || (hw != 0 && hw != 1))
            throw ArgError{HODLR_E_ARG, "corrupt problem header"};
        if (((uint64_t)t->n * (uint64_t)(t->ndim + hw) + 2 * (uint64_t)t->ndim) * sizeof(double) > r.left())
            throw ArgError{HODLR_E_ARG, "operator file is truncated (problem points)"};
        t->lo.resize(t->ndim);
        t->hi.resize(t->ndim);
        r.raw(t->lo.data(), sizeof(double) * t->ndim);
        r.raw(t->hi.data(), sizeof(double) * t->ndim);
        t->pts.resize((size_t)t->n * t->ndim);
        r.raw(t->pts.data(), sizeof(double) * t->pts.size());
        if (hw) {
            weights[p].resize(t->n);
            r.raw(weights[p].data(), sizeof(double) * t->n);
        }
        build_tree(*t);
        hodlr_kernel_desc kd;
        kd.family = family;
        kd.ndim = t->ndim;
        kd.scale = scale;
        kd.diag = diag;
        kd.weight = hw ? weights[p].data() : nullptr;
        op->trees.push_back(t.get());
        op->kds.push_back(kd);
        op->owned_trees.push_back(std::move(t));
    }
    const int64_t nb = r.get<int64_t>();
    if (nb < 0 || (uint64_t)nb * 2 * sizeof(int32_t) > r.left()) throw ArgError{HODLR_E_ARG, "corrupt block count"};
    op->preset_rank.resize(nb);
    op->preset_flags.resize(nb);
    r.raw(op->preset_rank.data(), sizeof(int32_t) * nb);
    r.raw(op->preset_flags.data(), sizeof(int32_t) * nb);
    const int64_t n_np = r.get<int64_t>(), n_up = r.get<int64_t>(), n_vp = r.get<int64_t>();
    op->fill_from = [&r, n_np, n_up, n_vp](hodlr_op &o) {
        if (o.n_np != n_np || o.n_up != n_up || o.n_vp != n_vp)
            throw ArgError{HODLR_E_ARG, "stored arenas do not match the rebuilt layout"};
        std::vector<char> buf;
        hipStream_t s = o.ctx->stream;
        r.device(o.NP64.p, sizeof(double) * n_np, buf, s);
        r.device(o.UP64.p, sizeof(double) * n_up, buf, s);
        r.device(o.VP64.p, sizeof(double) * n_vp, buf, s);
        if (o.sys_size >= 0) {
            const int64_t n_sk = r.get<int64_t>(), n_g = r.get<int64_t>();
            if (n_sk != o.n_sk || n_g != o.n_g)
                throw ArgError{HODLR_E_ARG, "stored skeleton sizes do not match the rebuilt layout"};
            r.device(o.SK.p, sizeof(int32_t) * 2 * n_sk, buf, s);
            o.GA.alloc(sizeof(double) * (size_t)n_g);
            r.device(o.GA.p, sizeof(double) * n_g, buf, s);
        }
        const uint64_t want = r.h;
        if (r.get<uint64_t>() != want) throw ArgError{HODLR_E_ARG, "operator file checksum mismatch"};
    };
    HIP_CHECK(hipSetDevice(ctx->device));
    assemble(*op);
    op->fill_from = nullptr;
    op->preset_rank.clear();
    op->preset_flags.clear();
    for (auto &k : op->kds) k.weight = nullptr;
    op->kd.weight = nullptr;
    return op.release();
}

}  // namespace hodlr
