// op.cpp -- operator assembly and the matvec driver (host side).
//
// Assembly (HDD_tree.hpp:76-83, HDD_node.hpp:118-147) runs in four steps,
// over one problem or over a batch of independent problems at once (their
// block-diagonal operator; positions and nodes numbered globally, problem
// after problem):
//   1. enumerate the admissible blocks in the reference's order (level,
//      node, interaction-list index), skipping empty nodes;
//   2. ACA rank pass over every block (batched, chunked scratch);
//   3. lay out the arenas from the exact ranks:
//        near field  NP: per leaf an nc x (sum of neighbour sizes) panel,
//        U panels    UP: per row node a the m_a x R_a matrix [U_ab1 | U_ab2 ..]
//                        (dense-fallback blocks contribute K_ab), stored
//                        leaf-blocked: each leaf's rows form a contiguous
//                        nc x R_a column-major slice;
//        V panels    VP: per column node b the n_b x T_b matrix of the V
//                        factors of every compressed block (., b), stored as
//                        256-column tiles, each a contiguous row-major block;
//   4. fill pass: near field, dense fallbacks, and the ACA again, writing
//      factors into place.
// With sys_size >= 0, compressed blocks whose column node holds more than
// sys_size points are skeletons (LowRank_matrix.hpp:22-36): they get no
// panel columns; the fill pass writes their pivot positions and the host
// forms G = C^+ of every r x r core (LowRank_matrix.hpp:685-698), and the
// matvec evaluates their entries on the fly (skeleton.hip).
#include "op.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>

namespace hodlr {

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// HODLR_VERBOSE=1 prints assembly progress to stderr
bool verbose() {
    static const bool v = [] {
        const char *e = std::getenv("HODLR_VERBOSE");
        return e && e[0] && e[0] != '0';
    }();
    return v;
}

// ACA kernel class of a block with rank budget cap: 8, 16, 32 = one wave
// per block with the factors in registers (rank capacity = class >= the
// budget, so max(m, n) <= 64 and min(m, n) <= 32); 64, 256, 1024 =
// workgroup of that size with the factors in the scratch pool
int aca_block_size(int m, int n, int cap) {
    const int mx = m > n ? m : n, mn = m < n ? m : n;
    const int c = mn < cap ? mn : cap;
    if (mx <= 64 && mn <= 32) return c <= 8 ? 8 : (c <= 16 ? 16 : 32);
    if (mx <= 128) return 64;
    if (mx <= 4096) return 256;
    return 1024;
}

bool aca_in_registers(int cls) { return cls < 64; }

// ratio of per-segment to flattened load rounds at or above which the 1-4 RHS
// panel product defaults to the flattened 16-byte form.  Measured (FP32 / FP64
// panel ms, per-segment vs flattened): it wins or ties from 2.6x (16^4
// Gauss-Legendre with 16-row leaves, 16-19x: 1.67 / 1.92 vs 0.66 / 1.08;
// 10^4 random points, 2.6-4.2x: 0.041 / 0.055 vs 0.027 / 0.053; 2-D random,
// 3.4-5.3x: 0.016 / 0.018 vs 0.012 / 0.018) and loses at 2.0x and below
// (the 100^3 grid 1.67 / 3.18 vs 1.77 / 3.31; 2 10^4 random points 0.052 /
// 0.097 vs 0.063 / 0.120); the 16^4 d' = 3 operator (2.0-2.5x) ties
constexpr double kFlatGain = 2.5;

// LDS a workgroup ACA may take for its block's points ((m + n)(NDIM + 1) 8
// bytes, plus 8 per column with weights): the 64 class always fits
// (m + n <= 256); the 256 class up to 32 KiB (about 1000 points in 3-D,
// four workgroups per CU); the 1024 class (max(m, n) > 4096) never.
// Measured on the N = 10^6 grid and a 200k-point build (kernel time): 256
// class 790 -> 772 ms (16 KiB: 774 ms), 64 class 99 -> 84 ms
int64_t aca_pts_lds_limit(int cls) {
    if (cls == 64) return 24576;
    if (cls == 256) return 32768;
    return 0;
}

// a block that may become a skeleton runs in aca_kernel in both passes
// (only that kernel writes pivot positions)
int aca_class(const hodlr_op &op, int64_t b, int cap) {
    const int c = aca_block_size(op.b_m[b], op.b_n[b], cap);
    return (op.sys_size >= 0 && op.b_n[b] > op.sys_size && c < 64) ? 64 : c;
}

// host threads one assembly may use (set per thread by callers that run
// several assemblies at once); 0 = up to 16
thread_local int tl_threads = 0;

// fn(begin, end) over [0, n) on the assembly's host threads, at least
// `grain` items per thread (inline for small n)
template <class F>
void parallel_for(int64_t n, F &&fn, int64_t grain = 1 << 16) {
    const int64_t kMin = grain;
    int nt = tl_threads > 0 ? tl_threads
                            : (int)std::min<int64_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
    nt = (int)std::min<int64_t>(nt, (n + kMin - 1) / kMin);
    if (nt <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t per = (n + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        const int64_t a = t * per, b = std::min(n, a + per);
        if (a < b) th.emplace_back([&fn, a, b] { fn(a, b); });
    }
    for (auto &x : th) x.join();
}

// A host table every entry of which is written before it is read: unlike a
// std::vector it is not value-initialised on allocation (layout tables run
// to GBs and are written by parallel passes).
template <class T>
struct Table {
    std::unique_ptr<T[]> p;
    size_t n = 0;
    explicit Table(size_t count) : p(new T[count > 0 ? count : 1]), n(count) {}
    T &operator[](size_t i) { return p[i]; }
    const T &operator[](size_t i) const { return p[i]; }
    T *data() { return p.get(); }
    const T *data() const { return p.get(); }
    size_t size() const { return n; }
};
template <class T>
void upload(DevBuf &b, const Table<T> &v, hipStream_t s = nullptr, bool temp = false) {
    upload(b, v.data(), v.size(), s, temp);
}

uint64_t block_seed(int level, int64_t row_node, int64_t col_node) {
    uint64_t s = 0x243F6A8885A308D3ULL ^ ((uint64_t)(level + 1) * 0x9E3779B97F4A7C15ULL);
    s ^= (uint64_t)row_node * 0xBF58476D1CE4E5B9ULL;
    s ^= (uint64_t)col_node * 0x94D049BB133111EBULL;
    return s;
}

// A_ij on the host from tree-order coordinates X and weights W of one
// problem (user_kernel.hpp:50-58,105-113,171-177; the same arithmetic as the
// device entry())
double host_entry(const hodlr_kernel_desc &k, int nd, const double *X, const double *W, int64_t i, int64_t j) {
    if (i == j) return k.diag;
    double s = 0.0;
    for (int d = 0; d < nd; d++) {
        const double t = X[i * nd + d] - X[j * nd + d];
        const double t2 = t * t;
        s = s + t2;
    }
    const double r = std::sqrt(s);
    const double f = k.family == 1 ? 1.0 / r : (k.family == 0 ? std::log(r) : std::exp(-r));
    double v = k.scale * f;
    if (W) v = v * W[j];
    return v;
}

// G = C^+ for the r x r core C of a skeleton block (column-major, a is
// destroyed): the linear map of the column-pivoted Householder QR solve the
// reference applies (ColPivHouseholderQR, LowRank_matrix.hpp:57,698).  The
// column of largest remaining norm is pivoted first; the rank is cut at the
// first |R_kk| <= eps r |R_00|; G = P [R11^-1 Q1^T ; 0].
void core_pinv(int r, std::vector<double> &a, double *G) {
    std::vector<double> tau(r, 0.0), w(r);
    std::vector<int> perm(r);
    for (int j = 0; j < r; j++) perm[j] = j;
    double maxpiv = 0.0;
    for (int k = 0; k < r; k++) {
        int best = k;
        double bn = -1.0;
        for (int j = k; j < r; j++) {
            double s = 0.0;
            for (int i = k; i < r; i++) s += a[(size_t)j * r + i] * a[(size_t)j * r + i];
            if (s > bn) {
                bn = s;
                best = j;
            }
        }
        if (best != k) {
            for (int i = 0; i < r; i++) std::swap(a[(size_t)k * r + i], a[(size_t)best * r + i]);
            std::swap(perm[k], perm[best]);
        }
        double *x = &a[(size_t)k * r];
        const double alpha = x[k];
        double s = 0.0;
        for (int i = k + 1; i < r; i++) s += x[i] * x[i];
        double beta = alpha, t = 0.0;
        if (s != 0.0) {
            beta = -std::copysign(std::sqrt(alpha * alpha + s), alpha);
            t = (beta - alpha) / beta;
            const double sc = 1.0 / (alpha - beta);
            for (int i = k + 1; i < r; i++) x[i] *= sc;
        }
        x[k] = beta;
        tau[k] = t;
        if (k == 0) maxpiv = std::fabs(beta);
        if (t == 0.0) continue;
        for (int j = k + 1; j < r; j++) {
            double *y = &a[(size_t)j * r];
            double d = y[k];
            for (int i = k + 1; i < r; i++) d += x[i] * y[i];
            d *= t;
            y[k] -= d;
            for (int i = k + 1; i < r; i++) y[i] -= d * x[i];
        }
    }
    const double thr = 2.220446049250313e-16 * (double)r * maxpiv;
    int rk = 0;
    while (rk < r && std::fabs(a[(size_t)rk * r + rk]) > thr) rk++;
    for (int c = 0; c < r; c++) {
        std::fill(w.begin(), w.end(), 0.0);
        w[c] = 1.0;
        for (int k = 0; k < r; k++) {  // w = Q^T e_c
            if (tau[k] == 0.0) continue;
            const double *v = &a[(size_t)k * r];
            double d = w[k];
            for (int i = k + 1; i < r; i++) d += v[i] * w[i];
            d *= tau[k];
            w[k] -= d;
            for (int i = k + 1; i < r; i++) w[i] -= d * v[i];
        }
        for (int i = rk - 1; i >= 0; i--) {
            double v = w[i];
            for (int j = i + 1; j < rk; j++) v -= a[(size_t)j * r + i] * w[j];
            w[i] = v / a[(size_t)i * r + i];
        }
        for (int i = 0; i < r; i++) G[(size_t)c * r + perm[i]] = i < rk ? w[i] : 0.0;
    }
}

// Runs one ACA pass over `ids` (block ids) with per-block caps, chunked so
// that the scratch pool stays within `budget` doubles.
void run_aca_pass(hodlr_op &op, const std::vector<int64_t> &ids, const std::vector<int32_t> &cap, int fill,
                  DevBuf &pool, int64_t budget, DevBuf &d_rank, DevBuf &d_flags, const DevBuf *d_dest,
                  hipStream_t s, const AcaStage &stage = AcaStage{nullptr, nullptr, 0, nullptr, nullptr, 0},
                  const std::vector<uint8_t> *stageable = nullptr) {
    // group by kernel class (largest class first), large blocks first
    // inside a class: bucket by class, then a counting sort on m + n
    static const int kClasses[] = {1024, 256, 64, 32, 16, 8};
    auto class_index = [](int c) {
        for (int k = 0; k < 6; k++)
            if (kClasses[k] == c) return k;
        return 0;
    };
    std::vector<int64_t> order;
    order.reserve(ids.size());
    {
        std::vector<std::vector<int64_t>> bucket(6);
        for (int64_t b : ids) bucket[class_index(aca_class(op, b, cap[b]))].push_back(b);
        for (auto &bk : bucket) {
            int64_t kmax = 0;
            for (int64_t b : bk) kmax = std::max<int64_t>(kmax, (int64_t)op.b_m[b] + op.b_n[b]);
            if (kmax > (1 << 16)) {
                std::stable_sort(bk.begin(), bk.end(), [&](int64_t a, int64_t b) {
                    return (int64_t)op.b_m[a] + op.b_n[a] > (int64_t)op.b_m[b] + op.b_n[b];
                });
                order.insert(order.end(), bk.begin(), bk.end());
                continue;
            }
            std::vector<int64_t> cnt(kmax + 2, 0);
            for (int64_t b : bk) cnt[kmax - (op.b_m[b] + op.b_n[b])]++;
            int64_t run = 0;
            for (auto &c : cnt) {
                const int64_t t = c;
                c = run;
                run += t;
            }
            const size_t base = order.size();
            order.resize(base + bk.size());
            for (int64_t b : bk) order[base + cnt[kmax - (op.b_m[b] + op.b_n[b])]++] = b;
        }
    }
    DevBuf d_blocks;
    size_t i = 0;
    std::vector<AcaBlock> chunk;
    std::vector<int64_t> offs;
    while (i < order.size()) {
        const int bs = aca_class(op, order[i], cap[order[i]]);
        // the chunk: consecutive blocks of one class within the scratch budget
        int64_t used = 0;
        const size_t i0 = i;
        offs.clear();
        while (i < order.size()) {
            const int64_t b = order[i];
            if (aca_class(op, b, cap[b]) != bs) break;
            const int64_t need = aca_in_registers(bs) ? 0 : aca_scratch_doubles(op.b_m[b], op.b_n[b], cap[b]);
            if (i > i0 && used + need > budget) break;
            offs.push_back(used);
            used += need;
            i++;
        }
        chunk.resize(i - i0);
        parallel_for((int64_t)chunk.size(), [&](int64_t q0, int64_t q1) {
            for (int64_t q = q0; q < q1; q++) {
                const int64_t b = order[i0 + q];
                const int p = op.b_prob[b];
                AcaBlock &a = chunk[q];
                a.row0 = op.node_start[op.b_row[b]];
                a.col0 = op.node_start[op.b_col[b]];
                a.row0l = a.row0 - op.pos_base[p];
                a.col0l = a.col0 - op.pos_base[p];
                a.scratch = offs[q];
                a.ord = op.ord.as<int32_t>() + op.ord_off[p] + (int64_t)op.b_level[b] * op.trees[p]->n;
                a.m = op.b_m[b];
                a.n = op.b_n[b];
                a.cap = cap[b];
                a.id = (int32_t)b;
                a.seed = block_seed(op.b_level[b], op.b_rowj[b], op.b_colj[b]);
                a.prob = p;
                a.stage = (stage.buf && stageable) ? (*stageable)[b] : 0;
            }
        });
        if ((int64_t)(pool.bytes / sizeof(double)) < used) pool.alloc((size_t)used * sizeof(double), s);
        upload(d_blocks, chunk, s, true);
        // the workgroup forms stage a block's points in LDS when they fit
        // in the class's allowance (the wave forms always do, statically);
        // a launch reserves what its largest fitting block needs, nothing
        // when none fits (the reservation costs occupancy)
        AcaStage st = stage;
        st.pts_lds = 0;
        if (!aca_in_registers(bs)) {
            const int64_t lim = aca_pts_lds_limit(bs);
            int64_t want = 0;
            for (size_t q = 0; q < chunk.size(); q++) {
                const int p = chunk[q].prob;
                const int64_t rec = (int64_t)(chunk[q].m + chunk[q].n) * (op.kds[p].ndim + 1) * 8 +
                                    (op.w_off[p] >= 0 ? (int64_t)chunk[q].n * 8 : 0);
                if (rec <= lim) want = std::max(want, rec);
            }
            st.pts_lds = (int32_t)want;
        }
        HIP_CHECK(op.entries->aca(bs, d_blocks.as<AcaBlock>(), (int64_t)chunk.size(), op.kdev.as<KernelDev>(), op.eps,
                             pool.as<double>(), d_rank.as<int32_t>(), d_flags.as<int32_t>(), fill,
                             d_dest ? d_dest->as<FillDest>() : nullptr, op.leaf_start.as<int64_t>(),
                             op.leaf_cnt.as<int32_t>(), op.UP64.as<double>(), op.VP64.as<double>(),
                             op.SK.as<int32_t>(), st, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (verbose())
            std::fprintf(stderr, "[hodlr] aca %s pass: %zu/%zu blocks (workgroup %d, scratch %.2f GB)\n",
                         fill ? "fill" : "rank", i, order.size(), bs, used * 8e-9);
    }
}

}  // namespace

void assemble(hodlr_op &op) {
    const double t_start = now_s();
    double t_mark = t_start;
    auto mark = [&](const char *phase) {  // HODLR_VERBOSE phase timing
        const double t = now_s();
        if (verbose()) std::fprintf(stderr, "[hodlr] %-12s %8.3f s\n", phase, t - t_mark);
        t_mark = t;
    };
    hipStream_t s = op.ctx->stream;
    const int P = (int)op.trees.size();
    op.tree = op.trees[0];
    op.kd = op.kds[0];
    if (!op.entries) op.entries = &builtin_entry_launchers();

    // ---- global numbering of positions and nodes ----
    std::vector<int64_t> node_base(P + 1, 0);
    op.pos_base.assign(P + 1, 0);
    for (int p = 0; p < P; p++) {
        op.pos_base[p + 1] = op.pos_base[p] + op.trees[p]->n;
        node_base[p + 1] = node_base[p] + op.trees[p]->level_off.back();
    }
    const int64_t N = op.pos_base[P];
    const int64_t nn = node_base[P];
    op.N = N;
    op.node_start.assign(nn, 0);
    std::vector<int64_t> count(nn, 0);
    for (int p = 0; p < P; p++) {
        const hodlr_tree &t = *op.trees[p];
        for (int64_t g = 0; g < t.level_off.back(); g++) {
            op.node_start[node_base[p] + g] = op.pos_base[p] + t.start[g];
            count[node_base[p] + g] = t.count[g];
        }
    }

    // ---- tree-order inputs: coordinates, weights, leaf of every position,
    //      per (problem, level) ascending-id order maps ----
    std::vector<double> X, W;  // host copies, kept for the skeleton cores
    std::vector<int64_t> xoff(P), woff(P, -1);
    {
        for (int p = 0; p < P; p++) {
            const hodlr_tree &t = *op.trees[p];
            xoff[p] = (int64_t)X.size();
            for (int64_t q = 0; q < t.n; q++)
                for (int k = 0; k < t.ndim; k++) X.push_back(t.pts[t.perm[q] * t.ndim + k]);
            if (op.kds[p].weight) {
                woff[p] = (int64_t)W.size();
                for (int64_t q = 0; q < t.n; q++) W.push_back(op.kds[p].weight[t.perm[q]]);
            }
        }
        upload(op.X, X, s);
        upload(op.W, W, s);
        op.w_off = woff;
        std::vector<KernelDev> kd(P);
        for (int p = 0; p < P; p++) {
            kd[p].family = op.kds[p].family;
            kd[p].ndim = op.trees[p]->ndim;
            kd[p].scale = op.kds[p].scale;
            kd[p].diag = op.kds[p].diag;
            kd[p].X = op.X.as<double>() + xoff[p];
            kd[p].W = woff[p] >= 0 ? op.W.as<double>() + woff[p] : nullptr;
            kd[p].pos0 = op.pos_base[p];
            kd[p].obj = op.kobj.p;
        }
        upload(op.kdev, kd, s);
        std::vector<int64_t> ls(N > 0 ? N : 1, 0);
        std::vector<int32_t> lc(N > 0 ? N : 1, 0);
        op.ord_off.assign(P, 0);
        int64_t ord_total = 0;
        for (int p = 0; p < P; p++) {
            op.ord_off[p] = ord_total;
            ord_total += (int64_t)op.trees[p]->nlevels * op.trees[p]->n;
        }
        std::vector<int32_t> ord((size_t)ord_total);
        // one sort per inner node with two or more points (leaves keep
        // ascending ids); nodes own disjoint ranges and ids are distinct, so
        // the sorts run in parallel with the result of a sequential pass
        struct SortTask {
            int p, l;
            int64_t g;
        };
        std::vector<SortTask> sorts;
        for (int p = 0; p < P; p++) {
            const hodlr_tree &t = *op.trees[p];
            const int L = t.nlevels - 1;
            const int64_t pb = op.pos_base[p];
            for (int64_t j = 0; j < t.level_size(L); j++) {
                const int64_t g = t.node(L, j);
                for (int64_t q = t.start[g]; q < t.start[g] + t.count[g]; q++) {
                    ls[pb + q] = pb + t.start[g];
                    lc[pb + q] = (int32_t)t.count[g];
                }
            }
            for (int l = 0; l <= L; l++) {
                int32_t *o = ord.data() + op.ord_off[p] + (size_t)l * t.n;
                for (int64_t q = 0; q < t.n; q++) o[q] = (int32_t)(pb + q);
                if (l == L) continue;
                for (int64_t j = 0; j < t.level_size(l); j++)
                    if (t.count[t.node(l, j)] > 1) sorts.push_back(SortTask{p, l, t.node(l, j)});
            }
        }
        parallel_for(
            (int64_t)sorts.size(),
            [&](int64_t k0, int64_t k1) {
                for (int64_t k = k0; k < k1; k++) {
                    const SortTask &T = sorts[k];
                    const hodlr_tree &t = *op.trees[T.p];
                    const int64_t pb = op.pos_base[T.p];
                    int32_t *o = ord.data() + op.ord_off[T.p] + (size_t)T.l * t.n;
                    std::sort(o + t.start[T.g], o + t.start[T.g] + t.count[T.g],
                              [&](int32_t a, int32_t b) { return t.perm[a - pb] < t.perm[b - pb]; });
                }
            },
            16);
        upload(op.leaf_start, ls, s);
        upload(op.leaf_cnt, lc, s);
        upload(op.ord, ord, s);
    }

    mark("inputs");
    // ---- 1. blocks in reference order, problem by problem ----
    {
        size_t tot = 0;
        for (int p = 0; p < P; p++) tot += op.trees[p]->il.size();
        for (auto *v : {&op.b_row, &op.b_col, &op.b_rowj, &op.b_colj}) v->reserve(tot);
        for (auto *v : {&op.b_m, &op.b_n}) v->reserve(tot);
        op.b_level.reserve(tot);
        op.b_prob.reserve(tot);
    }
    for (int p = 0; p < P; p++) {
        const hodlr_tree &t = *op.trees[p];
        for (int l = 0; l < t.nlevels; l++)
            for (int64_t j = 0; j < t.level_size(l); j++) {
                const int64_t g = t.node(l, j);
                if (!t.count[g]) continue;
                for (int64_t a = t.il_ptr[g]; a < t.il_ptr[g + 1]; a++) {
                    const int64_t gb = t.node(l, t.il[a]);
                    if (!t.count[gb]) continue;
                    op.b_level.push_back(l);
                    op.b_prob.push_back(p);
                    op.b_row.push_back(node_base[p] + g);
                    op.b_col.push_back(node_base[p] + gb);
                    op.b_rowj.push_back(j);
                    op.b_colj.push_back(t.il[a]);
                    op.b_m.push_back((int32_t)t.count[g]);
                    op.b_n.push_back((int32_t)t.count[gb]);
                }
            }
    }
    op.nblk = (int64_t)op.b_level.size();
    const int64_t nb = op.nblk;

    mark("block list");
    // ---- 2. rank pass ----
    const double t_rank0 = now_s();
    DevBuf d_rank, d_flags, pool;
    d_rank.alloc(sizeof(int32_t) * (nb + 1), s);
    d_flags.alloc(sizeof(int32_t) * (nb + 1), s);
    const size_t free_b = device_free_bytes();
    // ACA scratch per launch: at most 1/16 of free memory and 4 GiB, so
    // concurrent builds neither crowd out operators nor pay for clearing
    // freshly mapped memory on every chunk.  (Up to 32 GiB per launch cut
    // the ACA kernels' time at N = 10^6 from 1.09 s to 0.69 s, but the
    // assembly took longer: 1.7 s instead of 1.45 s for the rank pass, once
    // 7.1 s, allocating the larger pool in a process that followed another.)
    const int64_t budget = std::max<int64_t>(
        std::min<int64_t>((int64_t)(free_b / 16 / sizeof(double)), ((int64_t)4 << 30) / (int64_t)sizeof(double)),
        (int64_t)1 << 24);
    std::vector<int32_t> cap(nb);
    std::vector<int64_t> todo(nb);
    // first rank budget: 128, but 16 for blocks of the one-wave form with
    // 17-32 points on their short side.  Its registers hold two factor
    // values per rank slot and lane: with room for 32 the kernel takes 256
    // VGPRs (one wave per SIMD), with 16 it takes 195 (two), and ranks stay
    // below 16 for ~99% of such blocks (1/r at eps_ACA 1e-8).  A block that
    // reaches the budget is run again with min(m, n) (the retry below), which
    // repeats the same steps: the result does not depend on the budget.
    for (int64_t b = 0; b < nb; b++) {
        const int mn = std::min(op.b_m[b], op.b_n[b]);
        const int mx = std::max(op.b_m[b], op.b_n[b]);
        cap[b] = (mx <= 64 && mn > 16 && mn <= 32) ? 16 : std::min(mn, 128);
        todo[b] = b;
    }
    op.rank.assign(nb, 0);
    op.flags.assign(nb, 0);
    if (!op.preset_rank.empty()) {  // ranks of a stored operator
        if ((int64_t)op.preset_rank.size() != nb || (int64_t)op.preset_flags.size() != nb)
            throw ArgError{HODLR_E_ARG, "stored block count does not match the rebuilt tree"};
        // every stored rank and flag is checked against its block before it
        // drives the layout: flags 0 or 1 (dense), 0 <= rank <= min(m, n),
        // dense exactly when rank >= min(m, n) - 1 (LowRank_matrix.hpp:348)
        for (int64_t b = 0; b < nb; b++) {
            const int mn = std::min(op.b_m[b], op.b_n[b]);
            const int32_t r = op.preset_rank[b], f = op.preset_flags[b];
            if ((f != 0 && f != 1) || r < 0 || r > mn || (f == 1) != (r >= mn - 1))
                throw ArgError{HODLR_E_ARG, "stored rank or flag out of range for block " + std::to_string(b)};
        }
        op.rank = op.preset_rank;
        op.flags = op.preset_flags;
        todo.clear();
    }
    // Factor staging: the rank pass keeps the factors of the compressed
    // blocks that fit (skeleton candidates excluded), and the fill step
    // copies them into their panels instead of running the ACA again.  The
    // buffer takes at most a quarter of free memory (1/16 under an
    // admission hook, where builds run concurrently) or the context's cap;
    // blocks beyond it are recomputed by the fill pass as before.
    DevBuf stage_buf, stage_cur, stage_off;
    AcaStage stage{nullptr, nullptr, 0, nullptr};
    std::vector<uint8_t> stageable(nb, 0);
    if (!todo.empty() && op.ctx->stage_bytes != 0) {
        // a skeleton candidate (n > sys_size) keeps its pivots, the others
        // their factors
        int64_t need = 0;
        for (int64_t b = 0; b < nb; b++) {
            const bool sk = op.sys_size >= 0 && op.b_n[b] > op.sys_size;
            stageable[b] = sk ? 2 : 1;
            need += sk ? (int64_t)cap[b] : (int64_t)cap[b] * (op.b_m[b] + op.b_n[b]);
        }
        const int64_t room = op.ctx->stage_bytes > 0
                                 ? op.ctx->stage_bytes / (int64_t)sizeof(double)
                                 : (int64_t)(free_b / (op.admit ? 16 : 4) / sizeof(double));
        const int64_t capd = std::min(need, room);
        if (capd > 0) {
            try {
                stage_buf.alloc(sizeof(double) * (size_t)capd, s);
                stage_cur.alloc(sizeof(unsigned long long), s);
                stage_off.alloc(sizeof(int64_t) * (size_t)(nb > 0 ? nb : 1), s);
                HIP_CHECK(hipMemsetAsync(stage_cur.p, 0, sizeof(unsigned long long), s));
                HIP_CHECK(hipMemsetAsync(stage_off.p, 0xff, sizeof(int64_t) * (size_t)(nb > 0 ? nb : 1), s));
                stage = AcaStage{stage_buf.as<double>(), stage_cur.as<unsigned long long>(), capd,
                                 stage_off.as<int64_t>(), nullptr, 0};
            } catch (const ArgError &) {  // no room: the fill pass recomputes every block
                stage_buf.release();
                stage_cur.release();
                stage_off.release();
            }
        }
    }
    // entries each block's ACA evaluates (its last run: a block run again
    // with a larger rank budget repeats the same steps first)
    DevBuf d_evals;
    if (!todo.empty()) {
        d_evals.alloc(sizeof(int64_t) * (size_t)nb, s);
        HIP_CHECK(hipMemsetAsync(d_evals.p, 0, sizeof(int64_t) * (size_t)nb, s));
        stage.evals = d_evals.as<int64_t>();
    }
    // Retries (blocks whose rank outgrew the first budget of 128) are few
    // and large: 4-D d' = 3 puts 240 blocks of 4096 x 4096 with ranks up
    // to ~800 there, each a single-workgroup ACA of seconds.  One launch
    // takes all of them when memory allows (a per-launch scratch of up to a
    // quarter of free memory instead of 4 GiB), and the staging buffer grows
    // so their factors are copied, not recomputed, by the fill step.
    const int64_t retry_budget =
        op.admit ? budget : std::max<int64_t>(budget, (int64_t)(free_b / 4 / sizeof(double)));
    const int64_t stage_room = (int64_t)(free_b / (op.admit ? 16 : 4) / sizeof(double));
    bool first_pass = true;
    while (!todo.empty()) {
        const int64_t pass_budget = first_pass ? budget : retry_budget;
        std::vector<int64_t> big, rest;
        if (first_pass)
            for (int64_t b : todo) (aca_class(op, b, cap[b]) == 1024 ? big : rest).push_back(b);
        first_pass = false;
        if (!big.empty() && !rest.empty()) {
            // The largest blocks (workgroup-1024 class) put few workgroups on
            // the chip per launch; a second host thread drives them on a
            // second stream while this one runs the other classes.
            HIP_CHECK(hipStreamSynchronize(s));  // rank, flag and staging buffers ready
            hipStream_t s2 = nullptr;
            HIP_CHECK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            DevSlab *slab = tl_slab;
            const int threads = tl_threads;
            std::exception_ptr err2;
            std::thread th([&] {
                tl_slab = slab;
                tl_threads = threads;
                DevBuf pool2;
                try {
                    run_aca_pass(op, big, cap, 0, pool2, pass_budget, d_rank, d_flags, nullptr, s2, stage,
                                 &stageable);
                    (void)hipStreamSynchronize(s2);
                } catch (...) {
                    err2 = std::current_exception();
                }
            });
            std::exception_ptr err1;
            try {
                run_aca_pass(op, rest, cap, 0, pool, pass_budget, d_rank, d_flags, nullptr, s, stage, &stageable);
            } catch (...) {
                err1 = std::current_exception();
            }
            th.join();
            (void)hipStreamSynchronize(s2);
            (void)hipStreamDestroy(s2);
            if (err1) std::rethrow_exception(err1);
            if (err2) std::rethrow_exception(err2);
        } else {
            run_aca_pass(op, todo, cap, 0, pool, pass_budget, d_rank, d_flags, nullptr, s, stage, &stageable);
        }
        std::vector<int32_t> rk(nb), fl(nb);
        if (nb) {
            HIP_CHECK(hipMemcpyAsync(rk.data(), d_rank.p, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(fl.data(), d_flags.p, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        std::vector<int64_t> again;
        for (int64_t b : todo) {
            op.rank[b] = rk[b];
            op.flags[b] = fl[b];
            if (fl[b] & 2) {  // rank budget exhausted: retry with a larger cap
                const int mn = std::min(op.b_m[b], op.b_n[b]);
                if (cap[b] >= mn) throw ArgError{HODLR_E_RANK, "ACA rank budget exhausted"};
                cap[b] = std::min<int64_t>(mn, (int64_t)cap[b] * 8);
                again.push_back(b);
            }
        }
        if (stage.buf && !again.empty()) {
            // room for the retried blocks' factors: a larger buffer holding
            // what is staged so far, the cursor continuing from where it is
            int64_t extra = 0;
            for (int64_t b : again)
                if (stageable[b])
                    extra += stageable[b] == 2 ? (int64_t)cap[b] : (int64_t)cap[b] * (op.b_m[b] + op.b_n[b]);
            unsigned long long cur = 0;
            HIP_CHECK(hipMemcpyAsync(&cur, stage.cursor, sizeof(cur), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            const int64_t want = std::min<int64_t>((int64_t)cur + extra, stage_room);
            if (want > stage.cap) {
                DevBuf bigger;
                try {
                    bigger.alloc(sizeof(double) * (size_t)want, s);
                } catch (const ArgError &) {  // no room: those blocks are recomputed by the fill pass
                }
                if (bigger.p) {
                    const int64_t keep = std::min<int64_t>((int64_t)cur, stage.cap);
                    if (keep > 0)
                        HIP_CHECK(hipMemcpyAsync(bigger.p, stage.buf, sizeof(double) * (size_t)keep,
                                                 hipMemcpyDeviceToDevice, s));
                    HIP_CHECK(hipStreamSynchronize(s));
                    stage_buf = std::move(bigger);
                    stage.buf = stage_buf.as<double>();
                    stage.cap = want;
                }
            }
        }
        todo.swap(again);
    }
    std::vector<int64_t> soff;  // staging offset per block, -1: not staged
    if (stage.buf) {
        soff.resize(nb);
        HIP_CHECK(hipMemcpyAsync(soff.data(), stage.off, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    double aca_entries = 0.0;
    if (d_evals.p) {
        std::vector<int64_t> ev(nb);
        HIP_CHECK(hipMemcpyAsync(ev.data(), d_evals.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        for (int64_t v : ev) aca_entries += (double)v;
        d_evals.release();
        stage.evals = nullptr;
    }
    const double t_rank1 = now_s();
    if (verbose()) std::fprintf(stderr, "[hodlr] rank pass %.3f s, %lld blocks\n", t_rank1 - t_rank0, (long long)nb);

    mark("rank pass");
    // skeleton blocks: compressed (not dense) with more than sys_size
    // columns (LowRank_matrix.hpp:27); a dense fallback stays dense (:348-363)
    std::vector<uint8_t> skel(nb, 0);
    op.n_skel = 0;
    if (op.sys_size >= 0)
        for (int64_t b = 0; b < nb; b++)
            if (!(op.flags[b] & 1) && op.b_n[b] > op.sys_size) {
                skel[b] = 1;
                op.n_skel++;
            }
    // ---- 3. layout ----
    // U panel of row node a: R[a] columns, its blocks in block order (a
    // dense-fallback block contributes its n columns of K); t of row node a:
    // TR[a] entries, the low-rank blocks only.  Dense blocks multiply x
    // directly in the panel product.
    std::vector<int64_t> R(nn, 0), TR(nn, 0), Tb(nn, 0), up_base(nn, 0), t_base(nn, 0), vp_base(nn, 0),
        vcol_base(nn, 0);
    std::vector<int32_t> width(nb), coff(nb), voff(nb), tcoff(nb, 0);
    std::vector<int64_t> rb_ptr(nn + 1, 0), rb_list(nb > 0 ? nb : 1);  // blocks per row node, block order
    for (int64_t b = 0; b < nb; b++) {
        const bool dense = op.flags[b] & 1;
        width[b] = dense ? op.b_n[b] : (skel[b] ? 0 : op.rank[b]);
        const int64_t a = op.b_row[b];
        coff[b] = (int32_t)R[a];
        R[a] += width[b];
        rb_ptr[a + 1]++;
        if (!dense && !skel[b]) {
            tcoff[b] = (int32_t)TR[a];
            TR[a] += op.rank[b];
            voff[b] = (int32_t)Tb[op.b_col[b]];
            Tb[op.b_col[b]] += op.rank[b];
        }
    }
    for (int64_t g = 0; g < nn; g++) rb_ptr[g + 1] += rb_ptr[g];
    {
        std::vector<int64_t> fillp(rb_ptr.begin(), rb_ptr.end() - 1);
        for (int64_t b = 0; b < nb; b++) rb_list[fillp[op.b_row[b]]++] = b;
    }
    // V panels are padded to a multiple of 4 columns so that every row of a
    // projection tile starts on a 4-element boundary
    std::vector<int64_t> Tpad(nn, 0);
    op.n_up = op.n_vp = op.n_t = 0;
    int64_t nvcol = 0;
    op.e_v = 0;
    for (int64_t g = 0; g < nn; g++) {
        Tpad[g] = (Tb[g] + 3) / 4 * 4;
        up_base[g] = op.n_up;
        op.n_up += count[g] * R[g];
        t_base[g] = op.n_t;
        op.n_t += TR[g];
        vp_base[g] = op.n_vp;
        op.n_vp += count[g] * Tpad[g];
        vcol_base[g] = nvcol;
        nvcol += Tpad[g];
        op.e_v += (double)count[g] * Tb[g];
    }
    op.e_u = (double)op.n_up;
    mark("panel sizes");

    // skeleton slots: the pivot rows of the blocks (., c) of column node c
    // are contiguous (row slots), as are the pivot columns of the blocks
    // (a, .) of row node a (column slots); cores in block order
    std::vector<int64_t> skr_off(nb, -1), skc_off(nb, -1), g_off(nb, -1);
    std::vector<int64_t> Sc, Sa, sc_base, sa_base;
    std::vector<SkelCore> skcore;
    op.n_sk = op.n_g = 0;
    op.e_skel = 0;
    if (op.n_skel) {
        Sc.assign(nn, 0);
        Sa.assign(nn, 0);
        sc_base.assign(nn, 0);
        sa_base.assign(nn, 0);
        for (int64_t b = 0; b < nb; b++)
            if (skel[b]) {
                Sc[op.b_col[b]] += op.rank[b];
                Sa[op.b_row[b]] += op.rank[b];
            }
        int64_t rc = 0, ra = 0;
        for (int64_t g = 0; g < nn; g++) {
            sc_base[g] = rc;
            rc += Sc[g];
            sa_base[g] = ra;
            ra += Sa[g];
        }
        op.n_sk = rc;
        if (op.n_sk > INT32_MAX) throw ArgError{HODLR_E_ARG, "too many skeleton slots"};
        std::vector<int64_t> fc(sc_base), fa(sa_base);
        for (int64_t b = 0; b < nb; b++) {
            if (!skel[b]) continue;
            const int64_t r = op.rank[b];
            skr_off[b] = fc[op.b_col[b]];
            fc[op.b_col[b]] += r;
            skc_off[b] = fa[op.b_row[b]];
            fa[op.b_row[b]] += r;
            g_off[b] = op.n_g;
            op.n_g += r * r;
            op.e_skel += (double)r * (double)(op.b_m[b] + op.b_n[b]);
            if (r) skcore.push_back(SkelCore{g_off[b], skr_off[b], skc_off[b], (int32_t)r, 0});
        }
    }

    // near field and panel work per non-empty leaf.  Each row node's blocks
    // are cut into runs once: a run of low-rank blocks multiplies their
    // contiguous t entries, a run of dense blocks whose column nodes are
    // adjacent multiplies x directly; every leaf below the node takes the
    // runs as segments of its slice of the node's U panel.  The leaves are
    // laid out in two parallel passes (sizes, then entries at prefix-sum
    // offsets), so the tables are the same as one sequential walk's.
    struct Run {
        int64_t coff, vec;
        int32_t len, arena;
    };
    std::vector<int64_t> run_ptr(nn + 1, 0);
    std::vector<Run> runs;
    {
        auto cut = [&](int64_t ga, Run *out) {  // runs of row node ga (written when out is set)
            int64_t nr = 0;
            Run cur{0, 0, 0, -1};
            for (int64_t e = rb_ptr[ga]; e < rb_ptr[ga + 1]; e++) {
                const int64_t b = rb_list[e];
                if (!width[b]) continue;
                const bool dn = op.flags[b] & 1;
                const int arena = dn ? 2 : 1;
                const int64_t vec = dn ? op.node_start[op.b_col[b]] : t_base[ga] + tcoff[b];
                if (cur.arena == arena && cur.vec + cur.len == vec) {
                    cur.len += width[b];
                    continue;
                }
                if (cur.arena >= 0) {
                    if (out) out[nr] = cur;
                    nr++;
                }
                cur = Run{(int64_t)coff[b], vec, width[b], arena};
            }
            if (cur.arena >= 0) {
                if (out) out[nr] = cur;
                nr++;
            }
            return nr;
        };
        parallel_for(
            nn,
            [&](int64_t g0, int64_t g1) {
                for (int64_t g = g0; g < g1; g++) run_ptr[g + 1] = R[g] ? cut(g, nullptr) : 0;
            },
            1024);
        for (int64_t g = 0; g < nn; g++) run_ptr[g + 1] += run_ptr[g];
        runs.resize((size_t)run_ptr[nn]);
        parallel_for(
            nn,
            [&](int64_t g0, int64_t g1) {
                for (int64_t g = g0; g < g1; g++)
                    if (R[g]) cut(g, runs.data() + run_ptr[g]);
            },
            1024);
    }
    struct LeafRef {
        int64_t g, j;
        int p;
    };
    std::vector<LeafRef> leaves;
    for (int p = 0; p < P; p++) {
        const hodlr_tree &t = *op.trees[p];
        const int L = t.nlevels - 1;
        for (int64_t j = 0; j < t.level_size(L); j++) {
            const int64_t g = node_base[p] + t.node(L, j);
            if (count[g]) leaves.push_back(LeafRef{g, j, p});
        }
    }
    const int64_t nleaf = (int64_t)leaves.size();
    // The leaf's segments, passed to emit(index, segment) in order: its
    // near-field blocks (self first, then its neighbours; a neighbour
    // adjacent in x extends the previous segment), then the runs of every
    // ancestor's U panel (and of the leaf's own).  Column sources are
    // written when columns is set.  Returns the segment count.
    auto leaf_segments = [&](const LeafRef &lf, int64_t np_base, int64_t *columns, int64_t &ncols, auto &&emit) {
        const hodlr_tree &t = *op.trees[lf.p];
        const int L = t.nlevels - 1, d = t.ndim;
        const int64_t nc = count[lf.g];
        int64_t ns = 0;
        Segment pend{0, 0, 0, -1};  // the near segment still open
        ncols = 0;
        auto add_near = [&](int64_t gb) {
            const int64_t c = count[gb];
            if (!c) return;
            if (pend.arena == 0 && pend.vec + pend.len == op.node_start[gb]) {
                pend.len += (int32_t)c;
            } else {
                if (pend.arena == 0) emit(ns++, pend);
                pend = Segment{np_base + ncols * nc, op.node_start[gb], (int32_t)c, 0};
            }
            if (columns)
                for (int64_t q = 0; q < c; q++) columns[ncols + q] = op.node_start[gb] + q;
            ncols += c;
        };
        add_near(lf.g);
        const int64_t gl = t.node(L, lf.j);
        for (int64_t a = t.nbr_ptr[gl]; a < t.nbr_ptr[gl + 1]; a++) add_near(node_base[lf.p] + t.node(L, t.nbr[a]));
        if (pend.arena == 0) emit(ns++, pend);
        for (int l = 1; l <= L; l++) {
            const int64_t ga = node_base[lf.p] + t.node(l, lf.j >> (d * (L - l)));
            if (!R[ga]) continue;
            const int64_t slice = up_base[ga] + R[ga] * (op.node_start[lf.g] - op.node_start[ga]);
            for (int64_t e = run_ptr[ga]; e < run_ptr[ga + 1]; e++)
                emit(ns++, Segment{slice + runs[e].coff * nc, runs[e].vec, runs[e].len, runs[e].arena});
        }
        return ns;
    };
    // pass 1: sizes per leaf (segments, column sources, near entries, quads,
    // panel items) and the 32-bit limits of the quad and panel tables
    std::vector<int64_t> l_seg(nleaf + 1, 0), l_col(nleaf + 1, 0), l_np(nleaf + 1, 0), l_quad(nleaf + 1, 0),
        l_item(nleaf + 1, 0);
    std::atomic<int> too_wide{0};
    parallel_for(
        nleaf,
        [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; i++) {
                const LeafRef &lf = leaves[i];
                const int64_t nc = count[lf.g];
                // row tiles: 64-row tiles (G = 1), then one of h_last rows
                const int h_last = (int)(nc - (nc - 1) / 64 * 64), G_last = 64 / h_last;
                int64_t n4 = 0, ng_full = 0, ng_last = 0, ncols = 0;
                const int64_t ns = leaf_segments(lf, 0, nullptr, ncols, [&](int64_t, const Segment &S) {
                    n4 += (S.len + 3) / 4;
                    if (S.vec + 4 * ((S.len - 1) / 4) > INT32_MAX) too_wide = 1;
                    ng_full += S.len;
                    ng_last += (S.len + G_last - 1) / G_last;
                });
                if (n4 > INT32_MAX || (nc > 64 && ng_full > INT32_MAX) || ng_last > INT32_MAX) too_wide = 2;
                l_seg[i + 1] = ns;
                l_col[i + 1] = ncols;
                l_np[i + 1] = nc * ncols;
                l_quad[i + 1] = n4;
                l_item[i + 1] = (nc + 63) / 64;
            }
        },
        256);
    if (too_wide == 1) throw ArgError{HODLR_E_ARG, "vector offset beyond 32 bits"};
    if (too_wide == 2) throw ArgError{HODLR_E_ARG, "leaf panel too wide"};
    for (int64_t i = 0; i < nleaf; i++) {
        l_seg[i + 1] += l_seg[i];
        l_col[i + 1] += l_col[i];
        l_np[i + 1] += l_np[i];
        l_quad[i + 1] += l_quad[i];
        l_item[i + 1] += l_item[i];
    }
    Table<NearLeaf> near((size_t)nleaf);
    Table<int64_t> colsrc((size_t)l_col[nleaf]);
    Table<Segment> segs((size_t)l_seg[nleaf]);
    Table<QuadDesc> quads((size_t)l_quad[nleaf]);
    Table<PanelItem> panel((size_t)l_item[nleaf]);
    // pass 2: the entries
    parallel_for(
        nleaf,
        [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; i++) {
                const LeafRef &lf = leaves[i];
                const int64_t nc = count[lf.g];
                const int64_t seg0 = l_seg[i];
                Segment *S = segs.data() + seg0;
                int64_t ncols = 0;
                const int64_t nseg = leaf_segments(lf, l_np[i], colsrc.data() + l_col[i], ncols,
                                                   [&](int64_t q, const Segment &sg) { S[q] = sg; });
                NearLeaf nl;
                nl.np_base = l_np[i];
                nl.row0 = op.node_start[lf.g];
                nl.colsrc = l_col[i];
                nl.nc = (int32_t)nc;
                nl.ncols = (int32_t)ncols;
                nl.prob = lf.p;
                nl.pad = 0;
                near[i] = nl;
                // the leaf's quads of columns (shared by its row tiles)
                QuadDesc *qd = quads.data() + l_quad[i];
                for (int64_t q = 0; q < nseg; q++)
                    for (int32_t c = 0; c < S[q].len; c += 4) {
                        qd->mat = S[q].mat + (int64_t)c * nc;
                        qd->vec = (int32_t)(S[q].vec + c);
                        qd->info = S[q].arena | (std::min<int32_t>(4, S[q].len - c) << 8);
                        qd++;
                    }
                PanelItem *pi = panel.data() + l_item[i];
                for (int64_t r0 = 0; r0 < nc; r0 += 64, pi++) {
                    pi->y0 = op.node_start[lf.g] + r0;
                    pi->seg0 = seg0;
                    pi->nseg = (int32_t)nseg;
                    pi->h = (int32_t)std::min<int64_t>(64, nc - r0);
                    pi->ld = (int32_t)nc;
                    pi->r0 = (int32_t)r0;
                    const int G = 64 / pi->h;
                    int64_t n4 = 0, ng = 0;
                    for (int64_t q = 0; q < nseg; q++) {
                        n4 += (S[q].len + 3) / 4;
                        ng += (S[q].len + G - 1) / G;
                    }
                    pi->nq4 = (int32_t)n4;
                    pi->nqg = (int32_t)ng;
                    pi->q0 = l_quad[i];
                }
            }
        },
        256);
    op.n_np = l_np[nleaf];
    op.e_near = (double)op.n_np;
    mark("leaf panels");

    // projection tiles over the V panels; t index of every V column
    // t position of every V column (32-bit: read once per column per matvec)
    // every entry is written below: the low-rank blocks' columns cover
    // [0, Tb) of each V panel, the padding columns get 0
    Table<int32_t> tidx(nvcol > 0 ? nvcol : 1);
    Table<FillDest> dest(nb > 0 ? nb : 1);
    if (nvcol == 0) tidx[0] = 0;
    if (nb == 0) std::memset(dest.data(), 0, sizeof(FillDest));
    parallel_for(
        nn,
        [&](int64_t g0, int64_t g1) {
            for (int64_t g = g0; g < g1; g++)
                for (int64_t k = Tb[g]; k < Tpad[g]; k++) tidx[vcol_base[g] + k] = 0;
        },
        4096);
    std::atomic<int> t_over{0};
    parallel_for(nb, [&](int64_t b0, int64_t b1) {
        for (int64_t b = b0; b < b1; b++) {
            const int64_t a = op.b_row[b], c = op.b_col[b];
            FillDest fd;
            fd.up_base = up_base[a];
            fd.R = (int32_t)R[a];
            fd.coff = coff[b];
            fd.vp_base = vp_base[c];
            fd.T = (int32_t)Tpad[c];
            fd.voff = voff[b];
            fd.sk_row = skel[b] ? skr_off[b] : -1;
            fd.sk_col = skel[b] ? op.n_sk + skc_off[b] : -1;
            dest[b] = fd;
            if (!(op.flags[b] & 1) && !skel[b]) {
                if (t_base[a] + tcoff[b] + op.rank[b] > INT32_MAX) {
                    t_over = 1;
                    continue;
                }
                for (int k = 0; k < op.rank[b]; k++)
                    tidx[vcol_base[c] + voff[b] + k] = (int32_t)(t_base[a] + tcoff[b] + k);
            }
        }
    });
    if (t_over) throw ArgError{HODLR_E_ARG, "more than 2^31 low-rank coefficients per right-hand side"};
    std::vector<DenseBlock> dense;
    op.t_lowrank = 0;
    for (int64_t b = 0; b < nb; b++) {
        if (op.flags[b] & 1) {
            const int64_t a = op.b_row[b];
            DenseBlock db;
            db.row0 = op.node_start[a];
            db.col0 = op.node_start[op.b_col[b]];
            db.up_base = up_base[a];
            db.m = op.b_m[b];
            db.n = op.b_n[b];
            db.R = (int32_t)R[a];
            db.coff = coff[b];
            db.prob = op.b_prob[b];
            db.pad = 0;
            dense.push_back(db);
        } else if (!skel[b]) {
            op.t_lowrank += op.rank[b];
        }
    }
    mark("block dests");
    std::vector<ProjItem> proj;
    std::vector<ProjReduce> red;
    int64_t slots = 0;
    for (int64_t g = 0; g < nn; g++) {
        if (!Tb[g]) continue;
        const int64_t nr = count[g];
        for (int64_t c0 = 0; c0 < Tb[g]; c0 += kProjTile) {
            const int w = (int)std::min<int64_t>(kProjTile, Tb[g] - c0);
            const int64_t nch = (nr + kProjChunk - 1) / kProjChunk;
            if (nch > 1) {
                ProjReduce pr;
                pr.slot0 = slots;
                pr.nchunks = (int32_t)nch;
                pr.w = w;
                pr.tcol = (int32_t)(vcol_base[g] + c0);
                pr.pad = 0;
                red.push_back(pr);
            }
            for (int64_t ch = 0; ch < nch; ch++) {
                const int64_t j0 = ch * kProjChunk, j1 = std::min(nr, j0 + kProjChunk);
                const int64_t tw = std::min<int64_t>(kProjTile, Tpad[g] - c0);  // tile width in memory
                ProjItem it;
                it.vp = vp_base[g] + c0 * nr + j0 * tw;
                it.x0 = op.node_start[g] + j0;
                it.ld = (int32_t)tw;
                it.rows = (int32_t)(j1 - j0);
                it.w = w;
                it.tcol = (int32_t)(vcol_base[g] + c0);
                it.out = nch > 1 ? slots + ch : -1;
                proj.push_back(it);
            }
            if (nch > 1) slots += nch;
        }
    }
    // largest items first, by power-of-two size class: the hardware's
    // round-robin over XCDs then balances bytes per XCD, and inside a class
    // the items keep arena order, so concurrently running workgroups read
    // nearby memory.  per_wg items share a workgroup (item p runs in
    // workgroup p / per_wg).
    std::vector<ProjItem> projw;
    int64_t projw_big = 0;
    auto order_items = [](std::vector<ProjItem> &proj, int per_wg) {
        auto size_class = [](const ProjItem &a) {
            int64_t v = (int64_t)a.rows * a.ld;
            int l = 0;
            while (v > 1) {
                v >>= 1;
                l++;
            }
            return l;
        };
        std::stable_sort(proj.begin(), proj.end(),
                         [&](const ProjItem &a, const ProjItem &b) { return size_class(a) > size_class(b); });
        // Workgroup i runs on XCD i % 8.  Inside each class, deal the
        // items so that every XCD gets one contiguous run of them: nearby
        // column nodes then share an XCD's L2, where their scattered
        // writes to the same t lines (row nodes near them) combine.
        constexpr int kXcd = 8;
        std::vector<ProjItem> dealt(proj.size());
        size_t c0 = 0;
        while (c0 < proj.size()) {
            size_t c1 = c0;
            while (c1 < proj.size() && size_class(proj[c1]) == size_class(proj[c0])) c1++;
            size_t next = c0;  // next item of the class in arena order
            for (int x = 0; x < kXcd; x++)
                for (size_t p = c0; p < c1; p++)
                    if ((int)((p / per_wg) % kXcd) == x) dealt[p] = proj[next++];
            c0 = c1;
        }
        proj.swap(dealt);
    };
    {
        std::vector<ProjItem> big, small;
        for (const ProjItem &it : proj) (it.rows > kProjWaveRows ? big : small).push_back(it);
        order_items(big, 1);
        order_items(small, 4);
        projw_big = (int64_t)big.size();
        projw = std::move(big);
        projw.insert(projw.end(), small.begin(), small.end());
    }
    order_items(proj, 1);
    if (verbose()) {  // projection work by tile height and width
        const int rcut[4] = {64, 256, 2048, 1 << 30};
        for (int full = 1; full >= 0; full--)
            for (int rc = 0; rc < 4; rc++) {
                int64_t cnt = 0;
                double el = 0, lanes = 0;
                for (const ProjItem &it : proj) {
                    if ((it.ld == kProjTile) != (full == 1) || it.rows > rcut[rc] || (rc && it.rows <= rcut[rc - 1]))
                        continue;
                    cnt++;
                    el += (double)it.rows * it.ld;
                    lanes += (double)it.rows * kProjTile / 4;
                }
                if (cnt)
                    std::fprintf(stderr, "[hodlr] proj items %s rows<=%d: %lld, %.3g elements, lane use %.2f\n",
                                 full ? "full-width" : "narrow", rcut[rc], (long long)cnt, el, el / 4 / lanes);
            }
    }
    if (nvcol > INT32_MAX) throw ArgError{HODLR_E_ARG, "too many V columns for 32-bit tile offsets"};
    op.n_slots = slots;
    mark("proj items");

    // skeleton work: projection items per column node (kSkelTile pivot rows
    // x up to kSkelChunk sources), expand items per row node (kSkelTile
    // rows), by level
    std::vector<SkelProj> skproj;
    std::vector<SkelReduce> skred;
    std::vector<SkelExpand> skex;
    op.n_skslots = 0;
    op.skproj_l.clear();
    op.skex_l.clear();
    if (op.n_skel) {
        std::vector<int32_t> nprob(nn, 0), nlev(nn, 0);
        int nlevels = 0;
        for (int p = 0; p < P; p++) {
            const hodlr_tree &t = *op.trees[p];
            nlevels = std::max(nlevels, t.nlevels);
            for (int l = 0; l < t.nlevels; l++)
                for (int64_t j = 0; j < t.level_size(l); j++) {
                    nprob[node_base[p] + t.node(l, j)] = p;
                    nlev[node_base[p] + t.node(l, j)] = l;
                }
        }
        std::vector<std::vector<SkelExpand>> bylev(nlevels);
        for (int64_t g = 0; g < nn; g++) {
            if (Sc[g]) {
                const int64_t ns = count[g];
                const int64_t nch = (ns + kSkelChunk - 1) / kSkelChunk;
                for (int64_t q0 = 0; q0 < Sc[g]; q0 += kSkelTile) {
                    const int nt = (int)std::min<int64_t>(kSkelTile, Sc[g] - q0);
                    if (nch > 1) skred.push_back(SkelReduce{sc_base[g] + q0, op.n_skslots, nt, (int32_t)nch});
                    for (int64_t ch = 0; ch < nch; ch++) {
                        SkelProj it;
                        it.t0 = sc_base[g] + q0;
                        it.src0 = op.node_start[g] + ch * kSkelChunk;
                        it.out = nch > 1 ? op.n_skslots + ch * kSkelTile : -1;
                        it.ntgt = nt;
                        it.nsrc = (int32_t)std::min<int64_t>(kSkelChunk, ns - ch * kSkelChunk);
                        it.prob = nprob[g];
                        it.pad = 0;
                        skproj.push_back(it);
                    }
                    if (nch > 1) op.n_skslots += nch * kSkelTile;
                }
            }
            if (Sa[g]) {
                if (Sa[g] > INT32_MAX) throw ArgError{HODLR_E_ARG, "too many skeleton columns in one node"};
                for (int64_t r0 = 0; r0 < count[g]; r0 += kSkelTile)
                    bylev[nlev[g]].push_back(SkelExpand{op.node_start[g] + r0, sa_base[g],
                                                        (int32_t)std::min<int64_t>(kSkelTile, count[g] - r0),
                                                        (int32_t)Sa[g], nprob[g], 0});
            }
        }
        if (verbose()) {  // lanes doing counted work: one target per lane up to 64, else two
            auto slots = [](int64_t nt) { return nt > kWave ? 2 * kWave : kWave; };
            double up = 0, sp = 0, ue = 0, se = 0;
            size_t nex = 0;
            for (const SkelProj &it : skproj) {
                up += (double)it.ntgt * it.nsrc;
                sp += (double)slots(it.ntgt) * it.nsrc;
            }
            for (const auto &lv : bylev)
                for (const SkelExpand &it : lv) {
                    nex++;
                    ue += (double)it.h * it.nc;
                    se += (double)slots(it.h) * it.nc;
                }
            std::fprintf(stderr, "[hodlr] skeleton items: proj %zu, lane use %.3f; expand %zu, lane use %.3f\n",
                         skproj.size(), sp > 0 ? up / sp : 0.0, nex, se > 0 ? ue / se : 0.0);
        }
        // one launch per (kernel family, NDIM) group -- per tree level as well
        // for expand items -- with the longest items first
        auto key = [&](int p) { return 16 * op.kds[p].family + op.trees[p]->ndim; };
        std::stable_sort(skproj.begin(), skproj.end(), [&](const SkelProj &a, const SkelProj &b) {
            const int ka = key(a.prob), kb = key(b.prob);
            if (ka != kb) return ka < kb;
            return (int64_t)a.nsrc * a.ntgt > (int64_t)b.nsrc * b.ntgt;
        });
        for (size_t i = 0; i < skproj.size();) {
            size_t j = i;
            while (j < skproj.size() && key(skproj[j].prob) == key(skproj[i].prob)) j++;
            op.skproj_l.push_back(SkelLaunch{(int64_t)i, (int64_t)j, key(skproj[i].prob), 0});
            i = j;
        }
        for (int l = 0; l < nlevels; l++) {
            auto &lv = bylev[l];
            std::stable_sort(lv.begin(), lv.end(), [&](const SkelExpand &a, const SkelExpand &b) {
                const int ka = key(a.prob), kb = key(b.prob);
                if (ka != kb) return ka < kb;
                return (int64_t)a.nc * a.h > (int64_t)b.nc * b.h;
            });
            for (size_t i = 0; i < lv.size();) {
                size_t j = i;
                while (j < lv.size() && key(lv[j].prob) == key(lv[i].prob)) j++;
                op.skex_l.push_back(SkelLaunch{(int64_t)(skex.size() + i), (int64_t)(skex.size() + j),
                                               key(lv[i].prob), l});
                i = j;
            }
            skex.insert(skex.end(), lv.begin(), lv.end());
        }
    }

    mark("layout");
    if (op.admit) {
        double bytes = 8.0 * (double)(op.n_np + op.n_up + op.n_vp);
        if (op.prec_flags & HODLR_PREC_F32) bytes *= 1.5;
        bytes += 8.0 * (double)op.n_g + 8.0 * (double)op.n_sk + (double)sizeof(QuadDesc) * (double)quads.size() +
                 (double)sizeof(ProjItem) * (double)projw.size();
        op.admit(bytes);
        mark("admission");
    }
    // ---- 4. allocate and fill ----
    const double t_fill0 = now_s();
    if (stage.buf) {  // the staged factors must leave room for the arenas
        HIP_CHECK(hipStreamSynchronize(s));
        size_t fb = device_free_bytes();
        if (tl_slab) {  // arenas that do not fit the slab come from the driver
            size_t fr = 0, tot = 0;
            HIP_CHECK(hipMemGetInfo(&fr, &tot));
            fb = std::max(fb, fr);
        }
        double need = 8.0 * (double)(op.n_np + op.n_up + op.n_vp + 2 * kArenaPad);
        if (op.prec_flags & HODLR_PREC_F32) need *= 1.5;
        need += 8.0 * (double)op.n_g + 8.0 * (double)op.n_sk + (double)(1 << 30);
        if ((double)fb < need) {
            stage_buf.release();
            stage.buf = nullptr;
            soff.clear();
            if (verbose()) std::fprintf(stderr, "[hodlr] staged factors released for the arenas\n");
        }
    }
    if (verbose())
        std::fprintf(stderr, "[hodlr] layout: near %.3g, U %.3g, V %.3g entries (%.2f GB FP64)\n", (double)op.n_np,
                     (double)op.n_up, (double)op.n_vp, 8e-9 * (double)(op.n_np + op.n_up + op.n_vp));
    // near-field and U arenas carry zeroed tail padding for the 16-byte
    // panel loads that run past a tile's last row
    op.NP64.alloc(sizeof(double) * (op.n_np + kArenaPad));
    op.UP64.alloc(sizeof(double) * (op.n_up + kArenaPad));
    HIP_CHECK(hipMemsetAsync(op.NP64.as<double>() + op.n_np, 0, sizeof(double) * kArenaPad, s));
    HIP_CHECK(hipMemsetAsync(op.UP64.as<double>() + op.n_up, 0, sizeof(double) * kArenaPad, s));
    op.VP64.alloc(sizeof(double) * op.n_vp);
    HIP_CHECK(hipMemsetAsync(op.VP64.p, 0, sizeof(double) * op.n_vp, s));  // padding columns
    op.SK.alloc(sizeof(int32_t) * (size_t)(2 * op.n_sk));
    HIP_CHECK(hipMemsetAsync(op.SK.p, 0xff, sizeof(int32_t) * (size_t)(2 * op.n_sk), s));  // -1: unwritten
    DevBuf d_near;
    upload(d_near, near, s, true);
    if (op.fill_from) {
        op.fill_from(op);
    } else {
        DevBuf d_colsrc, d_dense, d_dest;
        upload(d_colsrc, colsrc, s, true);
        upload(d_dense, dense, s, true);
        upload(d_dest, dest, s, true);
        HIP_CHECK(op.entries->near_fill(d_near.as<NearLeaf>(), (int64_t)near.size(), d_colsrc.as<int64_t>(),
                                   op.kdev.as<KernelDev>(), op.NP64.as<double>(), s));
        HIP_CHECK(op.entries->dense_fill(d_dense.as<DenseBlock>(), (int64_t)dense.size(), op.kdev.as<KernelDev>(),
                                    op.leaf_start.as<int64_t>(), op.leaf_cnt.as<int32_t>(), op.UP64.as<double>(), s));
        HIP_CHECK(hipStreamSynchronize(s));
        std::vector<int64_t> lowrank, staged, staged_piv;
        for (int64_t b = 0; b < nb; b++)
            if (!(op.flags[b] & 1)) {
                if (stage.buf && soff[b] >= 0) (skel[b] ? staged_piv : staged).push_back(b);
                else lowrank.push_back(b);
            }
        if (!staged_piv.empty()) {  // skeleton blocks' pivots, kept by the rank pass
            std::vector<PivotCopy> items(staged_piv.size());
            parallel_for((int64_t)items.size(), [&](int64_t a0, int64_t a1) {
                for (int64_t a = a0; a < a1; a++) {
                    const int64_t b = staged_piv[(size_t)a];
                    items[(size_t)a] = PivotCopy{soff[b], dest[b].sk_row, dest[b].sk_col, op.rank[b], 0};
                }
            }, 1 << 14);
            DevBuf d_items;
            upload(d_items, items, s, true);
            HIP_CHECK(launch_stage_pivots(d_items.as<PivotCopy>(), (int64_t)items.size(), stage.buf,
                                          op.SK.as<int32_t>(), s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        if (!staged.empty()) {  // copy staged factors, in pieces of at most 64 Ki values
            constexpr int64_t kPiece = 1 << 16;
            std::vector<int64_t> first(staged.size() + 1, 0);
            for (size_t q = 0; q < staged.size(); q++) {
                const int64_t b = staged[q];
                const int64_t len = (int64_t)op.rank[b] * (op.b_m[b] + op.b_n[b]);
                first[q + 1] = first[q] + std::max<int64_t>(1, (len + kPiece - 1) / kPiece);
            }
            constexpr int64_t kBatch = (int64_t)1 << 20;  // items per launch
            std::vector<StageCopy> items;
            DevBuf d_items;
            for (size_t q0 = 0; q0 < staged.size();) {
                size_t q1 = q0;
                while (q1 < staged.size() && first[q1 + 1] - first[q0] <= kBatch) q1++;
                if (q1 == q0) q1 = q0 + 1;
                items.resize((size_t)(first[q1] - first[q0]));
                parallel_for(
                    (int64_t)(q1 - q0),
                    [&](int64_t a0, int64_t a1) {
                        for (int64_t a = a0; a < a1; a++) {
                            const size_t q = q0 + (size_t)a;
                            const int64_t b = staged[q];
                            const int p = op.b_prob[b];
                            const FillDest &D = dest[b];
                            StageCopy c;
                            c.src = soff[b];
                            c.row0 = op.node_start[op.b_row[b]];
                            c.col0 = op.node_start[op.b_col[b]];
                            c.row0l = c.row0 - op.pos_base[p];
                            c.col0l = c.col0 - op.pos_base[p];
                            c.ord = op.ord.as<int32_t>() + op.ord_off[p] + (int64_t)op.b_level[b] * op.trees[p]->n;
                            c.up_base = D.up_base;
                            c.vp_base = D.vp_base;
                            c.m = op.b_m[b];
                            c.n = op.b_n[b];
                            c.r = op.rank[b];
                            c.R = D.R;
                            c.coff = D.coff;
                            c.T = D.T;
                            c.voff = D.voff;
                            c.pad = 0;
                            const int64_t len = (int64_t)c.r * (c.m + c.n);
                            for (int64_t k = first[q]; k < first[q + 1]; k++) {
                                const int64_t e0 = (k - first[q]) * kPiece;
                                c.e0 = (int32_t)e0;
                                c.e1 = (int32_t)std::min(len, e0 + kPiece);
                                items[(size_t)(k - first[q0])] = c;
                            }
                        }
                    },
                    1 << 14);
                upload(d_items, items, s, true);
                HIP_CHECK(launch_stage_copy(d_items.as<StageCopy>(), (int64_t)items.size(), stage.buf,
                                            op.leaf_start.as<int64_t>(), op.leaf_cnt.as<int32_t>(),
                                            op.UP64.as<double>(), op.VP64.as<double>(), s));
                HIP_CHECK(hipStreamSynchronize(s));
                q0 = q1;
            }
            if (verbose())
                std::fprintf(stderr, "[hodlr] staged factors of %zu blocks and pivots of %zu copied, %zu blocks "
                             "recomputed\n", staged.size(), staged_piv.size(), lowrank.size());
        }
        stage_buf.release();
        // the fill pass repeats the rank pass with the same caps
        HIP_CHECK(hipMemsetAsync(d_rank.p, 0xff, sizeof(int32_t) * (nb + 1), s));
        run_aca_pass(op, lowrank, cap, 1, pool, budget, d_rank, d_flags, &d_dest, s);
        std::vector<int32_t> rk(nb > 0 ? nb : 1);
        if (nb) {
            HIP_CHECK(hipMemcpyAsync(rk.data(), d_rank.p, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        for (int64_t b : lowrank)
            if (rk[b] != op.rank[b])
                throw ArgError{HODLR_E_STATE, "ACA fill pass disagrees with the rank pass at block " +
                                                  std::to_string(b)};
    }
    if (op.n_skel) {
        // pivot positions (written by the fill pass or read from a file) must
        // lie in their block's problem before any kernel dereferences them
        std::vector<int32_t> skh((size_t)(2 * op.n_sk));
        if (op.n_sk) {
            HIP_CHECK(hipMemcpyAsync(skh.data(), op.SK.p, sizeof(int32_t) * skh.size(), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        std::vector<int64_t> skb;
        for (int64_t b = 0; b < nb; b++)
            if (skel[b] && op.rank[b] > 0) skb.push_back(b);
        for (int64_t b : skb) {
            const int p = op.b_prob[b];
            const int32_t *ri = skh.data() + skr_off[b];
            const int32_t *ci = skh.data() + op.n_sk + skc_off[b];
            for (int i = 0; i < op.rank[b]; i++)
                if (ri[i] < op.pos_base[p] || ri[i] >= op.pos_base[p + 1] || ci[i] < op.pos_base[p] ||
                    ci[i] >= op.pos_base[p + 1])
                    throw ArgError{HODLR_E_STATE, "skeleton pivot position out of range at block " + std::to_string(b)};
        }
        if (!op.fill_from) {  // the cores C = A(pivot rows, pivot columns) and G = C^+
            std::vector<double> Gh((size_t)op.n_g);
            parallel_for(
                (int64_t)skb.size(),
                [&](int64_t q0, int64_t q1) {
                    std::vector<double> Cm;
                    for (int64_t q = q0; q < q1; q++) {
                        const int64_t b = skb[q];
                        const int r = op.rank[b], p = op.b_prob[b];
                        const hodlr_tree &t = *op.trees[p];
                        const int64_t pb = op.pos_base[p];
                        const int32_t *ri = skh.data() + skr_off[b];
                        const int32_t *ci = skh.data() + op.n_sk + skc_off[b];
                        const double *Xp = X.data() + xoff[p];
                        const double *Wp = woff[p] >= 0 ? W.data() + woff[p] : nullptr;
                        Cm.resize((size_t)r * r);
                        for (int k = 0; k < r; k++)
                            for (int i = 0; i < r; i++)
                                Cm[(size_t)k * r + i] = host_entry(op.kds[p], t.ndim, Xp, Wp, ri[i] - pb, ci[k] - pb);
                        core_pinv(r, Cm, Gh.data() + g_off[b]);
                    }
                },
                16);
            upload(op.GA, Gh, s);
        }
    }
    pool.release();
    if (op.prec_flags & HODLR_PREC_F32) {
        op.NP32.alloc(sizeof(float) * (op.n_np + kArenaPad));
        op.UP32.alloc(sizeof(float) * (op.n_up + kArenaPad));
        HIP_CHECK(hipMemsetAsync(op.NP32.as<float>() + op.n_np, 0, sizeof(float) * kArenaPad, s));
        HIP_CHECK(hipMemsetAsync(op.UP32.as<float>() + op.n_up, 0, sizeof(float) * kArenaPad, s));
        op.VP32.alloc(sizeof(float) * op.n_vp);
        HIP_CHECK(launch_convert(op.NP64.as<double>(), op.NP32.as<float>(), op.n_np, s));
        HIP_CHECK(launch_convert(op.UP64.as<double>(), op.UP32.as<float>(), op.n_up, s));
        HIP_CHECK(launch_convert(op.VP64.as<double>(), op.VP32.as<float>(), op.n_vp, s));
        // the diagonal (entry(i, i) = diag for every i) is applied separately
        // in the working precision of the caller, so FP32 rounding of the
        // identity part of a second-kind operator does not limit refinement
        HIP_CHECK(launch_zero_self_diag(d_near.as<NearLeaf>(), (int64_t)near.size(), op.NP32.as<float>(), s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    upload(op.proj, proj, s);
    upload(op.projw, projw, s);
    op.n_projw_big = projw_big;
    upload(op.red, red, s);
    upload(op.panel, panel, s);
    upload(op.segs, segs, s);
    upload(op.quads, quads, s);
    upload(op.tidx, tidx, s);
    upload(op.skproj, skproj, s);
    upload(op.skred, skred, s);
    upload(op.skcore, skcore, s);
    upload(op.skex, skex, s);
    op.n_skproj = (int64_t)skproj.size();
    op.n_skred = (int64_t)skred.size();
    op.n_skcore = (int64_t)skcore.size();
    op.n_proj = (int64_t)proj.size();
    op.n_red = (int64_t)red.size();
    op.n_panel = (int64_t)panel.size();
    op.n_seg = (int64_t)segs.size();
    op.n_quad = (int64_t)quads.size();
    {
        // 1-4 RHS default panel form, from the load rounds of one wave the
        // two walks need (kU = 8 column steps per round, 4 waves, G = 64 /
        // ceil(h / R) columns per step, R = 16 / sizeof(T)): one segment at a
        // time, a segment takes ceil(len / (4 kU G)) rounds; over the
        // flattened walk a tile takes ceil(sum ceil(len / G) / (4 kU)).
        // Short segments under short tiles leave the per-segment walk's
        // rounds mostly empty; where they do not, the flattened walk's
        // per-step segment lookup costs more than it saves.
        for (int pi = 0; pi < 2; pi++) {
            const int R = pi == 0 ? 2 : 4;  // FP64, FP32
            double seg_rounds = 0.0, flat_rounds = 0.0;
            for (size_t pq = 0; pq < panel.size(); pq++) {
                const PanelItem &it = panel[pq];
                const int hl = (it.h + R - 1) / R, G = kWave / hl;
                int64_t steps = 0;
                for (int64_t q = it.seg0; q < it.seg0 + it.nseg; q++) {
                    seg_rounds += (double)((segs[q].len + 32 * G - 1) / (32 * G));
                    steps += (segs[q].len + G - 1) / G;
                }
                flat_rounds += (double)((steps + 31) / 32);
            }
            const double gain = flat_rounds > 0.0 ? seg_rounds / flat_rounds : 0.0;
            if (gain >= kFlatGain)
                for (int lg = 0; lg < 3; lg++)
                    if (op.mv_variant[pi][lg] == kMatvecPanelVec) op.mv_variant[pi][lg] |= kMatvecPanelFlat;
            if (verbose())
                std::fprintf(stderr, "[hodlr] panel %s: %zu segments, rounds per segment %.0f, flattened %.0f (%.2fx)\n",
                             pi == 0 ? "f64" : "f32", (size_t)segs.size(), seg_rounds, flat_rounds, gain);
        }
    }
    {  // per-position diagonal (a batch may mix diagonals) for the FP32 path
        std::vector<double> dg(N > 0 ? N : 1, 0.0);
        for (int p = 0; p < P; p++)
            for (int64_t q = op.pos_base[p]; q < op.pos_base[p + 1]; q++) dg[q] = op.kds[p].diag;
        upload(op.diag, dg, s);
        op.uniform_diag = true;
        for (int p = 1; p < P; p++) op.uniform_diag = op.uniform_diag && op.kds[p].diag == op.kds[0].diag;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    mark("fill");
    const double t_end = now_s();

    // ---- statistics (HDD_node.hpp:183-220, HDD_matrix.hpp:69-85) ----
    hodlr_op_stats &st = op.stats;
    std::memset(&st, 0, sizeof(st));
    st.e_near = op.e_near;
    st.n_flop_ref = op.e_near;
    for (int64_t b = 0; b < nb; b++) {
        const bool dn = op.flags[b] & 1;
        const int64_t mn = std::min(op.b_m[b], op.b_n[b]);
        const int64_t rr = dn ? mn : op.rank[b];
        st.n_flop_ref += (double)rr * (double)(op.b_m[b] + op.b_n[b]);
        st.max_rank = std::max<int64_t>(st.max_rank, rr);
        if (!skel[b])
            st.e_far += dn ? (double)op.b_m[b] * op.b_n[b] : (double)op.rank[b] * (op.b_m[b] + op.b_n[b]);
        st.n_dense += dn;
    }
    st.n_blocks = nb;
    op.e_far = st.e_far;
    st.n_skeleton = op.n_skel;
    st.e_skel = op.e_skel;
    st.aca_entries = aca_entries;
    // cores (FP64, shared by both precisions) and pivot positions
    st.bytes_core = 8.0 * (double)op.n_g + 8.0 * (double)op.n_sk;
    st.bytes_f64 = 8.0 * (double)(op.n_np + op.n_up + op.n_vp) + st.bytes_core;
    st.bytes_f32 = (op.prec_flags & HODLR_PREC_F32) ? 4.0 * (double)(op.n_np + op.n_up + op.n_vp) : 0.0;
    st.matvec_bytes_f64 = 8.0 * (st.e_near + st.e_far) + st.bytes_core + 2.0 * 8.0 * (double)N;
    st.matvec_bytes_f32 = 4.0 * (st.e_near + st.e_far) + st.bytes_core + 2.0 * 4.0 * (double)N;
    st.rank_seconds = t_rank1 - t_rank0;
    st.fill_seconds = t_end - t_fill0;
    st.assemble_seconds = t_end - t_start;
}

}  // namespace hodlr

void hodlr_op::ensure_work(int prec, int nrhs) {
    using namespace hodlr;
    if (n_skcore && cap_rhs_sk < nrhs) {
        skt1.alloc(sizeof(double) * (size_t)n_sk * nrhs);
        skt2.alloc(sizeof(double) * (size_t)n_sk * nrhs);
        skpart.alloc(sizeof(double) * (size_t)(n_skslots > 0 ? n_skslots : 1) * nrhs);
        skys.alloc(sizeof(double) * (size_t)N * nrhs);
        cap_rhs_sk = nrhs;
        work_gen++;
    }
    if (prec == HODLR_PREC_F64) {
        if (cap_rhs64 >= nrhs) return;
        t64.alloc(sizeof(double) * (size_t)(n_t > 0 ? n_t : 1) * nrhs);
        part64.alloc(sizeof(double) * (size_t)(n_slots > 0 ? n_slots : 1) * kProjTile * nrhs);
        cap_rhs64 = nrhs;
    } else {
        if (cap_rhs32 >= nrhs) return;
        t32.alloc(sizeof(float) * (size_t)(n_t > 0 ? n_t : 1) * nrhs);
        part32.alloc(sizeof(float) * (size_t)(n_slots > 0 ? n_slots : 1) * kProjTile * nrhs);
        cap_rhs32 = nrhs;
    }
    work_gen++;
}

hodlr_op::~hodlr_op() {
    for (auto &e : ev_pool) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
}

void hodlr_op::profile_collect() {
    using namespace hodlr;
    if (ev_pending.empty()) return;
    HIP_CHECK(hipEventSynchronize(ev_pool[ev_pending.back().second + 3]));
    for (auto &pe : ev_pending) {
        float a = 0, b = 0, c = 0;
        HIP_CHECK(hipEventElapsedTime(&a, ev_pool[pe.second], ev_pool[pe.second + 1]));
        HIP_CHECK(hipEventElapsedTime(&b, ev_pool[pe.second + 1], ev_pool[pe.second + 2]));
        if (n_skcore) HIP_CHECK(hipEventElapsedTime(&c, ev_pool[pe.second + 4], ev_pool[pe.second + 5]));
        const int p = pe.first == HODLR_PREC_F64 ? 0 : 1;
        prof_ms[p][0] += a;
        prof_ms[p][1] += b;
        prof_ms[p][2] += c;
        prof_count[p]++;
    }
    ev_pending.clear();
}

void hodlr_op::matvec_dev(int prec, int nrhs, const void *x, void *y, hipStream_t s, hipEvent_t *ev,
                          bool add_diag) {
    using namespace hodlr;
    ensure_work(prec, nrhs);
    if (profiling && !ev) {
        const size_t need = kMatvecEvents * (ev_pending.size() + 1);
        while (ev_pool.size() < need) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        const size_t first = kMatvecEvents * ev_pending.size();
        ev_pending.push_back({prec, first});
        ev = &ev_pool[first];
    }
    MatvecWork w;
    w.t = prec == HODLR_PREC_F64 ? t64.p : t32.p;
    w.part = prec == HODLR_PREC_F64 ? part64.p : part32.p;
    w.skt1 = skt1.as<double>();
    w.skt2 = skt2.as<double>();
    w.skpart = skpart.as<double>();
    w.skys = skys.as<double>();
    const int mode = n_skcore ? skel_schedule(prec, nrhs) : HODLR_SKEL_SERIAL;
    if (mode != HODLR_SKEL_SERIAL) {
        // not while s is being captured into a graph: the second stream
        // would have to join the capture
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_CHECK(hipStreamIsCapturing(s, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            if (!side) {
                HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
                HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
                HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
            }
            w.skel = mode;
        }
    }
    matvec_dev_on(prec, nrhs, x, y, s, w, ev, add_diag);
}

// The schedule of the skeleton phases for one (precision, RHS count).
// Without a caller's or the tuner's choice a cost model decides.  The stored
// passes move their algorithmic bytes at about 5.7 TB/s; the skeleton
// phases evaluate about 1.7e12 entries/s with every CU to themselves (1 RHS;
// each further RHS adds one FMA to ~12 instructions per entry).  Uncapped
// beside the stored passes they add about 0.65 of their own time to the
// product; capped they take about 2.2x their own time but leave the stored
// passes close to full speed.  So capping pays while 2.2 t_skel < t_stored +
// 0.65 t_skel.  Measured on the N = 10^6 grid operator (FP32: stored passes
// 2.88 ms, skeleton phases 1.51 ms alone): one stream 4.40 ms, uncapped
// 3.87 ms, capped 3.35 ms; FP64 (5.72 and 1.59 ms): 7.33, 6.48, 6.12 ms.
// An operator of mostly skeleton blocks keeps the whole machine for them.
int hodlr_op::skel_schedule(int prec, int nrhs) const {
    if (skel_mode != HODLR_SKEL_AUTO) return skel_mode;
    int lg = 0;
    while ((1 << lg) < nrhs && lg < 5) lg++;
    const int pi = prec == HODLR_PREC_F64 ? 0 : 1;
    if (skel_pick[pi][lg]) return skel_pick[pi][lg];
    const double t_stored = (pi == 0 ? stats.matvec_bytes_f64 : stats.matvec_bytes_f32) / 5.7e12;
    const double t_skel = e_skel * (11.0 + nrhs) / 12.0 / 1.7e12;
    return 1.55 * t_skel <= t_stored ? HODLR_SKEL_CAPPED : HODLR_SKEL_OVERLAP;
}

void hodlr_op::matvec_work_sizes(int prec, int nrhs, size_t sz[6]) const {
    using namespace hodlr;
    const size_t es = prec == HODLR_PREC_F64 ? sizeof(double) : sizeof(float);
    sz[0] = es * (size_t)(n_t > 0 ? n_t : 1) * nrhs;
    sz[1] = es * (size_t)(n_slots > 0 ? n_slots : 1) * kProjTile * nrhs;
    sz[2] = sz[3] = sz[4] = sz[5] = 0;
    if (n_skcore) {
        sz[2] = sz[3] = sizeof(double) * (size_t)n_sk * nrhs;
        sz[4] = sizeof(double) * (size_t)(n_skslots > 0 ? n_skslots : 1) * nrhs;
        sz[5] = sizeof(double) * (size_t)N * nrhs;
    }
}

void hodlr_op::matvec_dev_on(int prec, int nrhs, const void *x, void *y, hipStream_t s, const hodlr::MatvecWork &w,
                             hipEvent_t *ev, bool add_diag) {
    using namespace hodlr;
    const bool f64 = prec == HODLR_PREC_F64;
    if (!f64 && !(prec_flags & HODLR_PREC_F32)) throw ArgError{HODLR_E_STATE, "operator has no FP32 copy"};
    // skeleton phases: they read x and write only their own buffers (ys
    // last), so on the second stream they run beside the stored blocks'
    // passes; y += ys follows both.  Issued before the stored passes, so a
    // stream sharing a hardware queue with the second one never waits
    // ahead of the work it waits for.
    const bool skel = n_skcore > 0;
    const bool over = skel && w.skel != HODLR_SKEL_SERIAL;
    if (over && !side) throw ArgError{HODLR_E_STATE, "skeleton schedule needs the operator's second stream"};
    auto skeleton = [&](hipStream_t ss) {
        const SkeletonPlan p = skel_plan(w.skel == HODLR_SKEL_CAPPED ? kSkelCapLds : 0);
        if (ev) HIP_CHECK(hipEventRecord(ev[4], ss));
        if (f64)
            HIP_CHECK(launch_skeleton<double>(nrhs, p, static_cast<const double *>(x), w.skt1, w.skpart, w.skt2,
                                              w.skys, N, ss));
        else
            HIP_CHECK(launch_skeleton<float>(nrhs, p, static_cast<const float *>(x), w.skt1, w.skpart, w.skt2,
                                             w.skys, N, ss));
        if (ev) HIP_CHECK(hipEventRecord(ev[5], ss));
    };
    if (over) {
        HIP_CHECK(hipEventRecord(ev_fork, s));
        HIP_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
        skeleton(side);
        HIP_CHECK(hipEventRecord(ev_join, side));
    }
    if (f64)
        HIP_CHECK(launch_matvec<double>(nrhs, proj.as<ProjItem>(), n_proj, projw.as<ProjItem>(), n_projw_big,
                                        red.as<ProjReduce>(), n_red,
                                        panel.as<PanelItem>(), n_panel,
                                        segs.as<Segment>(), quads.as<QuadDesc>(), VP64.as<double>(), NP64.as<double>(),
                                        UP64.as<double>(), static_cast<const double *>(x), static_cast<double *>(w.t),
                                        static_cast<double *>(w.part), tidx.as<int32_t>(), static_cast<double *>(y),
                                        variant_for(prec, nrhs), ev, s));
    else
        HIP_CHECK(launch_matvec<float>(nrhs, proj.as<ProjItem>(), n_proj, projw.as<ProjItem>(), n_projw_big,
                                       red.as<ProjReduce>(), n_red,
                                       panel.as<PanelItem>(), n_panel,
                                       segs.as<Segment>(), quads.as<QuadDesc>(), VP32.as<float>(), NP32.as<float>(),
                                       UP32.as<float>(),
                                       static_cast<const float *>(x), static_cast<float *>(w.t),
                                       static_cast<float *>(w.part), tidx.as<int32_t>(), static_cast<float *>(y),
                                       variant_for(prec, nrhs), ev, s));
    if (skel) {
        if (over) HIP_CHECK(hipStreamWaitEvent(s, ev_join, 0));
        else skeleton(s);
        if (f64) HIP_CHECK(launch_skeleton_add<double>(static_cast<double *>(y), w.skys, N * nrhs, s));
        else HIP_CHECK(launch_skeleton_add<float>(static_cast<float *>(y), w.skys, N * nrhs, s));
    }
    if (ev) HIP_CHECK(hipEventRecord(ev[3], s));
    if (!f64 && add_diag)
        HIP_CHECK(launch_add_diag_f32(static_cast<float *>(y), static_cast<const float *>(x),
                                      diag.as<double>(), N, nrhs, s));
}

hodlr::SkeletonPlan hodlr_op::skel_plan(int lds_min) const {
    using namespace hodlr;
    SkeletonPlan p;
    p.proj = skproj.as<SkelProj>();
    p.red = skred.as<SkelReduce>();
    p.core = skcore.as<SkelCore>();
    p.ex = skex.as<SkelExpand>();
    p.nred = n_skred;
    p.ncore = n_skcore;
    p.proj_l = skproj_l.data();
    p.nproj_l = (int)skproj_l.size();
    p.ex_l = skex_l.data();
    p.nex_l = (int)skex_l.size();
    p.K = kdev.as<KernelDev>();
    p.SK = SK.as<int32_t>();
    p.nrow_slots = n_sk;
    p.G = GA.as<double>();
    p.lds_min = lds_min;
    return p;
}

int hodlr_op::variant_for(int prec, int nrhs) const {
    int lg = 0;
    while ((1 << lg) < nrhs && lg < 5) lg++;
    return mv_variant[prec == HODLR_PREC_F64 ? 0 : 1][lg];
}

void hodlr_op::tune_matvec(int prec_mask, int reps) {
    using namespace hodlr;
    hipStream_t s = ctx->stream;
    for (int pi = 0; pi < 2; pi++) {
        const int prec = pi == 0 ? HODLR_PREC_F64 : HODLR_PREC_F32;
        if (!(prec_mask & prec) || !(prec_flags & prec)) continue;
        const size_t es = prec == HODLR_PREC_F32 ? sizeof(float) : sizeof(double);
        for (int lg = 0; lg < 6; lg++) {
            const int k = 1 << lg;
            if (prec == HODLR_PREC_F64 && k > 16) continue;
            DevBuf dx, dy;
            dx.alloc(es * N * k);
            dy.alloc(es * N * k);
            HIP_CHECK(hipMemsetAsync(dx.p, 0, es * N * k, s));
            double best[2] = {1e300, 1e300};
            int pick[2] = {0, 0};
            // vector-ALU panel forms only up to 8 RHS: from 16 on their
            // per-lane accumulators spill past one wave per SIMD and they
            // run 4-20x slower than the matrix-core form (never chosen)
            std::vector<int> cands;
            if (k <= 8) cands = {0, kMatvecPanelFlat};
            if (k <= 8) cands.push_back(kMatvecPanelVec);
            if (k <= 8) cands.push_back(kMatvecPanelVec | kMatvecPanelFlat);
            if (k <= 2) cands.push_back(kMatvecProjPack | kMatvecPanelVec);
            if (k >= 4) cands.push_back(kMatvecProjMfma | kMatvecPanelMfma);
            if (k >= 16) cands.push_back(kMatvecPanelMfma);  // VALU projection, MFMA panel
            cands.push_back(kMatvecProjWave | (k >= 4 ? kMatvecPanelMfma | kMatvecPanelWide : kMatvecPanelVec));
            if (k >= 4) cands.push_back(kMatvecProjMfma | kMatvecPanelMfma | kMatvecPanelWide);
            for (int v : cands) {
                mv_variant[pi][lg] = v;
                hipEvent_t ev[kMatvecEvents];
                for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
                matvec_dev(prec, k, dx.p, dy.p, s);  // warm-up
                double ph[2] = {0, 0};
                for (int r = 0; r < reps; r++) {
                    matvec_dev(prec, k, dx.p, dy.p, s, ev);
                    HIP_CHECK(hipEventSynchronize(ev[3]));
                    float a = 0, b = 0;
                    HIP_CHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
                    HIP_CHECK(hipEventElapsedTime(&b, ev[1], ev[2]));
                    ph[0] += a;
                    ph[1] += b;
                }
                for (auto &e : ev) (void)hipEventDestroy(e);
                const int bits[2] = {kMatvecProjMfma | kMatvecProjPack | kMatvecProjWave,
                                     kMatvecPanelMfma | kMatvecPanelFlat | kMatvecPanelVec | kMatvecPanelWide};
                for (int q = 0; q < 2; q++)
                    if (ph[q] < best[q]) {
                        best[q] = ph[q];
                        pick[q] = v & bits[q];
                    }
            }
            mv_variant[pi][lg] = pick[0] | pick[1];
            work_gen++;
            // the skeleton phases' schedule, whole products timed
            int skel_best = 0;
            double skel_ms[4] = {0, 0, 0, 0};
            if (n_skcore && (prec == HODLR_PREC_F32 || k <= 16)) {
                const int keep = skel_mode;
                double best_t = 1e300;
                hipEvent_t ev[kMatvecEvents];
                for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
                for (int m = HODLR_SKEL_SERIAL; m <= HODLR_SKEL_CAPPED; m++) {
                    skel_mode = m;
                    matvec_dev(prec, k, dx.p, dy.p, s);  // warm-up
                    double t = 0;
                    for (int r = 0; r < reps; r++) {
                        matvec_dev(prec, k, dx.p, dy.p, s, ev);
                        HIP_CHECK(hipEventSynchronize(ev[3]));
                        float a = 0;
                        HIP_CHECK(hipEventElapsedTime(&a, ev[0], ev[3]));
                        t += a;
                    }
                    skel_ms[m] = t / reps;
                    if (t < best_t) {
                        best_t = t;
                        skel_best = m;
                    }
                }
                for (auto &e : ev) (void)hipEventDestroy(e);
                skel_mode = keep;
                skel_pick[pi][lg] = skel_best;
            }
            if (verbose())
                std::fprintf(stderr, "[hodlr] tune %s K=%d: proj %s %.3f ms, panel %s %.3f ms\n",
                             pi == 0 ? "f64" : "f32", k,
                             pick[0] & kMatvecProjMfma ? "mfma"
                                 : (pick[0] & kMatvecProjWave ? "wave" : (pick[0] & kMatvecProjPack ? "pack" : "valu")),
                             best[0] / reps,
                             pick[1] & kMatvecPanelMfma ? (pick[1] & kMatvecPanelWide ? "mfma-wide" : "mfma")
                                 : (pick[1] & kMatvecPanelVec ? (pick[1] & kMatvecPanelFlat ? "vec-flat" : "vec")
                                                              : (pick[1] ? "flat" : "seg")),
                             best[1] / reps);
            if (verbose() && skel_best)
                std::fprintf(stderr, "[hodlr] tune %s K=%d: skeleton schedule %d (whole product: one stream %.3f ms, "
                             "beside %.3f ms, capped beside %.3f ms)\n", pi == 0 ? "f64" : "f32", k, skel_best,
                             skel_ms[1], skel_ms[2], skel_ms[3]);
        }
    }
}

namespace hodlr {
void set_assembly_threads(int n) { tl_threads = n; }
}  // namespace hodlr
