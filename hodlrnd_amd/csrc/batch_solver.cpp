// batch_solver.cpp -- independent GMRES solves over a batch operator.
//
// Every (problem, right-hand side) pair of a batch operator is one GMRES
// instance with the semantics of iterSolver::GMRES (GMRES.hpp:160-303):
// residual check `<= tol` before the first cycle, `< tol` after each
// Arnoldi step and restart, Givens rotations (GMRES.hpp:83-129), back-solve
// update (GMRES.hpp:132-146).  The instances advance in lockstep: each
// Arnoldi step is one K-column matvec over the whole batch and one launch
// per CGS2 reduction, with per-instance sums over chunks that never cross a
// problem boundary (so an instance's arithmetic does not depend on the rest
// of the batch).  Converged instances stop changing: their next basis
// vectors are zero and their coefficients are dropped.
//
// The Givens rotations, the residual tests and the back-solves run on the
// device, one thread per instance (blas.hip); the host reads two counts per
// step (instances still running, breakdowns).  The device work of Arnoldi
// step i -- the matvec, both CGS2 passes, the norm of the new vector, the
// rotations of every instance, the next basis vector and the copy of the
// counts to pinned memory -- depends only on i (basis vectors, work buffers
// and the pinned staging are fixed per workspace), so it is captured once
// into a hipGraph per (operator precision, i) and replayed as one launch
// followed by one wait.  A graph is recaptured when the operator's work
// buffers or kernel forms change.  An instance that ends at a step keeps
// its basis vectors until the end of the cycle, where every instance that
// ran in it is updated by one product x += V y.
#include <chrono>
#include <cmath>
#include <cstring>

#include "op.h"
#include "solver.h"

namespace hodlr {

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int64_t kChunkRows = 2048;

}  // namespace

// device work space of a batched solve, cached on the operator
struct BatchWork {
    int K = 0, restart = 0;
    int64_t n = 0;
    int P = 0, nchunk = 0;
    std::vector<DevBuf> V;       // basis vectors, allocated on first use
    std::vector<double *> vptr;
    DevBuf vtab, w, r, d, rr, w64, x32, y32, partial, coef, chunks, pchunk, tab1, tabw, tabr;
    // device GMRES state of every instance: GmresDevState, the Hessenberg
    // matrices ((restart+1) x restart each), cs / sn / sv / y ((restart+1)
    // each), next-vector scales, ||b||^2, ||r||^2, tolerances, budgets and
    // the two step counts
    DevBuf gst, gH, gvec, gscale, gnb2, gr2, gtol, gbud, gcnt;
    double *hbuf = nullptr;      // pinned staging
    size_t hcap = 0;
    // captured Arnoldi steps [operator precision: 0 FP64, 1 FP32][step] and
    // the operator work generation each was captured at
    std::vector<hipGraphExec_t> gexec[2];
    std::vector<uint64_t> ggen[2];

    BatchWork(const hodlr_op &op, int K_, int restart_) : K(K_), restart(restart_) {
        n = op.N;
        P = (int)op.trees.size();
        std::vector<Chunk> ch;
        std::vector<int32_t> pc(P + 1, 0);
        for (int p = 0; p < P; p++) {
            pc[p] = (int32_t)ch.size();
            for (int64_t r0 = op.pos_base[p]; r0 < op.pos_base[p + 1]; r0 += kChunkRows) {
                Chunk c;
                c.row0 = r0;
                c.rows = (int32_t)std::min<int64_t>(kChunkRows, op.pos_base[p + 1] - r0);
                c.prob = p;
                ch.push_back(c);
            }
        }
        pc[P] = (int32_t)ch.size();
        nchunk = (int)ch.size();
        upload(chunks, ch);
        upload(pchunk, pc);
        V.resize(restart + 1);
        vptr.assign(restart + 1, nullptr);
        vtab.alloc(sizeof(double *) * (restart + 1));
        const size_t nk = (size_t)n * K;
        w.alloc(sizeof(double) * nk);
        r.alloc(sizeof(double) * nk);
        d.alloc(sizeof(double) * nk);
        rr.alloc(sizeof(double) * nk);
        w64.alloc(sizeof(double) * nk);
        x32.alloc(sizeof(float) * nk);
        y32.alloc(sizeof(float) * nk);
        partial.alloc(sizeof(double) * (size_t)std::max(nchunk, 1) * (restart + 1) * K);
        coef.alloc(sizeof(double) * (size_t)P * (restart + 1) * K * 3);
        tab1.alloc(sizeof(double *));
        tabw.alloc(sizeof(double *));
        tabr.alloc(sizeof(double *));
        double *wp = w.as<double>(), *rp = r.as<double>();
        HIP_CHECK(hipMemcpy(tabw.p, &wp, sizeof(double *), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(tabr.p, &rp, sizeof(double *), hipMemcpyHostToDevice));
        const size_t NI = (size_t)P * K, ld = (size_t)restart + 1;
        gst.alloc(sizeof(GmresDevState) * NI);
        gH.alloc(sizeof(double) * NI * ld * restart);
        gvec.alloc(sizeof(double) * NI * ld * 4);
        gscale.alloc(sizeof(double) * NI);
        gnb2.alloc(sizeof(double) * NI);
        gr2.alloc(sizeof(double) * NI);
        gtol.alloc(sizeof(double) * NI);
        gbud.alloc(sizeof(int32_t) * NI);
        gcnt.alloc(sizeof(int32_t) * 4);
        hcap = (size_t)P * (restart + 1) * K * 3 + 16;
        HIP_CHECK(hipHostMalloc(&hbuf, sizeof(double) * hcap));
        for (int k = 0; k < 2; k++) {
            gexec[k].assign(restart, nullptr);
            ggen[k].assign(restart, 0);
        }
    }
    ~BatchWork() {
        for (auto &g : gexec)
            for (auto &e : g)
                if (e) (void)hipGraphExecDestroy(e);
        if (hbuf) (void)hipHostFree(hbuf);
    }
    double *basis(int i) {
        if (!V[i].p) {
            V[i].alloc(sizeof(double) * (size_t)n * K);
            vptr[i] = V[i].as<double>();
            HIP_CHECK(hipMemcpy(vtab.p, vptr.data(), sizeof(double *) * vptr.size(), hipMemcpyHostToDevice));
        }
        return vptr[i];
    }
    const double *const *table() const { return vtab.as<const double *>(); }
};

namespace {

// per-instance sums of a[row, c] * b[row, c] -> host out[p * K + c].
// Uses the third coefficient slot on the device and the matching range of
// the pinned staging buffer, so Hessenberg columns staged in the first two
// stay intact; synchronizes the stream.
void inst_dot(BatchWork &bw, const double *a, const double *b, std::vector<double> &out, hipStream_t s) {
    const size_t off = (size_t)2 * bw.P * (bw.restart + 1) * bw.K;
    const double **slot = reinterpret_cast<const double **>(bw.hbuf + bw.hcap - 1);
    *slot = a;
    HIP_CHECK(hipMemcpyAsync(bw.tab1.p, slot, sizeof(double *), hipMemcpyHostToDevice, s));
    HIP_CHECK(launch_bmdot(bw.tab1.as<const double *>(), 1, b, bw.chunks.as<Chunk>(), bw.nchunk,
                           bw.partial.as<double>(), bw.pchunk.as<int32_t>(), bw.P, bw.K,
                           bw.coef.as<double>() + off, s));
    const size_t ni = (size_t)bw.P * bw.K;
    HIP_CHECK(hipMemcpyAsync(bw.hbuf + off, bw.coef.as<double>() + off, sizeof(double) * ni, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    out.assign(bw.hbuf + off, bw.hbuf + off + ni);
}

}  // namespace

// Batched GMRES: A applies the operator to n x K device vectors; gkey names
// the operator precision A uses (the step graphs are cached per key), < 0:
// no graphs.  Instance a runs at most budget[a] Arnoldi steps (GMRES.hpp:
// 298-302 with a per-instance max_iter); one that reaches it is updated with
// the columns built and reported unconverged.
void gmres_batch(const std::function<void(const double *, double *)> &A, BatchWork &bw, const double *b,
                 double *x, bool x0_zero, const std::vector<double> &tol, const std::vector<int> &budget, hipStream_t s,
                 std::vector<int> &iters, std::vector<int> &conv, std::vector<double> &resid, int &matvecs,
                 int gkey, uint64_t gen) {
    const int m = bw.restart, K = bw.K, P = bw.P;
    const int NI = P * K, ld = m + 1;
    const int64_t nk = bw.n * K;
    double *w = bw.w.as<double>(), *r = bw.r.as<double>();
    double *cf = bw.coef.as<double>();  // [0, cfs) h1, [cfs, 2 cfs) h2, then |w|^2 or the update coefficients
    const size_t cfs = (size_t)P * (m + 1) * K;
    GmresDevState *st = bw.gst.as<GmresDevState>();
    double *H = bw.gH.as<double>();
    double *cs = bw.gvec.as<double>(), *sn = cs + (size_t)NI * ld, *sv = sn + (size_t)NI * ld,
           *y = sv + (size_t)NI * ld;
    double *scale = bw.gscale.as<double>();
    int32_t *cnt = bw.gcnt.as<int32_t>();
    int32_t *hcnt = reinterpret_cast<int32_t *>(bw.hbuf + bw.hcap - 4);  // pinned: running, breakdowns
    const Chunk *ck = bw.chunks.as<Chunk>();
    auto inst_dot_dev = [&](const double *const *tab, const double *v, double *out) {  // per-instance v . v
        HIP_CHECK(launch_bmdot(tab, 1, v, ck, bw.nchunk, bw.partial.as<double>(), bw.pchunk.as<int32_t>(), P, K,
                               out, s));
    };
    auto read_counts = [&] {
        HIP_CHECK(hipMemcpyAsync(hcnt, cnt, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (hcnt[1]) throw ArgError{HODLR_E_BREAKDOWN, "non-finite Arnoldi norm"};
        return hcnt[0];
    };
    // inputs (the stream is idle: nothing queued reads them)
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int32_t> bud(budget.begin(), budget.end());
    HIP_CHECK(hipMemcpy(bw.tab1.p, &b, sizeof(double *), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(bw.gtol.p, tol.data(), sizeof(double) * NI, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(bw.gbud.p, bud.data(), sizeof(int32_t) * NI, hipMemcpyHostToDevice));
    inst_dot_dev(bw.tab1.as<const double *>(), b, bw.gnb2.as<double>());
    HIP_CHECK(launch_bgmres_init(NI, bw.gnb2.as<double>(), bw.gtol.as<double>(), bw.gbud.as<int32_t>(), st, s));
    if (x0_zero) {
        HIP_CHECK(hipMemcpyAsync(r, b, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    } else {
        A(x, w);
        matvecs++;
        HIP_CHECK(launch_sub(b, w, r, nk, s));
    }
    inst_dot_dev(bw.tabr.as<const double *>(), r, bw.gr2.as<double>());
    HIP_CHECK(launch_bgmres_cycle(NI, bw.gr2.as<double>(), ld, 1, sv, st, scale, cnt, s));
    int nactive = read_counts();
    while (nactive > 0) {
        HIP_CHECK(launch_bscale(r, scale, bw.basis(0), ck, bw.nchunk, K, s));  // v_0 = r / beta
        int i = 0;
        for (; i < m; i++) {
            const int nv = i + 1;
            bw.basis(i);
            bw.basis(i + 1);  // both allocated (and tabled) before any capture
            auto step = [&] {
                A(bw.vptr[i], w);
                HIP_CHECK(launch_bmdot(bw.table(), nv, w, ck, bw.nchunk, bw.partial.as<double>(),
                                       bw.pchunk.as<int32_t>(), P, K, cf, s));
                HIP_CHECK(launch_bgemv(bw.table(), nv, cf, -1.0, w, ck, bw.nchunk, K, s));
                HIP_CHECK(launch_bmdot(bw.table(), nv, w, ck, bw.nchunk, bw.partial.as<double>(),
                                       bw.pchunk.as<int32_t>(), P, K, cf + cfs, s));
                HIP_CHECK(launch_bgemv(bw.table(), nv, cf + cfs, -1.0, w, ck, bw.nchunk, K, s));
                HIP_CHECK(launch_bmdot(bw.tabw.as<const double *>(), 1, w, ck, bw.nchunk, bw.partial.as<double>(),
                                       bw.pchunk.as<int32_t>(), P, K, cf + 2 * cfs, s));
                HIP_CHECK(launch_bgmres_step(NI, K, cf, cf + cfs, cf + 2 * cfs, i, ld, H, cs, sn, sv, st, scale, cnt,
                                             s));
                HIP_CHECK(launch_bscale(w, scale, bw.vptr[i + 1], ck, bw.nchunk, K, s));
                HIP_CHECK(hipMemcpyAsync(hcnt, cnt, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
            };
            if (gkey >= 0) {
                hipGraphExec_t &ex = bw.gexec[gkey][i];
                if (!ex || bw.ggen[gkey][i] != gen) {
                    if (ex) HIP_CHECK(hipGraphExecDestroy(ex));
                    ex = nullptr;
                    hipGraph_t g = nullptr;
                    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    try {
                        step();
                    } catch (...) {
                        (void)hipStreamEndCapture(s, &g);
                        if (g) (void)hipGraphDestroy(g);
                        throw;
                    }
                    HIP_CHECK(hipStreamEndCapture(s, &g));
                    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    HIP_CHECK(e);
                    bw.ggen[gkey][i] = gen;
                }
                HIP_CHECK(hipGraphLaunch(ex, s));
            } else {
                step();
            }
            matvecs++;
            HIP_CHECK(hipStreamSynchronize(s));
            if (hcnt[1]) throw ArgError{HODLR_E_BREAKDOWN, "non-finite Arnoldi norm"};
            nactive = hcnt[0];
            if (nactive == 0) break;
        }
        // end of the cycle (GMRES.hpp:271 uses m - 1 columns; an instance
        // that ended earlier in it uses the columns it built)
        const int kmax = i < m ? i : m - 1;
        HIP_CHECK(launch_bgmres_backsolve(NI, K, kmax, ld, H, sv, st, y, cf + 2 * cfs, s));
        HIP_CHECK(launch_bgemv(bw.table(), kmax + 1, cf + 2 * cfs, 1.0, x, ck, bw.nchunk, K, s));
        if (nactive == 0) break;
        A(x, w);
        matvecs++;
        HIP_CHECK(launch_sub(b, w, r, nk, s));
        inst_dot_dev(bw.tabr.as<const double *>(), r, bw.gr2.as<double>());
        HIP_CHECK(launch_bgmres_cycle(NI, bw.gr2.as<double>(), ld, 0, sv, st, scale, cnt, s));
        nactive = read_counts();
    }
    std::vector<GmresDevState> out(NI);
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipMemcpy(out.data(), st, sizeof(GmresDevState) * NI, hipMemcpyDeviceToHost));
    iters.assign(NI, 0);
    conv.assign(NI, 0);
    resid.assign(NI, 0.0);
    for (int a = 0; a < NI; a++) {
        iters[a] = out[a].iters;
        conv[a] = out[a].converged;
        resid[a] = out[a].resid;
    }
}

void solve_batch(hodlr_op &op, int K, const double *b, double *x, const hodlr_gmres_params &p,
                 hodlr_gmres_info *infos) {
    const double t0 = now_s();
    hipStream_t s = op.ctx->stream;
    const int64_t n = op.N, nk = n * (int64_t)K;
    const int P = (int)op.trees.size(), NI = P * K;
    const int restart = p.restart > 0 ? p.restart : 100;
    if (restart + 1 > 128) throw ArgError{HODLR_E_ARG, "restart must be at most 127"};
    int64_t nmax_prob = 0;
    for (int q = 0; q < P; q++) nmax_prob = std::max(nmax_prob, op.pos_base[q + 1] - op.pos_base[q]);
    const int max_iter = p.max_iter > 0 ? p.max_iter : (int)std::min<int64_t>(nmax_prob, INT32_MAX);
    auto cached = std::static_pointer_cast<BatchWork>(op.batch_ws);
    if (!cached || cached->K != K || cached->restart != restart) {
        op.batch_ws.reset();
        cached = std::make_shared<BatchWork>(op, K, restart);
        op.batch_ws = cached;
    }
    BatchWork &bw = *cached;
    for (int a = 0; a < NI; a++) std::memset(&infos[a], 0, sizeof(hodlr_gmres_info));
    if (p.x0_is_b == 1) HIP_CHECK(hipMemcpyAsync(x, b, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    const bool x0_zero = p.x0_is_b == 2;
    if (x0_zero) HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * nk, s));
    auto A64 = [&](const double *in, double *out) { op.matvec_dev(HODLR_PREC_F64, K, in, out, s); };
    // work buffers exist before any step is captured (a capture may not allocate)
    op.ensure_work(HODLR_PREC_F64, K);
    if (op.prec_flags & HODLR_PREC_F32) op.ensure_work(HODLR_PREC_F32, K);
    const bool graphs = op.solver_graphs && !op.profiling;
    std::vector<int> it, cv;
    std::vector<double> rs;
    int mv64 = 0, mv32 = 0;
    if (p.mode == HODLR_SOLVE_F64) {
        std::vector<double> tol(NI, p.tol);
        std::vector<int> budget(NI, max_iter);
        gmres_batch(A64, bw, b, x, x0_zero, tol, budget, s, it, cv, rs, mv64, graphs ? 0 : -1, op.work_gen);
        for (int a = 0; a < NI; a++) {
            infos[a].iterations = it[a];
            infos[a].converged = cv[a];
            infos[a].resid = rs[a];
            infos[a].true_resid = -1.0;
        }
    } else {
        if (!(op.prec_flags & HODLR_PREC_F32)) throw ArgError{HODLR_E_STATE, "operator has no FP32 copy"};
        op.ensure_work(HODLR_PREC_F32, K);
        float *x32 = bw.x32.as<float>(), *y32 = bw.y32.as<float>();
        auto A32 = [&](const double *in, double *out) {  // FP32 off-diagonal, diagonal in FP64
            HIP_CHECK(launch_to_f32(in, x32, nk, s));
            op.matvec_dev(HODLR_PREC_F32, K, x32, y32, s, nullptr, false);
            HIP_CHECK(launch_f32_result(y32, in, op.diag.as<double>(), out, n, K, s));
        };
        double *rr = bw.rr.as<double>(), *d = bw.d.as<double>(), *w64 = bw.w64.as<double>();
        std::vector<double> nb, r2;
        inst_dot(bw, b, b, nb, s);
        if (x0_zero) {
            HIP_CHECK(hipMemcpyAsync(rr, b, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        } else {
            A32(x, w64);
            mv32++;
            HIP_CHECK(launch_sub(b, w64, rr, nk, s));
        }
        inst_dot(bw, rr, rr, r2, s);
        std::vector<double> res(NI);
        for (int a = 0; a < NI; a++) res[a] = std::sqrt(r2[a]) / (nb[a] == 0.0 ? 1.0 : std::sqrt(nb[a]));
        bool res_f64 = x0_zero;
        const int max_outer = p.max_outer > 0 ? p.max_outer : 20;
        const double floor = p.inner_tol > 0 ? p.inner_tol : 1e-14;
        // as the single-problem mixed solve: max_iter bounds each instance's
        // inner iterations over all sweeps, and an instance whose sweep made
        // no progress is not refined again
        std::vector<int> total(NI, 0), budget(NI);
        std::vector<char> stopped(NI, 0);
        int outer = 0;
        auto done = [&](int a) { return (res_f64 && res[a] < p.tol) || stopped[a] || total[a] >= max_iter; };
        auto all_done = [&] {
            if (!res_f64) return false;
            for (int a = 0; a < NI; a++)
                if (!done(a)) return false;
            return true;
        };
        while (!all_done() && outer < max_outer) {
            std::vector<double> target(NI);
            for (int a = 0; a < NI; a++) {  // a finished instance gets an unreachable-to-miss target
                target[a] = done(a) ? 1e300 : std::min(0.5, std::max(floor, 0.3 * p.tol / std::max(res[a], 1e-300)));
                budget[a] = done(a) ? 0 : max_iter - total[a];
            }
            HIP_CHECK(hipMemsetAsync(d, 0, sizeof(double) * nk, s));
            gmres_batch(A32, bw, rr, d, true, target, budget, s, it, cv, rs, mv32, graphs ? 1 : -1, op.work_gen);
            for (int a = 0; a < NI; a++) {
                total[a] += it[a];
                if (budget[a] > 0 && it[a] == 0) stopped[a] = 1;  // no progress possible at this precision
            }
            HIP_CHECK(launch_add(x, d, nk, s));
            A64(x, w64);
            mv64++;
            HIP_CHECK(launch_sub(b, w64, rr, nk, s));
            inst_dot(bw, rr, rr, r2, s);
            for (int a = 0; a < NI; a++) res[a] = std::sqrt(r2[a]) / (nb[a] == 0.0 ? 1.0 : std::sqrt(nb[a]));
            res_f64 = true;
            outer++;
        }
        if (!res_f64) {
            A64(x, w64);
            mv64++;
            HIP_CHECK(launch_sub(b, w64, rr, nk, s));
            inst_dot(bw, rr, rr, r2, s);
            for (int a = 0; a < NI; a++) res[a] = std::sqrt(r2[a]) / (nb[a] == 0.0 ? 1.0 : std::sqrt(nb[a]));
        }
        for (int a = 0; a < NI; a++) {
            infos[a].iterations = total[a];
            infos[a].outer = outer;
            infos[a].converged = res[a] < p.tol;
            infos[a].resid = res[a];
            infos[a].true_resid = res[a];
        }
    }
    HIP_CHECK(hipStreamSynchronize(s));
    const double t1 = now_s();
    for (int a = 0; a < NI; a++) {
        infos[a].matvecs_f64 = mv64;
        infos[a].matvecs_f32 = mv32;
        infos[a].seconds = t1 - t0;
    }
}

}  // namespace hodlr
