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
// The device work of Arnoldi step i -- the matvec, both CGS2 passes, the
// norm of the new vector and the copies of the Hessenberg column to pinned
// host memory -- depends only on i (basis vectors, work buffers and the
// pinned staging are fixed per workspace), so it is captured once into a
// hipGraph per (operator precision, i) and replayed as one launch followed
// by one wait; the host then applies the Givens rotations.  A graph is
// recaptured when the operator's work buffers or kernel forms change.
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

void gen_rot(double dx, double dy, double &cs, double &sn) {
    if (dy == 0.0) {
        cs = 1.0;
        sn = 0.0;
    } else if (std::fabs(dy) > std::fabs(dx)) {
        double temp = dx / dy;
        sn = 1.0 / std::sqrt(1.0 + temp * temp);
        cs = temp * sn;
    } else {
        double temp = dy / dx;
        cs = 1.0 / std::sqrt(1.0 + temp * temp);
        sn = temp * cs;
    }
}

void app_rot(double &dx, double &dy, double cs, double sn) {
    double temp = cs * dx + sn * dy;
    dy = (-sn) * dx + cs * dy;
    dx = temp;
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
    DevBuf vtab, w, r, d, rr, w64, x32, y32, partial, coef, chunks, pchunk, tab1, tabw;
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
        double *wp = w.as<double>();
        HIP_CHECK(hipMemcpy(tabw.p, &wp, sizeof(double *), hipMemcpyHostToDevice));
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

struct Inst {
    std::vector<double> H;  // (m+1) x m column-major
    std::vector<double> sv, cs, sn;
    double normb = 1.0, beta = 0.0, resid = 0.0;
    bool active = false;
    bool converged = false;
    int iterations = 0;
};

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
    const int NI = P * K;
    const int64_t nk = bw.n * K;
    std::vector<Inst> in(NI);
    for (auto &q : in) {
        q.H.assign((size_t)(m + 1) * m, 0.0);
        q.sv.assign(m + 1, 0.0);
        q.cs.assign(m + 1, 0.0);
        q.sn.assign(m + 1, 0.0);
    }
    auto Hc = [&](Inst &q, int row, int c) -> double & { return q.H[(size_t)c * (m + 1) + row]; };
    std::vector<double> nb, bt;
    inst_dot(bw, b, b, nb, s);
    double *w = bw.w.as<double>(), *r = bw.r.as<double>();
    if (x0_zero) {
        HIP_CHECK(hipMemcpyAsync(r, b, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    } else {
        A(x, w);
        matvecs++;
        HIP_CHECK(launch_sub(b, w, r, nk, s));
    }
    inst_dot(bw, r, r, bt, s);
    int nactive = 0;
    for (int i = 0; i < NI; i++) {
        Inst &q = in[i];
        q.normb = std::sqrt(nb[i]) == 0.0 ? 1.0 : std::sqrt(nb[i]);
        q.beta = std::sqrt(bt[i]);
        q.resid = q.beta / q.normb;
        q.converged = q.resid <= tol[i];  // GMRES.hpp:198-213
        q.active = !q.converged && budget[i] > 0;
        q.iterations = 0;
        nactive += q.active;
    }
    int max_iter = 0;
    for (int i = 0; i < NI; i++) max_iter = std::max(max_iter, budget[i]);
    double *cf = bw.coef.as<double>();  // [0, P*(m+1)*K) h1; then h2; then scratch
    const size_t cfs = (size_t)P * (m + 1) * K;
    std::vector<double> hc(cfs);
    // x += V_0..k y for the instances in `sel` (others get zero coefficients)
    auto update = [&](const std::vector<int> &ks) {
        std::fill(hc.begin(), hc.end(), 0.0);
        int kmax = -1;
        for (int i = 0; i < NI; i++) {
            const int k = ks[i];
            if (k < 0) continue;
            Inst &q = in[i];
            std::vector<double> y(q.sv.begin(), q.sv.begin() + k + 1);
            for (int a = k; a >= 0; a--) {  // GMRES.hpp:136-141
                y[a] /= Hc(q, a, a);
                for (int c = a - 1; c >= 0; c--) y[c] -= Hc(q, c, a) * y[a];
            }
            const int p = i / K, c = i % K;
            for (int a = 0; a <= k; a++) hc[((size_t)p * (m + 1) + a) * K + c] = y[a];
            kmax = std::max(kmax, k);
        }
        if (kmax < 0) return;
        // coefficients are laid out (p, v, c) with nv = kmax + 1
        std::vector<double> packed((size_t)P * (kmax + 1) * K);
        for (int p = 0; p < P; p++)
            for (int a = 0; a <= kmax; a++)
                for (int c = 0; c < K; c++)
                    packed[((size_t)p * (kmax + 1) + a) * K + c] = hc[((size_t)p * (m + 1) + a) * K + c];
        std::memcpy(bw.hbuf, packed.data(), sizeof(double) * packed.size());
        HIP_CHECK(hipMemcpyAsync(cf + 2 * cfs, bw.hbuf, sizeof(double) * packed.size(), hipMemcpyHostToDevice, s));
        HIP_CHECK(launch_bgemv(bw.table(), kmax + 1, cf + 2 * cfs, 1.0, x, bw.chunks.as<Chunk>(), bw.nchunk, K, s));
        HIP_CHECK(hipStreamSynchronize(s));
    };
    std::vector<double> scale(NI);
    int j = 1;
    while (nactive > 0 && j <= max_iter) {
        for (int i = 0; i < NI; i++) scale[i] = in[i].active ? 1.0 / in[i].beta : 0.0;
        std::memcpy(bw.hbuf, scale.data(), sizeof(double) * NI);
        HIP_CHECK(hipMemcpyAsync(cf + 2 * cfs, bw.hbuf, sizeof(double) * NI, hipMemcpyHostToDevice, s));
        HIP_CHECK(launch_bscale(r, cf + 2 * cfs, bw.basis(0), bw.chunks.as<Chunk>(), bw.nchunk, K, s));
        for (auto &q : in) {
            std::fill(q.sv.begin(), q.sv.end(), 0.0);
            q.sv[0] = q.beta;
        }
        int i = 0;
        for (; i < m && j <= max_iter && nactive > 0; i++, j++) {
            const int nv = i + 1;
            const size_t nh = (size_t)P * nv * K;
            bw.basis(i);
            bw.basis(i + 1);  // both allocated (and tabled) before any capture
            // step i's device work: new vector, CGS2, its norm, and the copies
            // of h1, h2 and |w|^2 to hbuf[0, nh), [nh, 2 nh), [2 cfs, 2 cfs + NI)
            auto step = [&] {
                A(bw.vptr[i], w);
                HIP_CHECK(launch_bmdot(bw.table(), nv, w, bw.chunks.as<Chunk>(), bw.nchunk, bw.partial.as<double>(),
                                       bw.pchunk.as<int32_t>(), P, K, cf, s));
                HIP_CHECK(launch_bgemv(bw.table(), nv, cf, -1.0, w, bw.chunks.as<Chunk>(), bw.nchunk, K, s));
                HIP_CHECK(launch_bmdot(bw.table(), nv, w, bw.chunks.as<Chunk>(), bw.nchunk, bw.partial.as<double>(),
                                       bw.pchunk.as<int32_t>(), P, K, cf + cfs, s));
                HIP_CHECK(launch_bgemv(bw.table(), nv, cf + cfs, -1.0, w, bw.chunks.as<Chunk>(), bw.nchunk, K, s));
                HIP_CHECK(launch_bmdot(bw.tabw.as<const double *>(), 1, w, bw.chunks.as<Chunk>(), bw.nchunk,
                                       bw.partial.as<double>(), bw.pchunk.as<int32_t>(), P, K, cf + 2 * cfs, s));
                HIP_CHECK(hipMemcpyAsync(bw.hbuf, cf, sizeof(double) * nh, hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipMemcpyAsync(bw.hbuf + nh, cf + cfs, sizeof(double) * nh, hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipMemcpyAsync(bw.hbuf + 2 * cfs, cf + 2 * cfs, sizeof(double) * NI, hipMemcpyDeviceToHost,
                                         s));
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
            std::vector<double> hn2(bw.hbuf + 2 * cfs, bw.hbuf + 2 * cfs + NI);
            std::vector<double> h1(bw.hbuf, bw.hbuf + nh), h2(bw.hbuf + nh, bw.hbuf + 2 * nh);
            std::vector<int> ks(NI, -1);
            int nconv = 0;
            for (int a = 0; a < NI; a++) {
                Inst &q = in[a];
                const int p = a / K, c = a % K;
                if (!q.active) {
                    scale[a] = 0.0;
                    continue;
                }
                for (int k = 0; k <= i; k++)
                    Hc(q, k, i) = h1[((size_t)p * nv + k) * K + c] + h2[((size_t)p * nv + k) * K + c];
                const double hn = std::sqrt(hn2[a]);
                Hc(q, i + 1, i) = hn;
                for (int k = 0; k < i; k++) app_rot(Hc(q, k, i), Hc(q, k + 1, i), q.cs[k], q.sn[k]);
                gen_rot(Hc(q, i, i), Hc(q, i + 1, i), q.cs[i], q.sn[i]);
                app_rot(Hc(q, i, i), Hc(q, i + 1, i), q.cs[i], q.sn[i]);
                app_rot(q.sv[i], q.sv[i + 1], q.cs[i], q.sn[i]);
                q.resid = std::fabs(q.sv[i + 1]) / q.normb;
                if (!std::isfinite(hn)) throw ArgError{HODLR_E_BREAKDOWN, "non-finite Arnoldi norm"};
                if (q.resid < tol[a] || j >= budget[a]) {  // converged, or its budget ends here
                    ks[a] = i;
                    q.iterations = j;
                    q.converged = q.resid < tol[a];
                    nconv++;
                    scale[a] = 0.0;
                } else {
                    scale[a] = hn != 0.0 ? 1.0 / hn : 0.0;
                }
            }
            std::memcpy(bw.hbuf, scale.data(), sizeof(double) * NI);
            HIP_CHECK(hipMemcpyAsync(cf + 2 * cfs, bw.hbuf, sizeof(double) * NI, hipMemcpyHostToDevice, s));
            HIP_CHECK(launch_bscale(w, cf + 2 * cfs, bw.basis(i + 1), bw.chunks.as<Chunk>(), bw.nchunk, K, s));
            HIP_CHECK(hipStreamSynchronize(s));
            if (nconv) {
                update(ks);
                for (int a = 0; a < NI; a++)
                    if (ks[a] >= 0) {
                        in[a].active = false;
                        nactive--;
                    }
            }
        }
        if (nactive == 0) break;
        // end of a cycle (or of the iteration budget): update with the columns
        // built in it (GMRES.hpp:271 uses m - 1; fewer exist when the budget
        // ends mid-cycle)
        std::vector<int> ks(NI, -1);
        for (int a = 0; a < NI; a++)
            if (in[a].active) ks[a] = i - 1;
        update(ks);
        A(x, w);
        matvecs++;
        HIP_CHECK(launch_sub(b, w, r, nk, s));
        std::vector<double> b2;
        inst_dot(bw, r, r, b2, s);
        for (int a = 0; a < NI; a++) {
            Inst &q = in[a];
            if (!q.active) continue;
            q.beta = std::sqrt(b2[a]);
            q.resid = q.beta / q.normb;
            if (q.resid < tol[a]) {
                q.active = false;
                q.converged = true;
                q.iterations = j;
                nactive--;
            }
        }
    }
    iters.assign(NI, 0);
    conv.assign(NI, 0);
    resid.assign(NI, 0.0);
    for (int a = 0; a < NI; a++) {
        iters[a] = in[a].active ? std::min(j - 1, budget[a]) : in[a].iterations;
        conv[a] = in[a].converged;
        resid[a] = in[a].resid;
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
