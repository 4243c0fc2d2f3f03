// server.cpp -- continuous-batching solve server on one GPU.
//
// Three parts:
//   builder thread   builds queued operators one at a time (tree on the
//                    host, assembly on the builder's own stream).  The
//                    assembly reports the exact arena footprint after its
//                    ACA rank pass; the builder admits it against the HBM
//                    budget, evicting idle resident operators (least
//                    recently used first) or waiting for running solves to
//                    finish.  Evicted operators keep their points and are
//                    rebuilt when a request asks for them again.
//   engines          one per resident operator with work: K slots (columns
//                    of n x K interleaved device blocks).  Each column is an
//                    independent mixed-precision solve: restarted GMRES
//                    cycles on the FP32 operator (FP64 diagonal, FP64
//                    Krylov vectors and Givens rotations, the reference's
//                    stopping semantics, GMRES.hpp:160-303) for the
//                    correction, then an FP64 residual; the column ends when
//                    that residual is below tol.  Columns sit at different
//                    Krylov steps; one matvec per tick serves all of them.
//   scheduler thread admits waiting columns into free slots (continuous
//                    batching) and runs ticks: launch every engine's device
//                    work on its own stream, then collect.  Engines overlap.
//
// A column's arithmetic never mixes with other columns' (matvec kernels
// and reductions are per column, in fixed order), so its result is
// bit-identical alone or in any mix for the same slot count.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "op.h"

namespace hodlr {
namespace {

using Clock = std::chrono::steady_clock;
constexpr int64_t kChunkRows = 2048;
// An FP64 residual pass reads the whole FP64 operator for all slots, so it
// waits until it serves a quarter of the engine's solving columns, or until
// the oldest waiting column has waited this many ticks (or nothing else runs)
constexpr int kRefineWait = 4;
// Krylov vectors reserved in an engine's admission (they are allocated as
// the columns' cycles grow; the mixed solves here take 5-15 inner steps)
constexpr int kBasisReserve = 8;

void gen_rot(double dx, double dy, double &cs, double &sn) {  // GMRES.hpp:83-110
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

void app_rot(double &dx, double &dy, double cs, double sn) {  // GMRES.hpp:112-129
    double temp = cs * dx + sn * dy;
    dy = (-sn) * dx + cs * dy;
    dx = temp;
}

struct Request {
    int64_t id = 0, op_id = 0;
    int nrhs = 0;
    int64_t n = 0;
    std::vector<double> b, x;  // point order, n x nrhs column-major
    std::vector<hodlr_gmres_info> infos;
    int done_cols = 0;
    int status = 0;  // 0 queued, 1 running, 2 done, < 0 error
    bool built = false;
    double t_submit = 0, t_start = -1, t_done = -1;
};

struct Column {
    std::shared_ptr<Request> req;
    int rc = -1;     // column of the request
    int phase = 0;   // 0 free, 1 start (x = 0), 2 needs FP64 residual, 3 inner cycle
    bool v0 = false; // inner cycle starts this tick: V_0 = r / beta
    bool finish = false;
    bool in_a = false;  // its FP64 residual / start is computed this tick
    int wait = 0;       // ticks waited for an FP64 residual pass
    int i = 0;       // Krylov step inside the cycle
    double normb = 1.0, beta = 0.0, res = 0.0, target = 0.0, hn = 0.0;
    int iterations = 0, outer = 0, mv32 = 0, mv64 = 0;
    int fail = 0;    // status code of a breakdown seen this tick (its request fails)
    std::vector<double> H, sv, cs, sn, y;
    double t_admit = 0;
};

}  // namespace

// Solve engine of one operator: K slots, device blocks n x K.
struct Engine {
    hodlr_op &op;
    int K, m;
    int64_t n;
    hipStream_t s = nullptr;
    int nchunk = 0;
    DevBuf chunks, pchunk, b, x, rr, w64, in, w, x32, y32, partial, coef, tsrc, tdst, tlim, tscale, tcoef, dptr,
        colbuf;
    std::vector<DevBuf> V;
    std::vector<double *> vptr;
    DevBuf vtab;
    double *hbuf = nullptr;
    size_t hcap = 0;
    std::vector<Column> cols;
    double bytes = 0;
    bool launched_a = false, launched_b = false;
    double tol, floor;
    int max_outer;

    // device bytes an engine of K slots takes at creation (its blocks and
    // the operator's K-column matvec work buffers not allocated yet), plus
    // room for `basis` Krylov vectors
    static double estimate_bytes(const hodlr_op &o, int K, int basis) {
        const double nk = (double)o.N * K;
        double b = 8.0 * nk * 6 + 4.0 * nk * 2 + 8.0 * (double)std::max<int64_t>(o.N, 1) + 8.0 * nk * basis;
        size_t sz[6];
        if (o.cap_rhs32 < K && (o.prec_flags & HODLR_PREC_F32)) {
            o.matvec_work_sizes(HODLR_PREC_F32, K, sz);
            b += (double)(sz[0] + sz[1]);
        }
        if (o.cap_rhs64 < K) {
            o.matvec_work_sizes(HODLR_PREC_F64, K, sz);
            b += (double)(sz[0] + sz[1]);
        }
        return b;
    }

    Engine(hodlr_op &o, int K_, int m_, double tol_, double floor_, int max_outer_)
        : op(o), K(K_), m(m_), n(o.N), tol(tol_), floor(floor_), max_outer(max_outer_) {
        // non-blocking: the engine's work is ordered on its own stream only,
        // never behind null-stream calls made by builders or callers
        HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            init();
        } catch (...) {  // a failed allocation must not leak the stream or pinned memory
            release_handles();
            throw;
        }
    }
    void release_handles() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
            s = nullptr;
        }
        if (hbuf) (void)hipHostFree(hbuf);
        hbuf = nullptr;
    }
    void init() {
        std::vector<Chunk> ch;
        for (int64_t r0 = 0; r0 < n; r0 += kChunkRows) {
            Chunk c;
            c.row0 = r0;
            c.rows = (int32_t)std::min<int64_t>(kChunkRows, n - r0);
            c.prob = 0;
            ch.push_back(c);
        }
        nchunk = (int)ch.size();
        std::vector<int32_t> pc = {0, nchunk};
        upload(chunks, ch);
        upload(pchunk, pc);
        const size_t nk = (size_t)n * K;
        // blocks start zeroed: a masked (zero) coefficient times a column
        // that was never written must give zero, not NaN from stale memory
        for (DevBuf *d : {&b, &x, &rr, &w64, &in, &w}) {
            d->alloc(sizeof(double) * nk);
            HIP_CHECK(hipMemsetAsync(d->p, 0, sizeof(double) * nk, s));
        }
        x32.alloc(sizeof(float) * nk);
        y32.alloc(sizeof(float) * nk);
        partial.alloc(sizeof(double) * (size_t)std::max(nchunk, 1) * (m + 1) * K);
        coef.alloc(sizeof(double) * (size_t)(m + 1) * K * 3);
        tsrc.alloc(sizeof(double *) * K);
        tdst.alloc(sizeof(double *) * K);
        tlim.alloc(sizeof(int32_t) * K);
        tscale.alloc(sizeof(double) * K);
        tcoef.alloc(sizeof(double) * (size_t)(m + 1) * K);
        colbuf.alloc(sizeof(double) * (size_t)std::max<int64_t>(n, 1));
        V.resize(m + 1);
        vptr.assign(m + 1, nullptr);
        vtab.alloc(sizeof(double *) * (m + 1));
        hcap = (size_t)(m + 1) * K * 3 + 8 * (size_t)K + 16;
        HIP_CHECK(hipHostMalloc(&hbuf, sizeof(double) * hcap));
        const double *pt[3] = {rr.as<double>(), b.as<double>(), w.as<double>()};
        dptr.alloc(sizeof(pt));
        HIP_CHECK(hipMemcpy(dptr.p, pt, sizeof(pt), hipMemcpyHostToDevice));
        cols.resize(K);
        // the blocks and the operator's K-column work buffers (t, partial
        // sums), which exist while the operator is resident
        bytes = estimate_bytes(op, K, 0);
        op.ensure_work(HODLR_PREC_F32, K);
        op.ensure_work(HODLR_PREC_F64, K);
    }
    ~Engine() { release_handles(); }

    double *basis(int i) {
        if (!V[i].p) {
            V[i].alloc(sizeof(double) * (size_t)n * K);
            HIP_CHECK(hipMemsetAsync(V[i].p, 0, sizeof(double) * (size_t)n * K, s));
            vptr[i] = V[i].as<double>();
            bytes += 8.0 * (double)n * K;
            HIP_CHECK(hipMemcpyAsync(vtab.p, vptr.data(), sizeof(double *) * vptr.size(), hipMemcpyHostToDevice, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        return vptr[i];
    }
    const double *const *table() const { return vtab.as<const double *>(); }

    int free_slots() const {
        int f = 0;
        for (auto &c : cols) f += c.phase == 0;
        return f;
    }
    bool busy() const {
        for (auto &c : cols)
            if (c.phase) return true;
        return false;
    }

    // pinned staging (disjoint areas).  Uploads are only rewritten after a
    // stream synchronisation that follows their copies; downloads are read
    // after the synchronisation that follows theirs.
    double **up_src() { return reinterpret_cast<double **>(hbuf); }          // K pointers
    double **up_dst() { return reinterpret_cast<double **>(hbuf + K); }      // K pointers
    double *up_scale() { return hbuf + 2 * (size_t)K; }                        // K
    int32_t *up_lim() { return reinterpret_cast<int32_t *>(hbuf + 3 * (size_t)K); }  // K ints
    double *up_coef() { return hbuf + 4 * (size_t)K; }                         // (m+1) K
    double *dn_h1() { return hbuf + 4 * (size_t)K + (size_t)(m + 1) * K; }     // (m+1) K
    double *dn_h2() { return dn_h1() + (size_t)(m + 1) * K; }                  // (m+1) K
    double *dn_norm() { return dn_h2() + (size_t)(m + 1) * K; }                // 4 K

    // per-column sums v[r, c] * v2[r, c] into out (device, K values);
    // which = 0 rr, 1 b, 2 w
    void inst_dot(int which, const double *v2, double *out) {
        HIP_CHECK(launch_bmdot(dptr.as<const double *>() + which, 1, v2, chunks.as<Chunk>(), nchunk,
                               partial.as<double>(), pchunk.as<int32_t>(), 1, K, out, s));
    }

    // A32: FP32 off-diagonal operator, diagonal applied in FP64
    void a32(const double *src, double *dst) {
        const int64_t nk = n * (int64_t)K;
        HIP_CHECK(launch_to_f32(src, x32.as<float>(), nk, s));
        op.matvec_dev(HODLR_PREC_F32, K, x32.as<float>(), y32.as<float>(), s, nullptr, false);
        HIP_CHECK(launch_f32_result(y32.as<float>(), src, op.diag.as<double>(), dst, n, K, s));
    }

    // ---- admission of one column: b_c (tree order, host) ----
    void admit(int c, std::shared_ptr<Request> req, int rc, const std::vector<double> &bt, double now) {
        Column &C = cols[c];
        C = Column();
        C.req = std::move(req);
        C.rc = rc;
        C.phase = 1;
        C.t_admit = now;
        C.H.assign((size_t)(m + 1) * m, 0.0);
        C.sv.assign(m + 1, 0.0);
        C.cs.assign(m + 1, 0.0);
        C.sn.assign(m + 1, 0.0);
        HIP_CHECK(hipStreamSynchronize(s));
        HIP_CHECK(hipMemcpyAsync(colbuf.p, bt.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIP_CHECK(launch_put_col(b.as<double>(), colbuf.as<double>(), n, K, c, s));
        HIP_CHECK(launch_put_col(x.as<double>(), nullptr, n, K, c, s));
        HIP_CHECK(hipStreamSynchronize(s));  // colbuf is reused by the next admission
    }

    // columns starting (x = 0: residual b, no matvec) join every phase A;
    // refining columns join when the FP64 pass is due (kRefineWait)
    bool plan_a() {
        int n1 = 0, n2 = 0, n3 = 0, maxw = 0;
        for (auto &c : cols) {
            n1 += c.phase == 1;
            n2 += c.phase == 2;
            n3 += c.phase == 3;
            if (c.phase == 2) maxw = std::max(maxw, c.wait);
        }
        const bool refine = n2 > 0 && (n3 == 0 || 4 * n2 >= n2 + n3 || maxw >= kRefineWait);
        for (auto &c : cols) {
            c.in_a = c.phase == 1 || (c.phase == 2 && refine);
            if (c.phase == 2) c.wait = refine ? 0 : c.wait + 1;
        }
        return n1 > 0 || refine;
    }
    bool needs_b() const {
        for (auto &c : cols)
            if (c.phase == 3) return true;
        return false;
    }

    // ---- phase A: FP64 residual rr = b - A64 x of starting / refining columns ----
    void launch_a() {
        launched_a = false;
        if (!plan_a()) return;
        HIP_CHECK(hipStreamSynchronize(s));  // staging buffers of the previous tick
        const int64_t nk = n * (int64_t)K;
        bool refine = false;
        for (auto &c : cols) refine = refine || (c.phase == 2 && c.in_a);
        if (refine) {  // columns just admitted have x = 0: their residual is b exactly
            op.matvec_dev(HODLR_PREC_F64, K, x.p, w64.p, s);
            HIP_CHECK(launch_sub(b.as<double>(), w64.as<double>(), rr.as<double>(), nk, s));
            for (auto &c : cols)
                if (c.phase == 2 && c.in_a) c.mv64++;
        } else {
            HIP_CHECK(hipMemcpyAsync(rr.p, b.p, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        }
        double *cf = coef.as<double>() + (size_t)(m + 1) * K * 2;
        inst_dot(0, rr.as<double>(), cf);
        inst_dot(1, b.as<double>(), cf + K);
        HIP_CHECK(hipMemcpyAsync(dn_norm(), cf, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, s));
        launched_a = true;
    }

    // completed columns are appended to `done`
    void finish_a(std::vector<int> &done) {
        if (!launched_a) return;
        HIP_CHECK(hipStreamSynchronize(s));
        const double *r2 = dn_norm(), *b2 = dn_norm() + K;
        for (int c = 0; c < K; c++) {
            Column &C = cols[c];
            if ((C.phase != 1 && C.phase != 2) || !C.in_a) continue;
            if (!std::isfinite(r2[c]) || !std::isfinite(b2[c])) {
                C.fail = HODLR_E_BREAKDOWN;  // non-finite right-hand side or residual
                continue;
            }
            if (C.phase == 1) C.normb = std::sqrt(b2[c]) == 0.0 ? 1.0 : std::sqrt(b2[c]);
            else C.outer++;
            C.beta = std::sqrt(r2[c]);
            C.res = C.beta / C.normb;
            if (C.res < tol || (C.phase == 1 && C.res <= tol) || C.outer >= max_outer) {
                done.push_back(c);
                continue;
            }
            // aim the cycle at the outer tolerance (never looser than 0.5,
            // never tighter than the floor)
            C.target = std::min(0.5, std::max(floor, 0.3 * tol / std::max(C.res, 1e-300)));
            C.phase = 3;
            C.v0 = true;
            C.i = 0;
            std::fill(C.sv.begin(), C.sv.end(), 0.0);
            C.sv[0] = C.beta;
        }
    }

    // ---- phase B: one Arnoldi step of every column in an inner cycle ----
    void launch_b() {
        launched_b = false;
        if (!needs_b()) return;
        HIP_CHECK(hipStreamSynchronize(s));
        double **hsrc = up_src(), **hdst = up_dst();
        double *hscale = up_scale();
        int32_t *hlim = up_lim();
        bool any_v0 = false;
        int nv = 0;
        for (int c = 0; c < K; c++) {
            Column &C = cols[c];
            hdst[c] = nullptr;
            hscale[c] = 0.0;
            if (C.phase == 3 && C.v0) {
                hdst[c] = basis(0);
                hscale[c] = C.beta != 0.0 ? 1.0 / C.beta : 0.0;
                any_v0 = true;
            }
        }
        for (int c = 0; c < K; c++) {
            Column &C = cols[c];
            hsrc[c] = C.phase == 3 ? basis(C.i) : nullptr;
            hlim[c] = C.phase == 3 ? C.i + 1 : 0;
            if (C.phase == 3) nv = std::max(nv, C.i + 1);
        }
        HIP_CHECK(hipMemcpyAsync(tdst.p, hdst, sizeof(double *) * K, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(tscale.p, hscale, sizeof(double) * K, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(tsrc.p, hsrc, sizeof(double *) * K, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(tlim.p, hlim, sizeof(int32_t) * K, hipMemcpyHostToDevice, s));
        if (any_v0)
            HIP_CHECK(launch_scatter_cols(rr.as<double>(), tscale.as<double>(), tdst.as<double *const>(), n, K, s));
        HIP_CHECK(launch_gather_cols(tsrc.as<const double *const>(), in.as<double>(), n, K, s));
        a32(in.as<double>(), w.as<double>());
        double *cf = coef.as<double>();
        double *cf2 = cf + (size_t)(m + 1) * K;
        double *cf3 = cf + (size_t)(m + 1) * K * 2;
        const Chunk *ck = chunks.as<Chunk>();
        // CGS2 against each column's own basis vectors 0..i (others masked)
        HIP_CHECK(launch_bmdot(table(), nv, w.as<double>(), ck, nchunk, partial.as<double>(), pchunk.as<int32_t>(),
                               1, K, cf, s));
        HIP_CHECK(launch_mask_coef(cf, nv, K, tlim.as<int32_t>(), s));
        HIP_CHECK(launch_bgemv(table(), nv, cf, -1.0, w.as<double>(), ck, nchunk, K, s));
        HIP_CHECK(launch_bmdot(table(), nv, w.as<double>(), ck, nchunk, partial.as<double>(), pchunk.as<int32_t>(),
                               1, K, cf2, s));
        HIP_CHECK(launch_mask_coef(cf2, nv, K, tlim.as<int32_t>(), s));
        HIP_CHECK(launch_bgemv(table(), nv, cf2, -1.0, w.as<double>(), ck, nchunk, K, s));
        inst_dot(2, w.as<double>(), cf3);
        HIP_CHECK(hipMemcpyAsync(dn_h1(), cf, sizeof(double) * nv * K, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(dn_h2(), cf2, sizeof(double) * nv * K, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(dn_norm(), cf3, sizeof(double) * K, hipMemcpyDeviceToHost, s));
        b_nv = nv;
        launched_b = true;
    }
    int b_nv = 0;

    void finish_b() {
        if (!launched_b) return;
        HIP_CHECK(hipStreamSynchronize(s));
        const int nv = b_nv;
        const double *h1 = dn_h1(), *h2 = dn_h2(), *hn2 = dn_norm();
        auto Hc = [&](Column &C, int row, int col) -> double & { return C.H[(size_t)col * (m + 1) + row]; };
        for (int c = 0; c < K; c++) {
            Column &C = cols[c];
            if (C.phase != 3) continue;
            C.v0 = false;
            C.mv32++;
            C.iterations++;
            const int i = C.i;
            for (int k = 0; k <= i; k++) Hc(C, k, i) = h1[(size_t)k * K + c] + h2[(size_t)k * K + c];
            (void)nv;
            const double hn = std::sqrt(hn2[c]);
            bool finite = std::isfinite(hn);
            for (int k = 0; k <= i && finite; k++) finite = std::isfinite(Hc(C, k, i));
            if (!finite) {  // this column's request fails; the other columns go on
                C.fail = HODLR_E_BREAKDOWN;
                continue;
            }
            Hc(C, i + 1, i) = hn;
            C.hn = hn;
            for (int k = 0; k < i; k++) app_rot(Hc(C, k, i), Hc(C, k + 1, i), C.cs[k], C.sn[k]);
            gen_rot(Hc(C, i, i), Hc(C, i + 1, i), C.cs[i], C.sn[i]);
            app_rot(Hc(C, i, i), Hc(C, i + 1, i), C.cs[i], C.sn[i]);
            app_rot(C.sv[i], C.sv[i + 1], C.cs[i], C.sn[i]);
            const double resid = std::fabs(C.sv[i + 1]) / (C.beta == 0.0 ? 1.0 : C.beta);
            C.finish = resid < C.target || i + 1 == m;
            if (C.finish) {  // back-solve (GMRES.hpp:132-146) over the columns built
                C.y.assign(C.sv.begin(), C.sv.begin() + i + 1);
                for (int a = i; a >= 0; a--) {
                    C.y[a] /= Hc(C, a, a);
                    for (int q = a - 1; q >= 0; q--) C.y[q] -= Hc(C, q, a) * C.y[a];
                }
            }
        }
    }

    // ---- phase C: next basis vectors; x += V y for finished cycles ----
    void launch_c() {
        if (!launched_b) return;
        launched_b = false;
        double **hdst = up_dst();
        double *hscale = up_scale();
        double *hcoef = up_coef();
        bool any_next = false;
        int nvu = 0;
        for (int c = 0; c < K; c++) {
            Column &C = cols[c];
            hdst[c] = nullptr;
            hscale[c] = 0.0;
            if (C.phase != 3 || C.fail) continue;
            if (C.finish) {
                nvu = std::max(nvu, C.i + 1);
            } else {
                hdst[c] = basis(C.i + 1);
                hscale[c] = C.hn != 0.0 ? 1.0 / C.hn : 0.0;
                any_next = true;
            }
        }
        if (nvu)
            for (int v = 0; v < nvu; v++)
                for (int c = 0; c < K; c++) {
                    const Column &C = cols[c];
                    hcoef[(size_t)v * K + c] = (C.phase == 3 && !C.fail && C.finish && v <= C.i) ? C.y[v] : 0.0;
                }
        if (any_next) {
            HIP_CHECK(hipMemcpyAsync(tdst.p, hdst, sizeof(double *) * K, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(tscale.p, hscale, sizeof(double) * K, hipMemcpyHostToDevice, s));
            HIP_CHECK(launch_scatter_cols(w.as<double>(), tscale.as<double>(), tdst.as<double *const>(), n, K, s));
        }
        if (nvu) {
            HIP_CHECK(hipMemcpyAsync(tcoef.p, hcoef, sizeof(double) * nvu * K, hipMemcpyHostToDevice, s));
            HIP_CHECK(launch_bgemv(table(), nvu, tcoef.as<double>(), 1.0, x.as<double>(), chunks.as<Chunk>(), nchunk,
                                   K, s));
        }
        for (auto &C : cols) {
            if (C.phase != 3 || C.fail) continue;
            if (C.finish) {
                C.phase = 2;
                C.finish = false;
            } else {
                C.i++;
            }
        }
    }

    // zeroes every device column of slot c (stream-ordered): a failed
    // column may leave non-finite values, and a later column in the slot
    // meets them through masked (zero) coefficients, where 0 * NaN = NaN
    // (a launch error is left to the next checked call of the tick, which
    // fails every request on the device)
    void clear_slot(int c) {
        for (DevBuf *d : {&b, &x, &rr, &w64, &in, &w}) (void)launch_put_col(d->as<double>(), nullptr, n, K, c, s);
        for (auto &v : V)
            if (v.p) (void)launch_put_col(v.as<double>(), nullptr, n, K, c, s);
    }

    // solution column c in tree order
    void read_x(int c, std::vector<double> &xt) {
        xt.resize(n);
        HIP_CHECK(hipStreamSynchronize(s));
        HIP_CHECK(launch_get_col(x.as<double>(), colbuf.as<double>(), n, K, c, s));
        HIP_CHECK(hipMemcpyAsync(xt.data(), colbuf.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
};

struct OpEntry {
    int64_t id = 0;
    int state = 0;  // 0 queued for build, 1 building, 2 ready, 3 evicted, < 0 failed (status code)
    int ndim = 0;
    int64_t n = 0;
    std::vector<double> pts, lo, hi, weight;
    hodlr_kernel_desc kd{};
    hodlr_tree *tree = nullptr;
    hodlr_op *op = nullptr;
    std::unique_ptr<Engine> engine;
    double bytes = 0;         // operator arenas (admitted)
    double engine_bytes = 0;  // engine blocks (accounted when created / grown)
    double last_use = 0;
    int refs = 0;             // requests queued or running on it
    bool in_tick = false;     // the scheduler is running its engine outside the lock
    bool engine_deferred = false;  // its engine waits for memory that running requests will free
    std::deque<std::pair<std::shared_ptr<Request>, int>> waiting;  // columns without a slot
    std::string error;
};

}  // namespace hodlr

struct hodlr_server {
    // device memory of the builds and engines (declared first: released last)
    std::unique_ptr<hodlr::DevSlab> slab;
    hodlr_server_config cfg{};
    int device = 0;
    std::vector<std::unique_ptr<hodlr_ctx>> bctx;  // one device + stream per builder
    std::mutex mu;
    std::condition_variable cv_build, cv_work, cv_done, cv_mem;
    std::map<int64_t, std::unique_ptr<hodlr::OpEntry>> ops;
    std::map<int64_t, std::shared_ptr<hodlr::Request>> reqs;
    int64_t next_op = 1, next_req = 1;
    bool stop = false;
    std::vector<std::thread> builders;
    std::thread scheduler;
    hodlr::Clock::time_point t0 = hodlr::Clock::now();
    hodlr_server_stats st{};
    double used = 0;  // admitted operator bytes + engine bytes
    double tick_t0 = 0;
    std::string last_error;

    double now() const { return std::chrono::duration<double>(hodlr::Clock::now() - t0).count(); }
    void build_loop(int idx);
    void sched_loop();
    void admit_memory(hodlr::OpEntry &e, double bytes);
    bool evict_one(int64_t keep);
    void free_op(hodlr::OpEntry &e);
    void fail_op(hodlr::OpEntry &e, int code, const std::string &msg);
    void complete_column(hodlr::OpEntry &e, int c, std::vector<double> &xt);
    void fail_request(hodlr::OpEntry &e, const std::shared_ptr<hodlr::Request> &r, int code, const std::string &msg);
    void fail_columns(hodlr::OpEntry &e);
    void fail_waiting(hodlr::OpEntry &e, int code, const std::string &msg);
    void retry_deferred();
};

using hodlr::Column;
using hodlr::Engine;
using hodlr::OpEntry;
using hodlr::Request;

// releases an operator's device memory (lock held; no engine work may be in
// flight for it: refs == 0 or the scheduler has dropped it)
void hodlr_server::free_op(OpEntry &e) {
    (void)hipSetDevice(device);
    // one device-wide wait, then the slab takes the ranges back without a
    // wait per buffer
    (void)hipDeviceSynchronize();
    hodlr::tl_slab_synced = true;
    e.engine.reset();
    if (e.op) hodlr_op_destroy(e.op);
    hodlr::tl_slab_synced = false;
    if (e.tree) hodlr_tree_destroy(e.tree);
    e.op = nullptr;
    e.tree = nullptr;
    used -= e.bytes + e.engine_bytes;
    e.bytes = e.engine_bytes = 0;
}

bool hodlr_server::evict_one(int64_t keep) {
    OpEntry *best = nullptr;
    for (auto &kv : ops) {
        OpEntry &e = *kv.second;
        if (e.id == keep || e.state != 2 || e.refs > 0 || e.in_tick) continue;
        if (!best || e.last_use < best->last_use) best = &e;
    }
    if (!best) return false;
    free_op(*best);
    best->state = 3;
    st.evictions++;
    return true;
}

// HBM admission of an assembly (builder thread, from inside assemble)
void hodlr_server::admit_memory(OpEntry &e, double bytes) {
    std::unique_lock<std::mutex> lk(mu);
    if (bytes > st.budget_bytes)
        throw hodlr::ArgError{HODLR_E_NOMEM, "operator needs " + std::to_string(bytes * 1e-9) +
                                                 " GB, more than the server budget"};
    while (used + bytes > st.budget_bytes) {
        if (evict_one(e.id)) continue;
        bool others = false;  // anything that will release memory later
        for (auto &kv : ops)
            if (kv.first != e.id && (kv.second->bytes > 0 || kv.second->engine_bytes > 0)) others = true;
        if (!others || stop)
            throw hodlr::ArgError{HODLR_E_NOMEM, "operator does not fit next to the resident ones"};
        cv_mem.wait_for(lk, std::chrono::milliseconds(50));
    }
    used += bytes;
    e.bytes = bytes;
    st.peak_bytes = std::max(st.peak_bytes, used);
}

void hodlr_server::fail_op(OpEntry &e, int code, const std::string &msg) {
    e.state = code;
    e.error = msg;
    for (auto &w : e.waiting) {
        Request &r = *w.first;
        if (r.status >= 0) {
            r.status = code;
            r.t_done = now();
            e.refs--;
        }
    }
    e.waiting.clear();
    last_error = msg;
    cv_done.notify_all();
}

void hodlr_server::build_loop(int idx) {
    (void)hipSetDevice(device);
    hodlr::tl_slab = slab.get();
    // concurrent builds share the host: 16 threads between them
    hodlr::set_assembly_threads(std::max(1, 16 / (int)bctx.size()));
    hodlr_ctx &bc = *bctx[idx];
    while (true) {
        OpEntry *e = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_build.wait(lk, [&] {
                if (stop) return true;
                for (auto &kv : ops)
                    if (kv.second->state == 0) return true;
                return false;
            });
            for (auto &kv : ops)
                if (kv.second->state == 0) {
                    e = kv.second.get();
                    break;
                }
            if (!e) return;  // stopping with nothing queued
            e->state = 1;
        }
        const double tb = now();
        hodlr_tree *t = nullptr;
        hodlr_op *op = nullptr;
        int rc = hodlr_tree_build(&bc, e->ndim, e->n, e->pts.data(), e->lo.data(), e->hi.data(), cfg.nmax,
                                  cfg.dprime, &t);
        std::string err = rc ? bc.last_error : "";
        if (rc == HODLR_OK) {
            hodlr_kernel_desc kd = e->kd;
            kd.weight = e->weight.empty() ? nullptr : e->weight.data();
            try {
                (void)hipSetDevice(device);
                op = new hodlr_op;
                op->ctx = &bc;
                op->trees.assign(1, t);
                op->kds.assign(1, kd);
                op->eps = cfg.eps_aca;
                op->prec_flags = HODLR_PREC_F64 | HODLR_PREC_F32;
                op->admit = [this, e](double bytes) { admit_memory(*e, bytes); };
                hodlr::assemble(*op);
                for (auto &k : op->kds) k.weight = nullptr;
                op->kd.weight = nullptr;
                op->admit = nullptr;
            } catch (const hodlr::ArgError &ex) {
                rc = ex.code;
                err = ex.msg;
            } catch (const hodlr::HipError &ex) {
                rc = HODLR_E_DEVICE;
                err = std::string(hipGetErrorString(ex.code)) + " in " + ex.where;
            } catch (const std::exception &ex) {
                rc = HODLR_E_STATE;
                err = ex.what();
            }
            if (rc != HODLR_OK && op) {
                delete op;
                op = nullptr;
            }
        }
        std::unique_lock<std::mutex> lk(mu);
        st.build_seconds += now() - tb;
        if (rc != HODLR_OK) {
            if (t) hodlr_tree_destroy(t);
            used -= e->bytes;
            e->bytes = 0;
            fail_op(*e, rc, err.empty() ? "operator build failed" : err);
        } else {
            e->tree = t;
            e->op = op;
            e->state = 2;
            e->last_use = now();
            st.builds++;
            cv_work.notify_all();
        }
    }
}

// one request fails (lock held): its status is set once and the operator's
// reference dropped once; its columns still waiting for a slot are dropped
// and its columns in slots are freed and zeroed.  Other requests, on this
// engine or others, are untouched.
void hodlr_server::fail_request(OpEntry &e, const std::shared_ptr<Request> &r, int code, const std::string &msg) {
    if (r->status >= 0) {
        r->status = code;
        r->t_done = now();
        e.refs--;
        e.last_use = now();
    }
    for (auto it = e.waiting.begin(); it != e.waiting.end();)
        it = it->first == r ? e.waiting.erase(it) : it + 1;
    if (e.engine)
        for (int c = 0; c < e.engine->K; c++) {
            Column &C = e.engine->cols[c];
            if (C.phase && C.req == r) {
                e.engine->clear_slot(c);
                C = Column();
            }
        }
    last_error = msg;
    retry_deferred();
    cv_done.notify_all();
    cv_mem.notify_all();
}

// every request of an operator still waiting for a slot fails (lock held)
void hodlr_server::fail_waiting(OpEntry &e, int code, const std::string &msg) {
    std::vector<std::shared_ptr<Request>> rs;
    for (auto &w : e.waiting)
        if (std::find(rs.begin(), rs.end(), w.first) == rs.end()) rs.push_back(w.first);
    for (auto &r : rs) fail_request(e, r, code, msg);
}

// a request ended: engines deferred for memory may fit now (lock held)
void hodlr_server::retry_deferred() {
    for (auto &kv : ops) kv.second->engine_deferred = false;
    cv_work.notify_all();
}

// requests of the columns that broke down in this tick's phases (lock held)
void hodlr_server::fail_columns(OpEntry &e) {
    std::vector<std::pair<std::shared_ptr<Request>, int>> bad;
    for (auto &C : e.engine->cols)
        if (C.phase && C.fail) bad.emplace_back(C.req, C.fail);
    for (auto &b : bad)
        fail_request(e, b.first, b.second,
                     "request " + std::to_string(b.first->id) + ": non-finite residual or Arnoldi norm");
}

void hodlr_server::complete_column(OpEntry &e, int c, std::vector<double> &xt) {
    Column &C = e.engine->cols[c];
    Request &r = *C.req;
    if (r.status < 0) {  // the request failed in another column
        C = Column();
        return;
    }
    const auto &perm = e.tree->perm;
    for (int64_t q = 0; q < r.n; q++) r.x[(size_t)C.rc * r.n + perm[q]] = xt[q];
    hodlr_gmres_info &in = r.infos[C.rc];
    in.iterations = C.iterations;
    in.outer = C.outer;
    in.matvecs_f32 = C.mv32;
    in.matvecs_f64 = C.mv64;
    in.converged = C.res < cfg.tol;
    in.resid = C.res;
    in.true_resid = C.res;
    in.seconds = now() - C.t_admit;
    r.done_cols++;
    st.columns_done++;
    if (r.done_cols == r.nrhs) {
        r.status = 2;
        r.t_done = now();
        st.requests_done++;
        e.refs--;
        e.last_use = now();
        retry_deferred();
        cv_done.notify_all();
        cv_mem.notify_all();
    }
    C = Column();
}

void hodlr_server::sched_loop() {
    (void)hipSetDevice(device);
    hodlr::tl_slab = slab.get();
    while (true) {
        std::vector<OpEntry *> active;
        std::vector<std::tuple<OpEntry *, int, std::shared_ptr<Request>, int>> admits;
        std::vector<OpEntry *> need_engine;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                admits.clear();
                active.clear();
                need_engine.clear();
                for (auto &kv : ops) {
                    OpEntry &e = *kv.second;
                    if (e.state == 2 && !e.waiting.empty() && !e.engine && !e.engine_deferred) need_engine.push_back(&e);
                }
                if (!need_engine.empty()) break;  // engines first, then admission
                for (auto &kv : ops) {
                    OpEntry &e = *kv.second;
                    if (e.state != 2 || !e.engine) continue;
                    // continuous batching: waiting columns take free slots now,
                    // whatever Krylov step the other columns are at
                    bool admitted = false;
                    for (int c = 0; c < e.engine->K && !e.waiting.empty(); c++) {
                        if (e.engine->cols[c].phase != 0) continue;
                        auto w = e.waiting.front();
                        e.waiting.pop_front();
                        admits.emplace_back(&e, c, w.first, w.second);
                        admitted = true;
                    }
                    if (e.engine->busy() || admitted) active.push_back(&e);
                }
                if (!admits.empty() || !active.empty()) {
                    for (OpEntry *e : active) e->in_tick = true;
                    break;
                }
                if (stop) return;
                cv_work.wait(lk);
            }
        }
        try {
            for (OpEntry *e : need_engine) {  // engines are created outside the lock
                {
                    // an engine's memory is admitted like an operator's:
                    // idle operators are evicted (least recently used
                    // first); when running requests hold the rest, the
                    // engine waits until one of them completes
                    std::unique_lock<std::mutex> lk(mu);
                    const double need = Engine::estimate_bytes(*e->op, cfg.slots, hodlr::kBasisReserve);
                    while (used + need > st.budget_bytes && evict_one(e->id)) {
                    }
                    if (used + need > st.budget_bytes) {
                        bool running = false;
                        for (auto &kv : ops)
                            if (kv.first != e->id && kv.second->engine && kv.second->engine->busy()) running = true;
                        if (running) {
                            e->engine_deferred = true;
                            continue;
                        }
                        // nothing will release memory: try (the slab keeps
                        // room beyond the budget)
                    }
                }
                std::unique_ptr<Engine> eng;
                try {
                    eng = std::make_unique<Engine>(*e->op, cfg.slots, cfg.restart, cfg.tol, cfg.inner_floor,
                                                   cfg.max_outer);
                } catch (const hodlr::ArgError &ex) {
                    // no device memory for this operator's engine: its
                    // waiting requests fail (never retried in a loop)
                    std::unique_lock<std::mutex> lk(mu);
                    fail_waiting(*e, ex.code, "no device memory for the solve engine of operator " +
                                                  std::to_string(e->id) + ": " + ex.msg);
                    continue;
                }
                std::unique_lock<std::mutex> lk(mu);
                e->engine_bytes = eng->bytes;
                used += eng->bytes;
                st.peak_bytes = std::max(st.peak_bytes, used);
                e->engine = std::move(eng);
            }
            if (!need_engine.empty()) continue;  // admit into the new slots first
            const double t = now();
            tick_t0 = t;
            for (auto &a : admits) {
                OpEntry &e = *std::get<0>(a);
                auto &req = std::get<2>(a);
                const int rc = std::get<3>(a);
                std::vector<double> bt(req->n);
                const auto &perm = e.tree->perm;
                for (int64_t q = 0; q < req->n; q++) bt[q] = req->b[(size_t)rc * req->n + perm[q]];
                e.engine->admit(std::get<1>(a), req, rc, bt, t);
                std::unique_lock<std::mutex> lk(mu);
                if (req->t_start < 0) req->t_start = t;
                if (req->status >= 0) req->status = 1;
            }
            const std::vector<OpEntry *> &run = active;
            // phase A on every engine, then collect
            for (OpEntry *e : run) e->engine->launch_a();
            std::vector<std::vector<int>> done(run.size());
            for (size_t q = 0; q < run.size(); q++) run[q]->engine->finish_a(done[q]);
            for (size_t q = 0; q < run.size(); q++) {
                for (int c : done[q]) {
                    std::vector<double> xt;
                    run[q]->engine->read_x(c, xt);
                    std::unique_lock<std::mutex> lk(mu);
                    complete_column(*run[q], c, xt);
                }
            }
            {
                std::unique_lock<std::mutex> lk(mu);
                for (OpEntry *e : run) fail_columns(*e);
            }
            // phase B (Arnoldi step) and C (next vectors / updates).  The
            // phases that allocate basis vectors fail only the requests of
            // an engine whose allocation failed
            auto guarded = [&](OpEntry *e, void (Engine::*phase)()) {
                try {
                    (e->engine.get()->*phase)();
                } catch (const hodlr::ArgError &ex) {
                    std::unique_lock<std::mutex> lk(mu);
                    std::vector<std::shared_ptr<Request>> inflight;
                    for (auto &C : e->engine->cols)
                        if (C.phase) inflight.push_back(C.req);
                    for (auto &r : inflight) fail_request(*e, r, ex.code, ex.msg);
                }
            };
            for (OpEntry *e : run) guarded(e, &Engine::launch_b);
            for (OpEntry *e : run) e->engine->finish_b();
            {
                std::unique_lock<std::mutex> lk(mu);
                for (OpEntry *e : run) fail_columns(*e);
            }
            for (OpEntry *e : run) guarded(e, &Engine::launch_c);
            std::unique_lock<std::mutex> lk(mu);
            st.ticks++;
            st.tick_seconds += now() - tick_t0;
            for (OpEntry *e : run) e->in_tick = false;
            for (OpEntry *e : run) {  // basis vectors allocated this tick
                used += e->engine->bytes - e->engine_bytes;
                e->engine_bytes = e->engine->bytes;
            }
            st.peak_bytes = std::max(st.peak_bytes, used);
        } catch (const hodlr::HipError &ex) {
            std::unique_lock<std::mutex> lk(mu);
            for (OpEntry *e : active) e->in_tick = false;
            last_error = std::string(hipGetErrorString(ex.code)) + " in " + ex.where;
            for (auto &kv : reqs)
                if (kv.second->status >= 0 && kv.second->status < 2) kv.second->status = HODLR_E_DEVICE;
            stop = true;
            cv_done.notify_all();
            return;
        } catch (const hodlr::ArgError &ex) {
            // an unexpected argument error inside a tick: the requests in
            // flight on the engines of this tick fail (each once); columns
            // that have not started keep waiting
            std::unique_lock<std::mutex> lk(mu);
            for (OpEntry *e : active) e->in_tick = false;
            for (OpEntry *e : active) {
                std::vector<std::shared_ptr<Request>> inflight;
                for (auto &C : e->engine->cols)
                    if (C.phase) inflight.push_back(C.req);
                for (auto &r : inflight) fail_request(*e, r, ex.code, ex.msg);
            }
        }
    }
}

namespace {

template <class F>
int sguard(hodlr_server *srv, F &&f) {
    try {
        f();
        return HODLR_OK;
    } catch (const hodlr::HipError &e) {
        if (srv) srv->last_error = std::string(hipGetErrorString(e.code)) + " in " + e.where;
        return HODLR_E_DEVICE;
    } catch (const hodlr::ArgError &e) {
        if (srv) srv->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (srv) srv->last_error = "host allocation failed";
        return HODLR_E_NOMEM;
    } catch (const std::exception &e) {
        if (srv) srv->last_error = e.what();
        return HODLR_E_STATE;
    }
}

}  // namespace

extern "C" {

int hodlr_server_create(int device, const hodlr_server_config *cfg, hodlr_server **out) {
    if (!cfg || !out || device < 0) return HODLR_E_ARG;
    *out = nullptr;
    const int sl = cfg->slots;
    if (sl != 1 && sl != 2 && sl != 4 && sl != 8 && sl != 16 && sl != 32) return HODLR_E_ARG;
    if (cfg->restart < 1 || cfg->restart > 127 || !(cfg->tol > 0) || !(cfg->eps_aca > 0 && cfg->eps_aca < 1) ||
        cfg->nmax < 1 || cfg->max_outer < 1 || cfg->builders < 1 || cfg->builders > 16)
        return HODLR_E_ARG;
    auto *srv = new hodlr_server;
    int rc = sguard(srv, [&] {
        HIP_CHECK(hipSetDevice(device));
        srv->cfg = *cfg;
        srv->device = device;
        for (int b = 0; b < cfg->builders; b++) {
            srv->bctx.push_back(std::make_unique<hodlr_ctx>());
            srv->bctx.back()->device = device;
            HIP_CHECK(hipStreamCreateWithFlags(&srv->bctx.back()->stream, hipStreamNonBlocking));
        }
        hodlr::keep_pool_memory(device);
        size_t fr = 0, tot = 0;
        HIP_CHECK(hipMemGetInfo(&fr, &tot));
        srv->st.budget_bytes = cfg->hbm_budget > 0 ? cfg->hbm_budget : 0.9 * (double)fr;
        // one slab for operators, engines and build scratch: the budget plus
        // room for concurrent builds (allocations that do not fit fall back
        // to the driver)
        const double room = std::max(16.0 * (1 << 30), 0.05 * (double)fr);
        const double sb = std::min(0.95 * (double)fr, srv->st.budget_bytes + room);
        if (sb >= (double)(1 << 30)) {
            try {
                srv->slab = std::make_unique<hodlr::DevSlab>((size_t)sb);
            } catch (const hodlr::ArgError &) {  // no slab: every buffer comes from the driver
                srv->slab.reset();
            }
        }
        for (int b = 0; b < cfg->builders; b++) srv->builders.emplace_back([srv, b] { srv->build_loop(b); });
        srv->scheduler = std::thread([srv] { srv->sched_loop(); });
    });
    if (rc != HODLR_OK) {
        delete srv;
        return rc;
    }
    *out = srv;
    return HODLR_OK;
}

void hodlr_server_destroy(hodlr_server *srv) {
    if (!srv) return;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->stop = true;
        for (auto &kv : srv->ops)  // queued builds are abandoned
            if (kv.second->state == 0) srv->fail_op(*kv.second, HODLR_E_STATE, "server stopped");
    }
    srv->cv_build.notify_all();
    srv->cv_work.notify_all();
    srv->cv_mem.notify_all();
    for (auto &t : srv->builders)
        if (t.joinable()) t.join();
    if (srv->scheduler.joinable()) srv->scheduler.join();
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        for (auto &kv : srv->ops) srv->free_op(*kv.second);
    }
    for (auto &c : srv->bctx)
        if (c->stream) (void)hipStreamDestroy(c->stream);
    delete srv;
}

int hodlr_server_submit_build(hodlr_server *srv, int ndim, int64_t n, const double *pts, const double *lo,
                              const double *hi, const hodlr_kernel_desc *k, int64_t *op_id) {
    if (!srv || !pts || !lo || !hi || !k || !op_id || n < 1 || ndim < 1 || ndim > hodlr::kMaxDim) return HODLR_E_ARG;
    if (k->ndim != ndim || k->family < 0 || k->family > 2) return HODLR_E_ARG;
    return sguard(srv, [&] {
        auto e = std::make_unique<OpEntry>();
        e->ndim = ndim;
        e->n = n;
        e->pts.assign(pts, pts + n * ndim);
        e->lo.assign(lo, lo + ndim);
        e->hi.assign(hi, hi + ndim);
        if (k->weight) e->weight.assign(k->weight, k->weight + n);
        e->kd = *k;
        e->kd.weight = nullptr;
        std::unique_lock<std::mutex> lk(srv->mu);
        if (srv->stop) throw hodlr::ArgError{HODLR_E_STATE, "server stopped"};
        e->id = srv->next_op++;
        *op_id = e->id;
        srv->ops[e->id] = std::move(e);
        srv->cv_build.notify_all();
    });
}

int hodlr_server_submit_solve(hodlr_server *srv, int64_t op_id, int nrhs, const double *b, int64_t *req_id) {
    if (!srv || !b || !req_id || nrhs < 1) return HODLR_E_ARG;
    return sguard(srv, [&] {
        std::unique_lock<std::mutex> lk(srv->mu);
        if (srv->stop) throw hodlr::ArgError{HODLR_E_STATE, "server stopped"};
        auto it = srv->ops.find(op_id);
        if (it == srv->ops.end()) throw hodlr::ArgError{HODLR_E_ARG, "unknown operator id"};
        OpEntry &e = *it->second;
        if (e.state < 0) throw hodlr::ArgError{e.state, "operator failed to build: " + e.error};
        auto r = std::make_shared<Request>();
        r->id = srv->next_req++;
        r->op_id = op_id;
        r->nrhs = nrhs;
        r->n = e.n;
        r->b.assign(b, b + (size_t)nrhs * e.n);
        r->x.assign((size_t)nrhs * e.n, 0.0);
        r->infos.assign(nrhs, hodlr_gmres_info{});
        r->t_submit = srv->now();
        r->built = e.state != 2;
        if (e.state == 3) {  // evicted: build again from the kept points
            e.state = 0;
            srv->cv_build.notify_all();
        }
        e.refs++;
        for (int c = 0; c < nrhs; c++) e.waiting.emplace_back(r, c);
        srv->reqs[r->id] = r;
        *req_id = r->id;
        srv->cv_work.notify_all();
    });
}

int hodlr_server_wait(hodlr_server *srv, int64_t req_id, double timeout_s, double *x, hodlr_request_info *info) {
    if (!srv) return HODLR_E_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(req_id);
    if (it == srv->reqs.end()) {
        srv->last_error = "unknown request id";
        return HODLR_E_ARG;
    }
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->status == 2 || r->status < 0; };
    if (timeout_s < 0) srv->cv_done.wait(lk, ready);
    else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return HODLR_E_STATE;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->op_id = r->op_id;
        info->nrhs = r->nrhs;
        info->status = r->status;
        info->built = r->built;
        info->t_submit = r->t_submit;
        info->t_start = r->t_start;
        info->t_done = r->t_done;
        info->converged = r->status == 2;
        for (auto &in : r->infos) {
            info->converged = info->converged && in.converged;
            info->iterations = std::max(info->iterations, in.iterations);
            info->outer = std::max(info->outer, in.outer);
            info->max_resid = std::max(info->max_resid, in.resid);
        }
    }
    if (r->status < 0) {
        srv->last_error = "request failed";
        return r->status;
    }
    if (x) std::memcpy(x, r->x.data(), sizeof(double) * r->x.size());
    return HODLR_OK;
}

int hodlr_server_get_stats(hodlr_server *srv, hodlr_server_stats *out) {
    if (!srv || !out) return HODLR_E_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    *out = srv->st;
    out->used_bytes = srv->used;
    out->resident_ops = 0;
    for (auto &kv : srv->ops) out->resident_ops += kv.second->state == 2;
    return HODLR_OK;
}

const char *hodlr_server_last_error(hodlr_server *srv) { return srv ? srv->last_error.c_str() : "null server"; }

}  // extern "C"
