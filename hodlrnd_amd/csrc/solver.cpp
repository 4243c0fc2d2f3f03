// solver.cpp -- restarted GMRES on the device (GMRES.hpp:160-303) and the
// FP32-inner / FP64-refined solve.
//
// Control flow follows iterSolver::GMRES: residual check `<= tol` before the
// first cycle, `< tol` after every Arnoldi step and after every restart,
// x updated by the back-solve of the rotated Hessenberg system
// (GMRES.hpp:132-146), real Givens rotations (GMRES.hpp:83-129).  Vectors,
// the Hessenberg matrix, the rotations and the residual test all stay on
// the device (blas.hip); per Arnoldi step the host reads a 56-byte state
// record to learn whether the solve goes on.
// A zero subdiagonal (lucky breakdown) leaves the next basis vector zero
// instead of dividing by zero (the reference has no guard, GMRES.hpp:242).
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

}  // namespace

Workspace::Workspace(int64_t n_, int restart_) : n(n_), restart(restart_) {
    V.alloc(sizeof(double) * (size_t)n * (restart + 1));
    w.alloc(sizeof(double) * n);
    r.alloc(sizeof(double) * n);
    partial.alloc(sizeof(double) * (size_t)mdot_chunks(n) * 130);
    small.alloc(sizeof(double) * 512);
    const size_t ld = (size_t)restart + 1;
    gh.alloc(sizeof(double) * (ld * restart + 4 * ld) + sizeof(GmresDevState) + 64);
    HIP_CHECK(hipHostMalloc(&hbuf, sizeof(double) * 512));
}

Workspace::~Workspace() {
    if (hbuf) (void)hipHostFree(hbuf);
}

double Workspace::norm2(const double *v, hipStream_t s) {
    HIP_CHECK(launch_mdot(v, n, 1, v, n, partial.as<double>(), small.as<double>(), 0, s));
    HIP_CHECK(hipMemcpyAsync(hbuf, small.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return std::sqrt(hbuf[0]);
}

int gmres_device(const Apply &A, Workspace &ws, const double *b, double *x, int max_iter, double tol,
                 bool x0_zero, hipStream_t s, GmresResult &res, std::vector<double> *hist, bool mgs) {
    const int64_t n = ws.n;
    const int m = ws.restart;
    const int ld = m + 1;
    double *V = ws.V.as<double>();
    double *w = ws.w.as<double>();
    double *r = ws.r.as<double>();
    double *hd = ws.small.as<double>();  // [0,128) h1, [128,256) h2, [256] |w|^2, [300] |b|^2, [301] |r|^2
    double *H = ws.gh.as<double>();
    double *cs = H + (size_t)ld * m, *sn = cs + ld, *sv = sn + ld, *y = sv + ld;
    GmresDevState *st = reinterpret_cast<GmresDevState *>(y + ld);
    // pinned record of the state after each step
    GmresDevState *slot = reinterpret_cast<GmresDevState *>(ws.hbuf + 400);

    HIP_CHECK(launch_mdot(b, n, 1, b, n, ws.partial.as<double>(), hd + 300, 0, s));
    HIP_CHECK(launch_gmres_init(hd + 300, tol, max_iter, st, s));
    if (x0_zero) {
        HIP_CHECK(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    } else {
        A(x, w);
        res.matvecs++;
        HIP_CHECK(launch_sub(b, w, r, n, s));
    }
    auto cycle_start = [&](int first) {  // beta, the residual test, sv = beta e_1 (GMRES.hpp:186-213,276-291)
        HIP_CHECK(launch_mdot(r, n, 1, r, n, ws.partial.as<double>(), hd + 301, 0, s));
        HIP_CHECK(launch_gmres_cycle(hd + 301, sv, ld, first, st, s));
        HIP_CHECK(hipMemcpyAsync(slot, st, sizeof(GmresDevState), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        return slot[0];
    };
    GmresDevState cur = cycle_start(1);
    res.resid = cur.resid;
    if (!cur.active) {  // GMRES.hpp:198-213 (or no iteration budget at all)
        res.iterations = 0;
        res.converged = cur.conv == -3;
        return 0;
    }
    auto launch_step = [&](int i) {
        double *vi = V + (size_t)i * n;
        A(vi, w);
        res.matvecs++;
        if (mgs) {
            // MGS (GMRES.hpp:236-240): h_k = <w, v_k>, w -= h_k v_k for k =
            // 0..i in turn, each coefficient read by the update on the
            // device; h2 = 0
            for (int k = 0; k <= i; k++) {
                const double *vk = V + (size_t)k * n;
                HIP_CHECK(launch_mdot(vk, n, 1, w, n, ws.partial.as<double>(), hd + k, 0, s));
                HIP_CHECK(launch_gemv(vk, n, 1, hd + k, -1.0, w, n, s));
            }
            HIP_CHECK(hipMemsetAsync(hd + 128, 0, sizeof(double) * (size_t)(i + 1), s));
        } else {
            // CGS2: h1 = V^T w, w -= V h1, h2 = V^T w, w -= V h2
            HIP_CHECK(launch_mdot(V, n, i + 1, w, n, ws.partial.as<double>(), hd, 0, s));
            HIP_CHECK(launch_gemv(V, n, i + 1, hd, -1.0, w, n, s));
            HIP_CHECK(launch_mdot(V, n, i + 1, w, n, ws.partial.as<double>(), hd + 128, 0, s));
            HIP_CHECK(launch_gemv(V, n, i + 1, hd + 128, -1.0, w, n, s));
        }
        HIP_CHECK(launch_mdot(w, n, 1, w, n, ws.partial.as<double>(), hd + 256, 0, s));
        HIP_CHECK(launch_gmres_step(H, cs, sn, sv, hd, i, ld, st, s));
        HIP_CHECK(launch_scale_by(w, &st->scale, V + (size_t)(i + 1) * n, n, s));
        HIP_CHECK(hipMemcpyAsync(slot, st, sizeof(GmresDevState), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    };
    // (Launching step i + 1 before step i's record was read measured slower
    // on small problems: the step launched after the end costs more than
    // the host wait it hides.)
    double resid = cur.resid;
    while (true) {
        HIP_CHECK(launch_scale_by(r, &st->scale, V, n, s));  // v_0 = r / beta
        int i = 0;
        while (true) {
            launch_step(i);
            cur = slot[0];
            if (hist) hist->push_back(resid);
            resid = cur.resid;
            res.resid = resid;
            if (!cur.active || i + 1 == m) break;
            i++;
        }
        if (cur.conv == -2) throw ArgError{HODLR_E_BREAKDOWN, "non-finite Arnoldi norm"};
        const int k = cur.active ? m - 1 : cur.conv;
        HIP_CHECK(launch_gmres_backsolve(H, sv, ld, k, y, s));
        HIP_CHECK(launch_gemv(V, n, k + 1, y, 1.0, x, n, s));
        if (!cur.active && cur.resid < tol) {  // GMRES.hpp:253-268
            HIP_CHECK(hipStreamSynchronize(s));
            res.iterations = cur.steps;
            res.converged = true;
            return cur.steps;
        }
        // end of the cycle or of the iteration budget: restart from the true
        // residual (GMRES.hpp:271-291); the reference's j has moved one past
        // the last step here
        A(x, w);
        res.matvecs++;
        HIP_CHECK(launch_sub(b, w, r, n, s));
        cur = cycle_start(0);
        resid = cur.resid;
        res.resid = resid;
        if (cur.conv == -3) {
            res.iterations = cur.steps + 1;
            res.converged = true;
            return cur.steps + 1;
        }
        if (!cur.active) {  // max_iter reached (GMRES.hpp:298-302)
            res.iterations = max_iter;
            res.converged = false;
            return max_iter;
        }
    }
}

struct SolveState {
    Workspace ws;
    DevBuf x32, y32, d, rr, w64;
    SolveState(int64_t n, int restart) : ws(n, restart) {
        x32.alloc(sizeof(float) * n);
        y32.alloc(sizeof(float) * n);
        d.alloc(sizeof(double) * n);
        rr.alloc(sizeof(double) * n);
        w64.alloc(sizeof(double) * n);
    }
};

void solve(hodlr_op &op, const double *b, double *x, const hodlr_gmres_params &p, hodlr_gmres_info &info) {
    const double t0 = now_s();
    hipStream_t s = op.ctx->stream;
    const int64_t n = op.N;
    const int restart = p.restart > 0 ? p.restart : 100;
    if (restart + 1 > 128) throw ArgError{HODLR_E_ARG, "restart must be at most 127"};
    const int max_iter = p.max_iter > 0 ? p.max_iter : (int)std::min<int64_t>(n, INT32_MAX);
    auto cached = std::static_pointer_cast<SolveState>(op.solver_ws);
    if (!cached || cached->ws.restart != restart) {
        op.solver_ws.reset();
        cached = std::make_shared<SolveState>(n, restart);
        op.solver_ws = cached;
    }
    Workspace &ws = cached->ws;
    std::memset(&info, 0, sizeof(info));
    std::vector<double> hist;
    if (p.x0_is_b == 1) HIP_CHECK(hipMemcpyAsync(x, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    const bool x0_zero = p.x0_is_b == 2;
    if (x0_zero) HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * n, s));

    if (p.ortho != HODLR_ORTHO_CGS2 && p.ortho != HODLR_ORTHO_MGS) throw ArgError{HODLR_E_ARG, "bad ortho"};
    const bool mgs = p.ortho == HODLR_ORTHO_MGS;
    Apply A64 = [&](const double *in, double *out) { op.matvec_dev(HODLR_PREC_F64, 1, in, out, s); };
    GmresResult gr;
    if (p.mode == HODLR_SOLVE_F64) {
        gmres_device(A64, ws, b, x, max_iter, p.tol, x0_zero, s, gr, &hist, mgs);
        info.iterations = gr.iterations;
        info.matvecs_f64 = gr.matvecs;
        info.converged = gr.converged;
        info.resid = gr.resid;
        info.true_resid = -1.0;
        if (p.verify) {
            A64(x, ws.w.as<double>());
            HIP_CHECK(launch_sub(b, ws.w.as<double>(), ws.r.as<double>(), n, s));
            double nb = ws.norm2(b, s);
            info.true_resid = ws.norm2(ws.r.as<double>(), s) / (nb == 0.0 ? 1.0 : nb);
        }
    } else {
        if (!(op.prec_flags & HODLR_PREC_F32)) throw ArgError{HODLR_E_STATE, "operator has no FP32 copy"};
        op.ensure_work(HODLR_PREC_F32, 1);
        DevBuf &x32 = cached->x32, &y32 = cached->y32, &d = cached->d, &rr = cached->rr, &w64 = cached->w64;
        // FP32 off-diagonal operator, diagonal applied in FP64
        Apply A32 = [&](const double *in, double *out) {
            HIP_CHECK(launch_to_f32(in, x32.as<float>(), n, s));
            op.matvec_dev(HODLR_PREC_F32, 1, x32.as<float>(), y32.as<float>(), s, nullptr, false);
            HIP_CHECK(launch_f32_result(y32.as<float>(), in, op.diag.as<double>(), out, n, 1, s));
        };
        const double nb0 = ws.norm2(b, s);
        const double nb = nb0 == 0.0 ? 1.0 : nb0;
        // The first correction is driven by r0 = b - A32 x0: its error
        // (~ FP32 rounding of the operator) is removed by the next sweep,
        // which, like every later one, starts from an FP64 residual.  The
        // loop always ends on an FP64 residual, so the reported residual is
        // certified against the FP64 operator.
        if (x0_zero) {  // r0 = b exactly
            HIP_CHECK(hipMemcpyAsync(rr.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        } else {
            A32(x, w64.as<double>());
            info.matvecs_f32++;
            HIP_CHECK(launch_sub(b, w64.as<double>(), rr.as<double>(), n, s));
        }
        double res = ws.norm2(rr.as<double>(), s) / nb;
        bool res_is_f64 = x0_zero;
        hist.push_back(res);
        const int max_outer = p.max_outer > 0 ? p.max_outer : 20;
        const double inner_floor = p.inner_tol > 0 ? p.inner_tol : 1e-14;
        int total = 0;
        while (!(res_is_f64 && res < p.tol) && info.outer < max_outer && total < max_iter) {
            GmresResult g;
            g.matvecs = 0;
            HIP_CHECK(hipMemsetAsync(d.p, 0, sizeof(double) * n, s));
            // aim the sweep at the outer tolerance: relative to ||r||, the inner
            // solve must reach 0.3 tol ||b|| / ||r|| (never looser than 0.5,
            // never tighter than the caller's floor inner_tol)
            const double target = std::min(0.5, std::max(inner_floor, 0.3 * p.tol / std::max(res, 1e-300)));
            gmres_device(A32, ws, rr.as<double>(), d.as<double>(), max_iter - total, target, true, s, g,
                         nullptr, mgs);
            total += g.iterations;
            info.matvecs_f32 += g.matvecs;
            HIP_CHECK(launch_add(x, d.as<double>(), n, s));
            A64(x, w64.as<double>());
            info.matvecs_f64++;
            HIP_CHECK(launch_sub(b, w64.as<double>(), rr.as<double>(), n, s));
            res = ws.norm2(rr.as<double>(), s) / nb;
            res_is_f64 = true;
            hist.push_back(res);
            info.outer++;
            if (g.iterations == 0) break;  // no progress possible at this precision
        }
        if (!res_is_f64) {  // converged before any sweep: certify in FP64
            A64(x, w64.as<double>());
            info.matvecs_f64++;
            HIP_CHECK(launch_sub(b, w64.as<double>(), rr.as<double>(), n, s));
            res = ws.norm2(rr.as<double>(), s) / nb;
        }
        info.iterations = total;
        info.converged = res < p.tol;
        info.resid = res;
        info.true_resid = res;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    if (p.hist && p.hist_cap > 0) {
        const int nh = (int)std::min<size_t>(hist.size(), (size_t)p.hist_cap);
        std::memcpy(p.hist, hist.data(), sizeof(double) * nh);
        info.hist_len = nh;
    }
    info.seconds = now_s() - t0;
}

}  // namespace hodlr
