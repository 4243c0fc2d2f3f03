// solver.h -- device GMRES (GMRES.hpp:160-303) and mixed-precision refinement.
#pragma once

#include <functional>
#include <vector>

#include "internal.h"

namespace hodlr {

// y = A x on device vectors (tree order)
using Apply = std::function<void(const double *, double *)>;

struct Workspace {
    int64_t n;
    int restart;
    DevBuf V, w, r, partial, small;
    // device GMRES state: Hessenberg matrix (restart+1) x restart, cs, sn,
    // the rotated right-hand side and the back-solved y (restart+1 each),
    // then the solve's GmresDevState
    DevBuf gh;
    double *hbuf = nullptr;  // pinned host staging
    Workspace(int64_t n, int restart);
    ~Workspace();
    double norm2(const double *v, hipStream_t s);
};

struct GmresResult {
    int iterations = 0;
    int matvecs = 0;
    bool converged = false;
    double resid = 0.0;
};

// mgs: modified Gram-Schmidt (GMRES.hpp:236-240) instead of CGS2
int gmres_device(const Apply &A, Workspace &ws, const double *b, double *x, int max_iter, double tol,
                 bool x0_zero, hipStream_t s, GmresResult &res, std::vector<double> *hist, bool mgs = false);

void solve(struct ::hodlr_op &op, const double *b, double *x, const hodlr_gmres_params &p,
           hodlr_gmres_info &info);

// every (problem, column) of a batch operator solved independently;
// b, x: n x K interleaved device arrays; infos: nprob * K entries
void solve_batch(struct ::hodlr_op &op, int K, const double *b, double *x, const hodlr_gmres_params &p,
                 hodlr_gmres_info *infos);

}  // namespace hodlr
