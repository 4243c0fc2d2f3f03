// HODLRnD.hpp -- the reference's C++ operator API over the gfx950 C ABI.
//
// Same shapes as kandapva/HODLRnD so a caller can switch by changing the
// include:
//   knobs NDIM, Nmax, eps_ACA, INTERACTION_TYPE_ALLOWED,  (myHeaders.hpp:35-49)
//         SYS_SIZE
//   ptsnD, uniform_nodes, Vec_ops::relative_error         (points_dt.hpp:9-13,
//                                                          myHeaders.hpp:111-136)
//   user kernels with getMatrixEntry / get_points          (user_kernel.hpp:10-179)
//   kernel_function<K>::getRow / getCol / getMatrix        (kernel_function.hpp:20-53)
//   HODLRdD_matrix<K>(ker, pts, x1, x2), Assemble_matrix_operators(),
//     operator*, solve, get_size, print_matrix_statistics  (HDD_matrix.hpp:20-85)
//   iterSolver<HODLRdD_matrix<K>>(restart, max_iter, tol).GMRES(A, x, b)
//                                                          (GMRES.hpp:44-50,160)
// The knobs are compile-time constants as in the reference; define
// HODLR_NDIM, HODLR_NMAX, HODLR_EPS_ACA or HODLR_DPRIME before including to
// change them.  A user kernel class must also provide
//   hodlr_kernel_desc device_kernel() const
// describing A_ij for evaluation on the GPU (the reference evaluates
// getMatrixEntry on the host; the device has no access to host code).
// Vectors are plain host arrays (the reference uses Eigen, which this API
// does not require).  Failures throw hodlr::Error.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hodlrnd.h"

#ifndef HODLR_NDIM
#define HODLR_NDIM 3
#endif
#ifndef HODLR_NMAX
#define HODLR_NMAX 64
#endif
#ifndef HODLR_EPS_ACA
#define HODLR_EPS_ACA 1e-10
#endif
#ifndef HODLR_DPRIME
#define HODLR_DPRIME 0
#endif
#ifndef HODLR_SYS_SIZE
#define HODLR_SYS_SIZE (2 * HODLR_NMAX)
#endif

// ---- knobs (myHeaders.hpp:35-49) ----
constexpr int NDIM = HODLR_NDIM;
constexpr int Nmax = HODLR_NMAX;
constexpr int INTERACTION_TYPE_ALLOWED = HODLR_DPRIME;  // d'
const double eps_ACA = HODLR_EPS_ACA;
// blocks with more columns keep only their skeleton and are evaluated on
// the fly (LowRank_matrix.hpp:22-36); define HODLR_SYS_SIZE -1 to store
// every factor
constexpr int SYS_SIZE = HODLR_SYS_SIZE;
inline int numPoints = 10;
inline size_t N = 1;
const double PI = 4.0 * std::atan(1.0);

using dtype = double;
using dtype_base = double;

namespace hodlr {
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
}  // namespace hodlr

// Minimal dense vector with the members the reference's callers use.
class Vec {
    std::vector<double> d_;

public:
    Vec() = default;
    explicit Vec(size_t n, double v = 0.0) : d_(n, v) {}
    static Vec Zero(size_t n, size_t = 1) { return Vec(n, 0.0); }
    // Eigen's VectorXd::Random: uniform in [-1, 1] from rand()
    static Vec Random(size_t n, size_t = 1) {
        Vec v(n);
        for (size_t i = 0; i < n; i++) v.d_[i] = -1.0 + 2.0 * (double)std::rand() / (double)RAND_MAX;
        return v;
    }
    size_t size() const { return d_.size(); }
    double &operator()(size_t i) { return d_[i]; }
    double operator()(size_t i) const { return d_[i]; }
    double &operator[](size_t i) { return d_[i]; }
    double operator[](size_t i) const { return d_[i]; }
    double *data() { return d_.data(); }
    const double *data() const { return d_.data(); }
    double dot(const Vec &o) const {
        double s = 0.0;
        for (size_t i = 0; i < d_.size(); i++) s += d_[i] * o.d_[i];
        return s;
    }
    double norm() const { return std::sqrt(dot(*this)); }
    Vec operator-(const Vec &o) const {
        Vec r(*this);
        for (size_t i = 0; i < d_.size(); i++) r.d_[i] -= o.d_[i];
        return r;
    }
};

// ---- points (points_dt.hpp:9-39) ----
struct ptsnD {
    double x[NDIM];
    size_t id = (size_t)-1;
};

namespace nd_points {
inline double euclidean_distance(const ptsnD &a, const ptsnD &b) {
    double dist = 0.0;
    for (int i = 0; i < NDIM; i++) dist += (a.x[i] - b.x[i]) * (a.x[i] - b.x[i]);
    return std::sqrt(dist);
}
inline double max_norm_distance(const ptsnD &a, const ptsnD &b) {
    double dist = std::fabs(a.x[0] - b.x[0]);
    for (int i = 0; i < NDIM; i++) dist = std::fabs(a.x[i] - b.x[i]) > dist ? std::fabs(a.x[i] - b.x[i]) : dist;
    return dist;
}
}  // namespace nd_points

// myHeaders.hpp:111-120
inline Vec uniform_nodes(double a, double b, int n) {
    Vec X(n);
    double l = (b - a);
    for (int k = 0; k < n; k++) X(k) = a + l / (double)(n - 1) * k;
    return X;
}

namespace Vec_ops {
inline double relative_error(const Vec &X_ori, const Vec &X_comp) { return (X_ori - X_comp).norm() / X_ori.norm(); }
}  // namespace Vec_ops

// ---- test kernels (user_kernel.hpp:10-179): tensor grid, dim 0 outermost ----
namespace hodlr_detail {
inline std::vector<ptsnD> *tensor_grid(double a, double b) {
    auto *g = new std::vector<ptsnD>;
    Vec nodes = uniform_nodes(a, b, numPoints);
    size_t total = 1;
    for (int d = 0; d < NDIM; d++) total *= (size_t)numPoints;
    g->reserve(total);
    for (size_t p = 0; p < total; p++) {
        ptsnD t;
        size_t rem = p;
        for (int d = NDIM - 1; d >= 0; d--) {
            t.x[d] = nodes((size_t)(rem % numPoints));
            rem /= numPoints;
        }
        t.id = p;
        g->push_back(t);
    }
    return g;
}
// Gauss-Legendre rule on [-1, 1] by Newton iteration on P_n
inline void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; i++) {
        long double z = std::cos(3.141592653589793238L * (i + 0.75L) / (n + 0.5L)), dp = 0;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            long double dz = p1 / dp;
            z -= dz;
            if (std::fabs((double)dz) < 1e-19) break;
        }
        x[i] = -(double)z;
        x[n - 1 - i] = (double)z;
        w[i] = w[n - 1 - i] = (double)(2.0L / ((1 - z * z) * dp * dp));
    }
}
// integral4d.hpp:18-62: -exp(-|x|)/(4 pi) over a 4-D box by a 32^4 rule
inline double quadruple_integral(const double *a, const double *b) {
    std::vector<double> x, w;
    gauss_legendre(32, x, w);
    const double c = -1.0 / (4 * PI);
    double cx[4], L[4];
    for (int d = 0; d < 4; d++) {
        cx[d] = 0.5 * (b[d] - a[d]);
        L[d] = 0.5 * (b[d] + a[d]);
    }
    double result = 0.0;
    for (int i = 0; i < 32; i++)
        for (int j = 0; j < 32; j++)
            for (int k = 0; k < 32; k++)
                for (int l = 0; l < 32; l++) {
                    double t0 = cx[0] * x[i] + L[0], t1 = cx[1] * x[j] + L[1];
                    double t2 = cx[2] * x[k] + L[2], t3 = cx[3] * x[l] + L[3];
                    result += w[i] * w[j] * w[k] * w[l] * c * std::exp(-std::sqrt(t0 * t0 + t1 * t1 + t2 * t2 + t3 * t3));
                }
    return result * cx[0] * cx[1] * cx[2] * cx[3];
}
}  // namespace hodlr_detail

class kernel_2d_test {
    double c = -1.0 / (2 * PI);

public:
    std::vector<ptsnD> *gridPoints;
    kernel_2d_test() { gridPoints = hodlr_detail::tensor_grid(-1, 1); }
    void get_points(std::vector<ptsnD> *&src) { src = gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return c * std::log(x); }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return 0.0;
        return Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
    hodlr_kernel_desc device_kernel() const { return {HODLR_KERNEL_LOG, NDIM, c, 0.0, nullptr}; }
};

class kernel_3d_test {
    double c = 1.0 / (4 * PI);

public:
    std::vector<ptsnD> *gridPoints;
    kernel_3d_test() { gridPoints = hodlr_detail::tensor_grid(-1, 1); }
    void get_points(std::vector<ptsnD> *&src) { src = gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return c / x; }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return 0.0;
        return Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
    hodlr_kernel_desc device_kernel() const { return {HODLR_KERNEL_INV_R, NDIM, c, 0.0, nullptr}; }
};

class kernel_4d_test {
    double c = -1.0 / (4 * PI);
    double h = 1.0 / numPoints;
    double h4 = std::pow(h, 4.0);
    double kii = 0.0;

public:
    std::vector<ptsnD> *gridPoints;
    kernel_4d_test() {
        double a[4] = {0, 0, 0, 0}, b[4] = {h * 0.5, h * 0.5, h * 0.5, h * 0.5};
        kii = 1.0 + hodlr_detail::quadruple_integral(a, b);  // second kind
        gridPoints = hodlr_detail::tensor_grid(-1, 1);
    }
    void get_points(std::vector<ptsnD> *&src) { src = gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return c * std::exp(-x); }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return kii;
        return h4 * Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
    hodlr_kernel_desc device_kernel() const { return {HODLR_KERNEL_EXP, NDIM, h4 * c, kii, nullptr}; }
};

// ---- kernel_function (kernel_function.hpp:11-68): host access to entries ----
template <class Operator>
class kernel_function {
    Operator *kernel;

public:
    explicit kernel_function(Operator *&A) : kernel(A) {}
    Vec getRow(int j, std::vector<size_t> &targets) {
        Vec row(targets.size());
        for (size_t k = 0; k < targets.size(); k++) row(k) = kernel->getMatrixEntry(j, (int)targets[k]);
        return row;
    }
    Vec getCol(int k, std::vector<size_t> &sources) {
        Vec col(sources.size());
        for (size_t j = 0; j < sources.size(); j++) col(j) = kernel->getMatrixEntry((int)sources[j], k);
        return col;
    }
    std::vector<double> getMatrix(std::vector<size_t> &sources, std::vector<size_t> &targets) {
        std::vector<double> m(sources.size() * targets.size());  // column-major
        for (size_t k = 0; k < targets.size(); k++)
            for (size_t j = 0; j < sources.size(); j++)
                m[k * sources.size() + j] = kernel->getMatrixEntry((int)sources[j], (int)targets[k]);
        return m;
    }
    Operator *get() { return kernel; }
};

template <class Op>
class iterSolver;

// ---- the operator (HDD_matrix.hpp:10-89) ----
template <class Kernel>
class HODLRdD_matrix {
    size_t N_ = 0;
    hodlr_ctx *ctx_ = nullptr;
    hodlr_tree *tree_ = nullptr;
    hodlr_op *op_ = nullptr;
    kernel_function<Kernel> *kernel_func;
    Kernel *kernel_;

    void check(int rc, const char *what) {
        if (rc != HODLR_OK && rc != HODLR_E_NOCONV)
            throw hodlr::Error(rc, std::string(what) + ": " + hodlr_last_error(ctx_));
    }
    friend class iterSolver<HODLRdD_matrix<Kernel>>;

public:
    HODLRdD_matrix(Kernel *&userkernel, std::vector<ptsnD> *&gPoints, const Vec &x1, const Vec &x2, int device = 0) {
        std::printf("HMATRIX START\n");
        kernel_ = userkernel;
        kernel_func = new kernel_function<Kernel>(userkernel);
        N_ = gPoints->size();
        std::vector<double> pts(N_ * NDIM);
        for (size_t i = 0; i < N_; i++)
            for (int d = 0; d < NDIM; d++) pts[i * NDIM + d] = (*gPoints)[i].x[d];
        check(hodlr_ctx_create(device, &ctx_), "ctx_create");
        check(hodlr_tree_build(ctx_, NDIM, (int64_t)N_, pts.data(), x1.data(), x2.data(), Nmax,
                               INTERACTION_TYPE_ALLOWED, &tree_),
              "tree_build");
        hodlr_tree_info ti;
        hodlr_tree_get_info(tree_, &ti);
        std::printf("Tree Formed with depth %d\nHMATRIX DONE\n", ti.nlevels - 1);
    }
    HODLRdD_matrix(const HODLRdD_matrix &) = delete;
    HODLRdD_matrix &operator=(const HODLRdD_matrix &) = delete;
    ~HODLRdD_matrix() {
        hodlr_op_destroy(op_);
        hodlr_tree_destroy(tree_);
        hodlr_ctx_destroy(ctx_);
        delete kernel_func;
    }
    size_t get_size() { return N_; }
    void Assemble_matrix_operators() {
        hodlr_kernel_desc kd = kernel_->device_kernel();
        check(hodlr_op_assemble_skeleton(ctx_, 1, &tree_, &kd, eps_ACA, SYS_SIZE, HODLR_PREC_F64 | HODLR_PREC_F32,
                                         &op_),
              "op_assemble");
        std::printf("Matrix operators formed...\n");
    }
    Vec operator*(const Vec &x) {
        Vec b(N_);
        check(hodlr_matvec(op_, HODLR_PREC_F64, 1, x.data(), b.data()), "matvec");
        return b;
    }
    // GMRES(100), max_iter = N, tol = eps_ACA, x0 = b (HDD_matrix.hpp:51-62)
    Vec solve(Vec &b);
    void print_matrix_statistics() {
        hodlr_op_stats s;
        hodlr_op_get_stats(op_, &s);
        std::printf("Memory (in GB) : %g\n", 8e-9 * s.n_flop_ref);
        std::printf("Number of FLOP : %g\n", s.n_flop_ref);
        std::printf("Compression Ratio : %g\n", s.n_flop_ref / ((double)N_ * (double)N_));
        std::printf("Maximum rank across the Tree : %lld\n", (long long)s.max_rank);
        std::printf("Time to Initialize (in s) : %g\n", s.assemble_seconds);
    }
    hodlr_op *handle() { return op_; }
};

// GMRES.hpp:30-53 for the HODLR operator: the solve runs on the device.
template <class Kernel>
class iterSolver<HODLRdD_matrix<Kernel>> {
    int restart, max_iter, num_iter = 0;
    double tol, resid = 0.0;
    std::string fname;
    bool ptofile = false;

public:
    iterSolver(int restart_, int max_iter_, double tol_) : restart(restart_), max_iter(max_iter_), tol(tol_) {}
    iterSolver(int max_iter_, double tol_) : restart(max_iter_), max_iter(max_iter_), tol(tol_) {}
    void set_output_file(std::string f) {
        fname = f;
        ptofile = true;
    }
    double getResidual() { return resid; }
    int getMaxIterations() { return num_iter; }
    // x holds the initial guess on entry and the solution on exit
    int GMRES(HODLRdD_matrix<Kernel> *&A, Vec &x, Vec &b) {
        std::vector<double> hist(2 * (size_t)max_iter + 8);
        hodlr_gmres_params p{};
        p.restart = restart;
        p.max_iter = max_iter;
        p.tol = tol;
        p.mode = HODLR_SOLVE_F64;
        p.x0_is_b = 0;
        p.hist_cap = (int)hist.size();
        p.hist = hist.data();
        hodlr_gmres_info info{};
        int rc = hodlr_gmres(A->op_, b.data(), x.data(), &p, &info);
        A->check(rc, "gmres");
        resid = info.resid;
        num_iter = info.iterations;
        if (ptofile) {  // GMRES.hpp:166-172,227-228
            FILE *f = std::fopen(fname.c_str(), "w");
            if (f) {
                std::fprintf(f, "Restarted GMRES Parameters \nMaximum Iterations : %d\nTolerance :%g\n"
                                "Restart Iteration at %d\n",
                             max_iter, tol, restart);
                for (int i = 0; i < info.hist_len; i++) std::fprintf(f, "Iter[%d] : %.8g\n", i + 1, hist[i]);
                std::fprintf(f, "Residual : %g\n", info.resid);
                std::fclose(f);
            }
        }
        return info.iterations;
    }
};

template <class Kernel>
Vec HODLRdD_matrix<Kernel>::solve(Vec &b) {
    Vec x = b;
    HODLRdD_matrix<Kernel> *mat_obj = this;
    iterSolver<HODLRdD_matrix<Kernel>> solve_obj(100, (int)N_, eps_ACA);
    solve_obj.set_output_file("gmres_out.txt");
    int k = solve_obj.GMRES(mat_obj, x, b);
    std::printf("GMRES status : %d\n", k);
    return x;
}
