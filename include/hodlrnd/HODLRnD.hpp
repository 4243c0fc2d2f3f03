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
//   iterSolver<Op> for any Op with Vec operator*(Vec)      (GMRES.hpp:31-303)
// The knobs are compile-time constants as in the reference; define
// HODLR_NDIM, HODLR_NMAX, HODLR_EPS_ACA or HODLR_DPRIME before including to
// change them.
//
// User kernel classes written for the reference compile unchanged: the
// device cannot call their host getMatrixEntry, so the operator identifies
// the entries instead.  It probes getMatrixEntry and accepts
//   A_ij = diag (i == j),  w_j * f(|x_i - x_j|) (i != j),
//   w_j a constant or a per-column weight,
// when every probed entry agrees to 1e-13 relative, in this order:
//   1. the class's own description, hodlr_kernel_desc device_kernel() const,
//      when it has one (no probing);
//   2. f one of the built-in log r, 1/r, exp(-r);
//   3. f the class's own Kernel_Fun, run on the GPU: compiled with hipcc
//      (-x hip -ffp-contract=off), the classes included between
//        #include "hodlrnd/device_kernels_begin.hpp"
//        #include "user_kernel.hpp"        (or the classes themselves)
//        #include "hodlrnd/device_kernels_end.hpp"
//      are host and device code, and the library's assembly runs entry
//      kernels instantiated for their Kernel_Fun in the caller's program
//      (hodlr_op_assemble_custom; every factor is stored, whatever SYS_SIZE).
//      Kernel_Fun may read the object's own members (the object's bytes are
//      copied to the device) but not follow its pointers or host globals.
// A kernel of another form fails with a message.
// Vectors are plain host arrays (the reference uses Eigen, which this API
// does not require; the Eigen vector names its user kernels use are
// provided unless Eigen itself is included first).  Failures throw
// hodlr::Error.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
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
// 4 atan(1) (myHeaders.hpp:30), constexpr so device code may use it
constexpr double PI = 3.14159265358979323846;

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
    Vec operator+(const Vec &o) const {
        Vec r(*this);
        for (size_t i = 0; i < d_.size(); i++) r.d_[i] += o.d_[i];
        return r;
    }
    Vec operator*(double a) const {
        Vec r(*this);
        for (auto &v : r.d_) v *= a;
        return r;
    }
    friend Vec operator*(double a, const Vec &v) { return v * a; }
    Vec &operator-=(const Vec &o) {
        for (size_t i = 0; i < d_.size(); i++) d_[i] -= o.d_[i];
        return *this;
    }
    Vec &operator+=(const Vec &o) {
        for (size_t i = 0; i < d_.size(); i++) d_[i] += o.d_[i];
        return *this;
    }
    double squaredNorm() const { return dot(*this); }
};

// the Eigen vector names the reference's user kernels use
// (user_kernel.hpp:17-27: VectorXd, Eigen::VectorXd)
#if !defined(EIGEN_WORLD_VERSION) && !defined(HODLR_NO_EIGEN_NAMES)
using VectorXd = Vec;
namespace Eigen {
using VectorXd = ::Vec;
}
#endif

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

// myHeaders.hpp:122-127
inline std::string getTimeStamp() { return "\n" + std::string(__DATE__) + "," + std::string(__TIME__) + "\n"; }

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

// integral_equation_4D/integral4d.hpp:18 (used by the reference's 4-D kernel)
inline double quadruple_integral(double *a, double *b) { return hodlr_detail::quadruple_integral(a, b); }

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

// ---- entries of a user kernel on the device ----
namespace hodlr_detail {
template <class K, class = void>
struct has_device_kernel : std::false_type {};
template <class K>
struct has_device_kernel<K, std::void_t<decltype(std::declval<K &>().device_kernel())>> : std::true_type {};

inline double family_value(int family, double r) {
    switch (family) {
    case HODLR_KERNEL_LOG: return std::log(r);
    case HODLR_KERNEL_INV_R: return 1.0 / r;
    default: return std::exp(-r);
    }
}

// The device evaluates A_ij = diag (i == j), scale * f(r_ij) * w_j (i != j)
// (hodlr_kernel_desc).  A reference-shaped kernel class only has the host
// callback getMatrixEntry (kernel_function.hpp:26,38,51), so its entries are
// matched to that form for a given f: the diagonal from getMatrixEntry(i, i)
// (equal on every probed i), then a weight per column w_j = A(i_j, j) / f(r)
// from one partner row i_j, checked on 512 further pairs (i, j) to 1e-13
// relative.  Equal weights become the scale; otherwise they are passed per
// point in `weights`.  Returns false when the entries are not of that form.
template <class K, class F>
bool match_form(K *ker, const std::vector<ptsnD> &P, F f, hodlr_kernel_desc &kd, std::vector<double> &weights) {
    const int64_t n = (int64_t)P.size();
    auto A = [&](int64_t i, int64_t j) { return (double)ker->getMatrixEntry((int)i, (int)j); };
    auto r_of = [&](int64_t i, int64_t j) { return nd_points::euclidean_distance(P[(size_t)i], P[(size_t)j]); };
    weights.clear();
    kd.scale = 1.0;
    kd.weight = nullptr;
    if (n < 2) return true;
    // a deterministic stream of off-diagonal probe pairs
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto next = [&](int64_t m) {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return (int64_t)(st % (uint64_t)m);
    };
    const int64_t step = n > 64 ? n / 64 : 1;
    // per-column weight from a partner row whose f(r) is finite and non-zero
    auto weight = [&](int64_t j, double &w) {
        for (int t = 0; t < 8; t++) {
            int64_t i = t == 0 ? (j + n / 2) % n : next(n);
            if (i == j) i = (j + 1) % n;
            const double fv = f(r_of(i, j));
            if (std::isfinite(fv) && fv != 0.0) {
                w = A(i, j) / fv;
                return std::isfinite(w);
            }
        }
        return false;
    };
    std::vector<double> w;
    bool uniform = true;
    double w0 = 0.0;
    const int64_t nprobe = n < 64 ? n : 64;
    for (int64_t q = 0; q < nprobe; q++) {  // a sample of columns first
        double wj;
        if (!weight((q * step) % n, wj)) return false;
        if (q == 0) w0 = wj;
        uniform = uniform && std::fabs(wj - w0) <= 1e-13 * std::fabs(w0);
    }
    if (!uniform) {
        w.resize((size_t)n);
        for (int64_t j = 0; j < n; j++)
            if (!weight(j, w[(size_t)j])) return false;
    }
    for (int q = 0; q < 512; q++) {
        int64_t i = next(n), j = next(n);
        if (i == j) continue;
        const double fv = f(r_of(i, j));
        const double want = A(i, j), got = (uniform ? w0 : w[(size_t)j]) * fv;
        if (!(std::isfinite(fv) && std::fabs(got - want) <= 1e-13 * std::fabs(want) + 1e-300)) return false;
    }
    if (uniform) {
        kd.scale = w0;
    } else {
        weights = std::move(w);
        kd.weight = weights.data();
    }
    return true;
}

// the diagonal, equal on every probed i
template <class K>
double match_diag(K *ker, int64_t n) {
    if (n == 0) return 0.0;
    const double d = ker->getMatrixEntry(0, 0);
    const int64_t step = n > 64 ? n / 64 : 1;
    for (int64_t i = 0; i < n; i += step)
        if (ker->getMatrixEntry((int)i, (int)i) != d)
            throw hodlr::Error(HODLR_E_ARG, "kernel has no single diagonal value (getMatrixEntry(i, i) varies); "
                                            "describe it with device_kernel()");
    return d;
}

// f among the built-in families; throws when none matches
template <class K>
hodlr_kernel_desc identify_kernel(K *ker, const std::vector<ptsnD> &P, std::vector<double> &weights) {
    hodlr_kernel_desc kd{HODLR_KERNEL_INV_R, NDIM, 0.0, 0.0, nullptr};
    kd.diag = match_diag(ker, (int64_t)P.size());
    if (P.size() < 2) return kd;
    const int fams[3] = {HODLR_KERNEL_INV_R, HODLR_KERNEL_LOG, HODLR_KERNEL_EXP};
    for (int family : fams) {
        if (match_form(ker, P, [&](double r) { return family_value(family, r); }, kd, weights)) {
            kd.family = family;
            return kd;
        }
    }
    throw hodlr::Error(HODLR_E_ARG, "kernel entries are not diag / w_j f(|x_i - x_j|) with f = log r, 1/r or "
                                    "exp(-r); describe the kernel with device_kernel(), or compile its class as "
                                    "device code (hodlrnd/device_kernels_begin.hpp ... _end.hpp, hipcc)");
}

// f = the class's own Kernel_Fun (host side of the device form); throws
// when the entries are not diag / w_j Kernel_Fun(r)
template <class K>
hodlr_kernel_desc identify_kernel_fun(K *ker, const std::vector<ptsnD> &P, std::vector<double> &weights) {
    hodlr_kernel_desc kd{HODLR_KERNEL_CUSTOM, NDIM, 1.0, 0.0, nullptr};
    kd.diag = match_diag(ker, (int64_t)P.size());
    if (!match_form(ker, P, [&](double r) { return (double)ker->Kernel_Fun(r); }, kd, weights))
        throw hodlr::Error(HODLR_E_ARG, "kernel entries are not diag / w_j Kernel_Fun(|x_i - x_j|)");
    return kd;
}

// Specialized by hodlrnd/device_kernels_end.hpp for the classes whose
// Kernel_Fun is device code: launchers() is then their entry-kernel table.
template <class K, class = void>
struct device_kernel_fun {
    static constexpr bool value = false;
    static const void *launchers() { return nullptr; }
};
}  // namespace hodlr_detail

// ---- GMRES for any operator (GMRES.hpp:31-303) ----
// Restarted GMRES(m) with modified Gram-Schmidt, real Givens rotations and
// back substitution, on host vectors, for any Op with Vec operator*(const
// Vec&) (the reference's iterSolver is generic the same way).  The HODLR
// operator has its own specialization below that runs on the device.
template <class Op>
class iterSolver {
    int restart, max_iter, num_iter = 0;
    double tol, resid = 0.0;
    std::string fname;
    bool ptofile = false;

    static void gen_rot(double &dx, double &dy, double &cs, double &sn) {  // GMRES.hpp:83-110
        if (dy == 0.0) {
            cs = 1.0;
            sn = 0.0;
        } else if (std::fabs(dy) > std::fabs(dx)) {
            const double t = dx / dy;
            sn = 1.0 / std::sqrt(1.0 + t * t);
            cs = t * sn;
        } else {
            const double t = dy / dx;
            cs = 1.0 / std::sqrt(1.0 + t * t);
            sn = t * cs;
        }
    }
    static void app_rot(double &dx, double &dy, double cs, double sn) {  // GMRES.hpp:112-129
        const double t = cs * dx + sn * dy;
        dy = -sn * dx + cs * dy;
        dx = t;
    }

public:
    iterSolver(int restart_, int max_iter_, double tol_) : restart(restart_), max_iter(max_iter_), tol(tol_) {}
    iterSolver(int max_iter_, double tol_) : restart(max_iter_), max_iter(max_iter_), tol(tol_) {}
    void set_output_file(std::string f) {
        fname = std::move(f);
        ptofile = true;
    }
    double getResidual() { return resid; }
    int getMaxIterations() { return num_iter; }
    // x holds the initial guess on entry and the solution on exit; returns
    // the iteration count (max_iter when tol was not reached)
    int GMRES(Op *&A, Vec &x, Vec &b) {
        FILE *f = ptofile ? std::fopen(fname.c_str(), "w") : nullptr;
        if (f)
            std::fprintf(f, "Restarted GMRES Parameters \nMaximum Iterations : %d\nTolerance :%g\nRestart Iteration at %d\n",
                         max_iter, tol, restart);
        double mv_time = 0.0;
        int mv_count = 0;
        auto apply = [&](const Vec &v) {
            const auto t0 = std::chrono::steady_clock::now();
            Vec y = (*A) * v;
            mv_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            mv_count++;
            return y;
        };
        auto finish = [&](const char *why) {
            if (!f) return;
            std::fprintf(f, "%s\nResidual : %g\nMat-Vec Time : %g\n", why, resid, mv_time / mv_count);
            std::fclose(f);
            f = nullptr;
        };
        const int m = restart;
        std::vector<double> H((size_t)(m + 1) * m, 0.0), s(m + 1), cs(m + 1), sn(m + 1);
        auto h = [&](int r, int c) -> double & { return H[(size_t)c * (m + 1) + r]; };
        double normb = b.norm();
        Vec r = b - apply(x);
        double beta = r.norm();
        if (normb == 0.0) normb = 1.0;
        resid = beta / normb;
        if (resid <= tol) {
            num_iter = 0;
            finish("Exited before any Progress");
            return 0;
        }
        std::vector<Vec> v(m + 1);
        auto update = [&](int k) {  // GMRES.hpp:132-146: x += V y, H y = s by back substitution
            std::vector<double> y(s.begin(), s.begin() + k + 1);
            for (int i = k; i >= 0; i--) {
                y[i] /= h(i, i);
                for (int j = i - 1; j >= 0; j--) y[j] -= h(j, i) * y[i];
            }
            for (int j = 0; j <= k; j++) x += v[j] * y[j];
        };
        int j = 1;
        while (j <= max_iter) {
            v[0] = r * (1.0 / beta);
            std::fill(s.begin(), s.end(), 0.0);
            s[0] = beta;
            int i = 0;
            for (; i < m && j <= max_iter; i++, j++) {
                if (f) std::fprintf(f, "Iter[%d] : %.8g\n", j, resid);
                Vec w = apply(v[i]);
                for (int k = 0; k <= i; k++) {  // modified Gram-Schmidt
                    h(k, i) = w.dot(v[k]);
                    w -= v[k] * h(k, i);
                }
                h(i + 1, i) = w.norm();
                // a zero norm (exact solution in the space) gives a zero vector
                v[i + 1] = h(i + 1, i) != 0.0 ? w * (1.0 / h(i + 1, i)) : Vec(w.size());
                for (int k = 0; k < i; k++) app_rot(h(k, i), h(k + 1, i), cs[k], sn[k]);
                gen_rot(h(i, i), h(i + 1, i), cs[i], sn[i]);
                app_rot(h(i, i), h(i + 1, i), cs[i], sn[i]);
                app_rot(s[i], s[i + 1], cs[i], sn[i]);
                resid = std::fabs(s[i + 1]) / normb;
                if (resid < tol) {
                    update(i);
                    num_iter = j;
                    finish("Reached Solution before Max_Iterations");
                    return j;
                }
            }
            update(i - 1);  // the columns of this cycle (m of them unless max_iter ended it)
            r = b - apply(x);
            beta = r.norm();
            resid = beta / normb;
            if (resid < tol) {
                num_iter = j;
                finish("Reached Solution before Max_Iterations");
                return j;
            }
        }
        if (f) std::fclose(f);
        num_iter = max_iter;
        return max_iter;
    }
};

// ---- the operator (HDD_matrix.hpp:10-89) ----
template <class Kernel>
class HODLRdD_matrix {
    size_t N_ = 0;
    hodlr_ctx *ctx_ = nullptr;
    hodlr_tree *tree_ = nullptr;
    hodlr_op *op_ = nullptr;
    kernel_function<Kernel> *kernel_func;
    Kernel *kernel_;
    std::vector<ptsnD> *points_;
    std::vector<double> lo_, hi_, weights_;
    double matvec_seconds_ = 0.0;

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
        points_ = gPoints;
        N_ = gPoints->size();
        lo_.assign(x1.data(), x1.data() + NDIM);
        hi_.assign(x2.data(), x2.data() + NDIM);
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
    // the kernel's description for the device: its own device_kernel()
    // when it has one, else identified from getMatrixEntry against the
    // built-in families, then (device-compiled classes) against its own
    // Kernel_Fun, in which case *entries receives its entry-kernel table;
    // per-point weights are held by this object until the next call
    hodlr_kernel_desc device_description(const void **entries = nullptr) {
        if (entries) *entries = nullptr;
        if constexpr (hodlr_detail::has_device_kernel<Kernel>::value) {
            return kernel_->device_kernel();
        } else {
            if constexpr (hodlr_detail::device_kernel_fun<Kernel>::value) {
                try {
                    return hodlr_detail::identify_kernel(kernel_, *points_, weights_);
                } catch (const hodlr::Error &) {
                }
                hodlr_kernel_desc kd = hodlr_detail::identify_kernel_fun(kernel_, *points_, weights_);
                if (entries) *entries = hodlr_detail::device_kernel_fun<Kernel>::launchers();
                return kd;
            }
            return hodlr_detail::identify_kernel(kernel_, *points_, weights_);
        }
    }
    void Assemble_matrix_operators() {
        const void *entries = nullptr;
        hodlr_kernel_desc kd = device_description(&entries);
        if (kd.family == HODLR_KERNEL_CUSTOM) {
            if (!entries)
                throw hodlr::Error(HODLR_E_ARG, "custom kernel family without device-compiled Kernel_Fun");
            check(hodlr_op_assemble_custom(ctx_, tree_, &kd, static_cast<const void *>(kernel_),
                                           (int64_t)sizeof(Kernel), entries, eps_ACA,
                                           HODLR_PREC_F64 | HODLR_PREC_F32, &op_),
                  "op_assemble");
        } else {
            check(hodlr_op_assemble_skeleton(ctx_, 1, &tree_, &kd, eps_ACA, SYS_SIZE,
                                             HODLR_PREC_F64 | HODLR_PREC_F32, &op_),
                  "op_assemble");
        }
        std::printf("Matrix operators formed...\n");
    }
    Vec operator*(const Vec &x) {
        Vec b(N_);
        const auto t0 = std::chrono::steady_clock::now();
        check(hodlr_matvec(op_, HODLR_PREC_F64, 1, x.data(), b.data()), "matvec");
        matvec_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return b;
    }
    // GMRES(100), max_iter = N, tol = eps_ACA, x0 = b (HDD_matrix.hpp:51-62)
    Vec solve(Vec &b);
    // HDD_matrix.hpp:63-67 -> HDD_tree.hpp:117-132, HDD_node.hpp:100-113:
    // every node's point count, box, neighbours and interaction list
    void print_matrix_details() {
        std::printf("============HODLRnD=================\n");
        hodlr_tree_info ti;
        hodlr_tree_get_info(tree_, &ti);
        std::vector<int64_t> loff(ti.nlevels + 1), start(ti.nnodes), count(ti.nnodes), nptr(ti.nnodes + 1),
            iptr(ti.nnodes + 1);
        hodlr_tree_nodes(tree_, loff.data(), start.data(), count.data());
        hodlr_tree_lists(tree_, 0, nptr.data(), nullptr);
        hodlr_tree_lists(tree_, 1, iptr.data(), nullptr);
        std::vector<int64_t> nidx(nptr[ti.nnodes] > 0 ? nptr[ti.nnodes] : 1),
            iidx(iptr[ti.nnodes] > 0 ? iptr[ti.nnodes] : 1);
        hodlr_tree_lists(tree_, 0, nptr.data(), nidx.data());
        hodlr_tree_lists(tree_, 1, iptr.data(), iidx.data());
        std::printf("Tree Depth:%d\n______________________________\n\n", ti.nlevels);
        for (int l = 0; l < ti.nlevels; l++) {
            std::printf("Level[%d]\n", l);
            for (int64_t j = 0; j < loff[l + 1] - loff[l]; j++) {
                const int64_t q = loff[l] + j;
                std::printf("**************\nNode [%lld] , Nunmber particles = %lld\n", (long long)j,
                            (long long)count[q]);
                // the box: midpoint splits from the root box, dimension 0 in
                // the most significant bit of each level's child index
                std::printf("\nBounding box : ");
                for (int d = 0; d < NDIM; d++) {
                    double a = lo_[d], b = hi_[d];
                    for (int t = l - 1; t >= 0; t--) {
                        const double mid = 0.5 * (a + b);
                        if ((j >> (t * NDIM + (NDIM - 1 - d))) & 1) a = mid;
                        else b = mid;
                    }
                    std::printf(d + 1 < NDIM ? "[%g,%g]x" : "[%g,%g]", a, b);
                }
                std::printf("\nNeighbours -");
                for (int64_t k = nptr[q]; k < nptr[q + 1]; k++) std::printf(" %lld", (long long)nidx[k]);
                std::printf("\nInteraction List -");
                for (int64_t k = iptr[q]; k < iptr[q + 1]; k++) std::printf(" %lld", (long long)iidx[k]);
                std::printf("\n**************\n\n");
            }
            std::printf("______________________________\n");
        }
        std::printf("====================================\n");
    }
    void print_matrix_statistics() {  // HDD_matrix.hpp:69-85
        hodlr_op_stats s;
        hodlr_op_get_stats(op_, &s);
        std::printf("%s\n", getTimeStamp().c_str());
        std::printf("Memory (in GB) : %g\n", 8e-9 * s.n_flop_ref);
        std::printf("Number of FLOP : %g\n", s.n_flop_ref);
        std::printf("Compression Ratio : %g\n", s.n_flop_ref / ((double)N_ * (double)N_));
        std::printf("Maximum rank across the Tree : %lld\n", (long long)s.max_rank);
        std::printf("Time to Initialize (in s) : %g\n", s.assemble_seconds);
        // wall time of the last operator* (host vectors in and out, as the
        // reference's timer around Tree::mat_vec, HDD_tree.hpp:95-99)
        std::printf("Matrix-Vector time (in s) : %g\n", matvec_seconds_);
    }
    hodlr_op *handle() { return op_; }
};

// GMRES.hpp:30-303 for the HODLR operator: the solve runs on the device.
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
        p.ortho = HODLR_ORTHO_MGS;  // the reference's Arnoldi order (GMRES.hpp:236-240)
        hodlr_gmres_info info{};
        int rc = hodlr_gmres(A->op_, b.data(), x.data(), &p, &info);
        A->check(rc, "gmres");
        // the log's mean matvec time: the device product timed with HIP
        // events (3 products on the operator's stream)
        double mv_ms[4] = {0, 0, 0, 0};
        if (ptofile) A->check(hodlr_matvec_timing(A->op_, HODLR_PREC_F64, 1, 3, mv_ms), "matvec_timing");
        resid = info.resid;
        num_iter = info.iterations;
        if (ptofile) {  // GMRES.hpp:166-172,206-209,227-228,260-263,287-290
            FILE *f = std::fopen(fname.c_str(), "w");
            if (f) {
                std::fprintf(f, "Restarted GMRES Parameters \nMaximum Iterations : %d\nTolerance :%g\n"
                                "Restart Iteration at %d\n",
                             max_iter, tol, restart);
                for (int i = 0; i < info.hist_len && i < info.iterations; i++)
                    std::fprintf(f, "Iter[%d] : %.8g\n", i + 1, hist[i]);
                if (info.iterations == 0 && info.converged)
                    std::fprintf(f, "Exited before any Progress\n");
                else if (info.converged)
                    std::fprintf(f, "Reached Solution before Max_Iterations\n");
                if (info.converged || info.iterations == 0)
                    std::fprintf(f, "Residual : %g\nMat-Vec Time : %g\n", info.resid, mv_ms[0] * 1e-3);
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
