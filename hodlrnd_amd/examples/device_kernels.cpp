// Kernel classes in the reference's shape (user_kernel.hpp:62-115: points
// built in the constructor, Kernel_Fun, getMatrixEntry over
// gridPoints->at(i)) whose Kernel_Fun is none of the built-in families,
// compiled unchanged as host and device code (hipcc -x hip): the operator
// identifies getMatrixEntry as diag / w_j Kernel_Fun(r) and assembles with
// entry kernels instantiated here for each Kernel_Fun.
//
//   device_kernels check [numPoints] [outdir]
//     per kernel: how it was identified, A*x against the dense product of
//     getMatrixEntry (main.cpp:42-57), the GMRES solution error, the
//     statistics; with outdir, the points, the identified description, the
//     block ranks and A*x are written there for the oracle's comparison
#define HODLR_NDIM 3
#define HODLR_NMAX 64
#define HODLR_DPRIME 0
#define HODLR_EPS_ACA 1e-10
#include <cstring>
#include <string>

#include "hodlrnd/device_kernels_begin.hpp"

// ---- kernel classes, written as for the reference ----

// the reference's kernel_3d_test: 1/(4 pi r), zero diagonal (a built-in form)
class green_3d {
    double c = 1.0 / (4 * PI);

public:
    std::vector<ptsnD> *gridPoints;
    green_3d() {
        gridPoints = new std::vector<ptsnD>;
        VectorXd loc_dir[NDIM];
        Eigen::VectorXd X(NDIM), Y(NDIM);
        for (int i = 0; i < NDIM; ++i) {
            X(i) = -1;
            Y(i) = 1;
        }
        int cts = 0;
        for (int i = 0; i < NDIM; ++i) loc_dir[i] = uniform_nodes(X(i), Y(i), numPoints);
        for (int i = 0; i < numPoints; i++)
            for (int j = 0; j < numPoints; j++)
                for (int k = 0; k < numPoints; k++) {
                    ptsnD temp;
                    temp.x[0] = loc_dir[0][i];
                    temp.x[1] = loc_dir[1][j];
                    temp.x[2] = loc_dir[2][k];
                    temp.id = cts++;
                    gridPoints->push_back(temp);
                }
    }
    void get_points(std::vector<ptsnD> *&src) { src = this->gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return c / x; }
    dtype_base getMatrixEntry(int i, int j) {
        double r;
        if (i == j)
            r = 0.0;
        else
            r = Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
        return r;
    }
    ~green_3d() {}
};

// screened Coulomb (Yukawa) potential exp(-r) / (4 pi r), second kind,
// midpoint weight h^3
class yukawa_3d {
    double c = 1.0 / (4 * PI);
    double h = 2.0 / numPoints;
    double h3 = h * h * h;

public:
    std::vector<ptsnD> *gridPoints;
    yukawa_3d() {
        green_3d g;
        gridPoints = g.gridPoints;
    }
    void get_points(std::vector<ptsnD> *&src) { src = this->gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return c * exp(-x) / x; }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return 1.0;
        return h3 * Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
    ~yukawa_3d() {}
};

// Matern-3/2 covariance (1 + s r) exp(-s r), s = sqrt(3) / 0.5, with a
// unit nugget on the diagonal, scaled by 1/N
class matern_3d {
    double s = std::sqrt(3.0) / 0.5;
    double w;

public:
    std::vector<ptsnD> *gridPoints;
    matern_3d() {
        green_3d g;
        gridPoints = g.gridPoints;
        w = 1.0 / (double)gridPoints->size();
    }
    void get_points(std::vector<ptsnD> *&src) { src = this->gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return (1.0 + s * x) * exp(-s * x); }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return 1.0 + w;
        return w * Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
    ~matern_3d() {}
};

// Gaussian exp(-r^2 / l2) with quadrature weights that vary per point, on
// pseudo-random points
class weighted_gauss {
    double l2 = 0.5;
    std::vector<double> wt;

public:
    std::vector<ptsnD> *gridPoints;
    explicit weighted_gauss(int n) {
        gridPoints = new std::vector<ptsnD>;
        uint64_t st = 2024;
        auto u = [&] {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return -1.0 + 2.0 * (double)(st >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int p = 0; p < n; p++) {
            ptsnD t;
            for (int d = 0; d < NDIM; d++) t.x[d] = u();
            t.id = p;
            gridPoints->push_back(t);
            wt.push_back((1.0 + 0.5 * u()) / n);
        }
    }
    void get_points(std::vector<ptsnD> *&src) { src = this->gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return exp(-x * x / l2); }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return 1.0;
        return wt[j] * Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
    ~weighted_gauss() {}
};

#include "hodlrnd/device_kernels_end.hpp"

static void save(const std::string &dir, const std::string &name, const std::vector<double> &v) {
    if (dir.empty()) return;
    if (hodlr_vec_save((dir + "/" + name).c_str(), (int64_t)v.size(), v.data()) != HODLR_OK)
        throw hodlr::Error(HODLR_E_ARG, "cannot write " + dir + "/" + name);
}

template <class K>
static void check(const char *name, K *ker, const std::string &dir) {
    Vec X(NDIM), Y(NDIM);
    for (int i = 0; i < NDIM; ++i) {
        X(i) = -1;
        Y(i) = 1;
    }
    std::vector<ptsnD> *pts;
    ker->get_points(pts);
    HODLRdD_matrix<K> Kmat(ker, pts, X, Y);
    Kmat.Assemble_matrix_operators();
    // the description the assembly used (its weights stay valid until the
    // next call)
    const void *entries = nullptr;
    const hodlr_kernel_desc kd = Kmat.device_description(&entries);
    const size_t n = Kmat.get_size();
    std::printf("%s form=%s family=%d scale=%.17g diag=%.17g weights=%s\n", name,
                kd.family == HODLR_KERNEL_CUSTOM ? "device_kernel_fun" : "builtin", kd.family, kd.scale, kd.diag,
                kd.weight ? "per-point" : "none");
    Vec x(n);
    for (size_t i = 0; i < n; i++) x(i) = std::cos(0.11 * (double)i);
    Vec b1 = Kmat * x;
    kernel_function<K> kf(ker);
    std::vector<size_t> all(n);
    for (size_t i = 0; i < n; i++) all[i] = i;
    Vec b2(n);
    for (size_t i = 0; i < n; i++) b2(i) = kf.getRow((int)i, all).dot(x);
    std::printf("%s matvec vs dense getMatrixEntry product %.3g\n", name, Vec_ops::relative_error(b2, b1));
    if (kd.diag != 0.0) {
        Vec x_sol = Kmat.solve(b1);
        std::printf("%s gmres solution error %.3g\n", name, Vec_ops::relative_error(x, x_sol));
    }
    Kmat.print_matrix_statistics();
    if (dir.empty()) return;
    std::vector<double> p(n * NDIM), desc{(double)kd.family, kd.scale, kd.diag}, w;
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < NDIM; d++) p[i * NDIM + d] = (*pts)[i].x[d];
    if (kd.weight) w.assign(kd.weight, kd.weight + n);
    hodlr_op_stats st;
    hodlr_op_get_stats(Kmat.handle(), &st);
    std::vector<int32_t> rk((size_t)st.n_blocks), dn((size_t)st.n_blocks);
    hodlr_op_block_ranks(Kmat.handle(), rk.data(), dn.data(), st.n_blocks);
    std::vector<double> ranks(rk.begin(), rk.end()), dense(dn.begin(), dn.end());
    const std::string s = name;
    save(dir, s + "_pts", p);
    save(dir, s + "_desc", desc);
    save(dir, s + "_weights", w);
    save(dir, s + "_ranks", ranks);
    save(dir, s + "_dense", dense);
    save(dir, s + "_y", std::vector<double>(b1.data(), b1.data() + n));
}

int main(int argc, char *argv[]) {
    if (argc < 2 || std::strcmp(argv[1], "check") != 0) {
        std::fprintf(stderr, "usage: device_kernels check [numPoints] [outdir]\n");
        return 2;
    }
    numPoints = argc > 2 ? std::atoi(argv[2]) : 12;
    const std::string dir = argc > 3 ? argv[3] : "";
    try {
        check("green_3d", new green_3d(), dir);
        check("yukawa_3d", new yukawa_3d(), dir);
        check("matern_3d", new matern_3d(), dir);
        check("weighted_gauss", new weighted_gauss(numPoints * numPoints * numPoints), dir);
    } catch (const hodlr::Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
