// User kernel classes in the reference's shape (user_kernel.hpp:10-179:
// points built in the constructor with Eigen vectors, Kernel_Fun,
// getMatrixEntry over gridPoints->at(i)), with no device description: the
// operator identifies their entries from getMatrixEntry.
//
//   user_kernels identify   host only: the identified descriptions, a kernel
//                           that is refused, and the generic iterSolver on a
//                           dense host operator
//   user_kernels full [n]   on the GPU: assemble, A*x against the dense
//                           product of getMatrixEntry, GMRES, statistics
#define HODLR_NDIM 3
#define HODLR_NMAX 64
#define HODLR_DPRIME 0
#define HODLR_EPS_ACA 1e-10
// every factor stored: the accuracy bars below are those of the stored form
// (the skeleton form, SYS_SIZE = 2 Nmax, is covered by its own tests)
#define HODLR_SYS_SIZE -1
#include <cstring>

#include "hodlrnd/HODLRnD.hpp"

// 3-D Laplace Green's function on a tensor grid, zero diagonal
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
        if (i == j) return 0.0;
        return Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
};

// second-kind equation with midpoint-rule weights that vary per point:
// A_ij = 1 (i == j), w_j / |x_i - x_j| otherwise, points pseudo-random
class weighted_second_kind {
    std::vector<double> w;

public:
    std::vector<ptsnD> *gridPoints;
    explicit weighted_second_kind(int n) {
        gridPoints = new std::vector<ptsnD>;
        uint64_t s = 12345;
        auto u = [&] {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            return -1.0 + 2.0 * (double)(s >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int p = 0; p < n; p++) {
            ptsnD t;
            for (int d = 0; d < NDIM; d++) t.x[d] = u();
            t.id = p;
            gridPoints->push_back(t);
            w.push_back((1.0 + 0.5 * u()) / n);
        }
    }
    void get_points(std::vector<ptsnD> *&src) { src = this->gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return 1.0 / x; }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return 1.0;
        return w[j] * Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
};

// Yukawa potential exp(-r)/r: not of the device's forms, refused
class yukawa_3d {
public:
    std::vector<ptsnD> *gridPoints;
    yukawa_3d() {
        numPoints = 6;
        green_3d g;
        gridPoints = g.gridPoints;
    }
    void get_points(std::vector<ptsnD> *&src) { src = this->gridPoints; }
    dtype_base Kernel_Fun(dtype_base x) { return std::exp(-x) / x; }
    dtype_base getMatrixEntry(int i, int j) {
        if (i == j) return 1.0;
        return Kernel_Fun(nd_points::euclidean_distance(gridPoints->at(i), gridPoints->at(j)));
    }
};

// any operator with Vec operator*(const Vec&) solves through iterSolver
class dense_operator {
public:
    size_t n;
    std::vector<double> a;  // row-major
    template <class K>
    dense_operator(K *k, size_t n_, double shift) : n(n_), a(n_ * n_) {
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) a[i * n + j] = k->getMatrixEntry((int)i, (int)j) / n + (i == j ? shift : 0.0);
    }
    Vec operator*(const Vec &x) {
        Vec y(n);
        for (size_t i = 0; i < n; i++) {
            double s = 0.0;
            for (size_t j = 0; j < n; j++) s += a[i * n + j] * x(j);
            y(i) = s;
        }
        return y;
    }
};

static const char *fam_name(int f) {
    return f == HODLR_KERNEL_LOG ? "log" : (f == HODLR_KERNEL_INV_R ? "inv_r" : "exp");
}

template <class K>
static void report(const char *name, K *k) {
    std::vector<double> w;
    try {
        hodlr_kernel_desc kd = hodlr_detail::identify_kernel(k, *k->gridPoints, w);
        std::printf("identified %s family=%s scale=%.17g diag=%.17g weights=%zu\n", name, fam_name(kd.family),
                    kd.scale, kd.diag, w.size());
    } catch (const hodlr::Error &e) {
        std::printf("refused %s code=%d: %s\n", name, e.code, e.what());
    }
}

static int identify_mode() {
    numPoints = 8;
    green_3d g;
    report("green_3d", &g);
    weighted_second_kind ws(700);
    report("weighted_second_kind", &ws);
    yukawa_3d y;
    report("yukawa_3d", &y);
    // the generic solver (GMRES.hpp:160-303) on host operators
    numPoints = 7;  // 343 points
    green_3d g2;
    dense_operator D(&g2, g2.gridPoints->size(), 1.0);
    dense_operator *Dp = &D;
    Vec xt(D.n);
    for (size_t i = 0; i < D.n; i++) xt(i) = std::sin(0.37 * (double)i);
    Vec b = D * xt;
    for (int restart : {100, 5}) {
        Vec x = Vec::Zero(D.n);
        iterSolver<dense_operator> S(restart, (int)D.n, 1e-12);
        S.set_output_file(restart == 5 ? "gmres_generic_r5.txt" : "gmres_generic.txt");
        int it = S.GMRES(Dp, x, b);
        std::printf("generic gmres restart=%d iterations=%d resid=%.3g error=%.3g\n", restart, it, S.getResidual(),
                    Vec_ops::relative_error(xt, x));
    }
    Vec x0 = b;  // A = I: exact at entry
    dense_operator I(&g2, g2.gridPoints->size(), 0.0);
    std::fill(I.a.begin(), I.a.end(), 0.0);
    for (size_t i = 0; i < I.n; i++) I.a[i * I.n + i] = 1.0;
    dense_operator *Ip = &I;
    iterSolver<dense_operator> S0(100, (int)I.n, 1e-12);
    S0.set_output_file("gmres_generic_exact.txt");
    std::printf("generic gmres identity iterations=%d\n", S0.GMRES(Ip, x0, b));
    return 0;
}

template <class K>
static void dense_check(const char *name, K *ker, std::vector<ptsnD> *pts, Vec &X, Vec &Y) {
    HODLRdD_matrix<K> Kmat(ker, pts, X, Y);
    Kmat.Assemble_matrix_operators();
    const size_t n = Kmat.get_size();
    Vec x = Vec::Zero(n);
    for (size_t i = 0; i < n; i++) x(i) = std::cos(0.11 * (double)i);
    Vec b1 = Kmat * x;
    kernel_function<K> kf(ker);
    std::vector<size_t> all(n);
    for (size_t i = 0; i < n; i++) all[i] = i;
    Vec b2(n);
    for (size_t i = 0; i < n; i++) b2(i) = kf.getRow((int)i, all).dot(x);
    std::printf("%s matvec vs dense getMatrixEntry product %.3g\n", name, Vec_ops::relative_error(b2, b1));
    Vec x_sol = Kmat.solve(b1);
    std::printf("%s gmres solution error %.3g\n", name, Vec_ops::relative_error(x, x_sol));
    Kmat.print_matrix_statistics();
}

static int full_mode(int np) {
    numPoints = np;
    Vec X(NDIM), Y(NDIM);
    for (int i = 0; i < NDIM; ++i) {
        X(i) = -1;
        Y(i) = 1;
    }
    std::vector<ptsnD> *pts;
    green_3d *g = new green_3d();
    g->get_points(pts);
    {
        // first kind with a zero diagonal: checked by its product only
        HODLRdD_matrix<green_3d> Kmat(g, pts, X, Y);
        Kmat.Assemble_matrix_operators();
        const size_t n = Kmat.get_size();
        Vec x(n);
        for (size_t i = 0; i < n; i++) x(i) = std::cos(0.11 * (double)i);
        Vec b1 = Kmat * x;
        kernel_function<green_3d> kf(g);
        std::vector<size_t> all(n);
        for (size_t i = 0; i < n; i++) all[i] = i;
        Vec b2(n);
        for (size_t i = 0; i < n; i++) b2(i) = kf.getRow((int)i, all).dot(x);
        std::printf("green_3d matvec vs dense getMatrixEntry product %.3g\n", Vec_ops::relative_error(b2, b1));
        Kmat.print_matrix_statistics();
        numPoints = 5;  // a small tree (depth 1) for the details printout
        green_3d small;
        std::vector<ptsnD> *sp;
        small.get_points(sp);
        green_3d *smp = &small;
        HODLRdD_matrix<green_3d> S(smp, sp, X, Y);
        S.print_matrix_details();
    }
    weighted_second_kind *ws = new weighted_second_kind(np * np * np);
    ws->get_points(pts);
    dense_check("weighted_second_kind", ws, pts, X, Y);
    yukawa_3d *yk = new yukawa_3d();
    yk->get_points(pts);
    HODLRdD_matrix<yukawa_3d> Ky(yk, pts, X, Y);
    try {
        Ky.Assemble_matrix_operators();
        std::printf("yukawa_3d assembled (unexpected)\n");
    } catch (const hodlr::Error &e) {
        std::printf("yukawa_3d refused code=%d\n", e.code);
    }
    return 0;
}

int main(int argc, char *argv[]) {
    if (argc > 1 && std::strcmp(argv[1], "identify") == 0) return identify_mode();
    if (argc > 1 && std::strcmp(argv[1], "full") == 0) return full_mode(argc > 2 ? std::atoi(argv[2]) : 16);
    std::fprintf(stderr, "usage: user_kernels identify | full [numPoints]\n");
    return 2;
}
