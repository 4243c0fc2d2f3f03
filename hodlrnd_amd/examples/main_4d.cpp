// The reference's driver (include/main.cpp:12-59) on the gfx950 library:
// kernel_4d_test on a numPoints^4 tensor grid, assemble, A*x, GMRES solve,
// statistics, and the dense self-check for N < 500000.  Knobs as in the
// reference's myHeaders.hpp (NDIM = 4, Nmax = 1000, d' = 1, eps = 1e-10).
#define HODLR_NDIM 4
#define HODLR_NMAX 1000
#define HODLR_DPRIME 1
#define HODLR_EPS_ACA 1e-10
#include "hodlrnd/HODLRnD.hpp"

int main(int argc, char *argv[]) {
    std::vector<ptsnD> *gridPoints;
    numPoints = argc > 1 ? std::atoi(argv[1]) : 10;
    N = (size_t)std::pow(numPoints, NDIM);
    Vec X(NDIM), Y(NDIM);
    for (int i = 0; i < NDIM; ++i) {
        X(i) = -1;
        Y(i) = 1;
    }
    kernel_4d_test *ker = new kernel_4d_test();
    ker->get_points(gridPoints);
    kernel_function<kernel_4d_test> *kernelfunc = new kernel_function<kernel_4d_test>(ker);
    HODLRdD_matrix<kernel_4d_test> Kmat(ker, gridPoints, X, Y);
    Kmat.Assemble_matrix_operators();
    Vec x_test = Vec::Random(N, 1);
    std::printf("The size of K matrix %zu\n", Kmat.get_size());
    Vec b1 = Kmat * x_test;
    Vec x_sol = Kmat.solve(b1);
    Kmat.print_matrix_statistics();
    if (N < 500000) {
        Vec b2 = Vec::Zero(N, 1);
        std::vector<size_t> v3;
        for (size_t i = 0; i < N; i++) v3.push_back(i);
        for (size_t i = 0; i < N; i++) b2(i) = kernelfunc->getRow((int)i, v3).dot(x_test);
        std::printf("Relative Error in hmatrix matvec ... %g\n", Vec_ops::relative_error(b2, b1));
        std::printf("Relative Error in GMRES solution ... %g\n", Vec_ops::relative_error(x_sol, x_test));
    }
    delete kernelfunc;
    return 0;
}
