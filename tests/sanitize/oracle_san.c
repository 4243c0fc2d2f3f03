/*
 * oracle_san.c -- drives every entry point of the FP64 CPU oracle under
 * AddressSanitizer and UndefinedBehaviorSanitizer (built by the Makefile
 * beside this file, run by tests/test_sanitizers.py).
 *
 * The cases cover what the oracle's callers reach: 1-D to 4-D trees on
 * random, grid (coordinate ties) and clustered (empty boxes) points, a tree
 * of depth zero, every list, ACA with pivots and a dense fallback, stored
 * and skeleton operators (sys_size -1 / 2 Nmax / 0), sampled assembly, the
 * tree matvec against the dense product, GMRES on the operator and on a
 * dense matrix, the CPQR solve (full rank and rank deficient), the
 * Gauss-Legendre rule and the 4-D quadrature.  The numerical checks are
 * loose: the point is that every path runs clean under the sanitizers.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/hodlr_oracle.h"

static int fails = 0;

#define CHECK(c, ...)                                                  \
    do {                                                               \
        if (!(c)) {                                                    \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);       \
            fprintf(stderr, __VA_ARGS__);                              \
            fputc('\n', stderr);                                       \
            fails++;                                                   \
        }                                                              \
    } while (0)

static double relerr(const double *a, const double *b, int64_t n) {
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < n; i++) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

/* points: 0 uniform random, 1 tensor grid (ties on every split), 2 points
 * clustered in one corner (most boxes empty) */
static double *make_points(int nd, int64_t *n, int kind, uint64_t seed) {
    double *pts;
    if (kind == 1) {
        int m = (int)floor(pow((double)*n, 1.0 / nd) + 1e-9);
        if (m < 2) m = 2;
        double *nodes = malloc(sizeof(double) * m);
        or_uniform_nodes(-1.0, 1.0, m, nodes);
        int64_t cnt = 1;
        for (int d = 0; d < nd; d++) cnt *= m;
        pts = malloc(sizeof(double) * cnt * nd);
        or_tensor_grid(nd, m, nodes, pts);
        free(nodes);
        *n = cnt;
        return pts;
    }
    pts = malloc(sizeof(double) * (*n) * nd);
    or_splitmix_uniform(seed, (*n) * nd, pts);
    if (kind == 2)
        for (int64_t i = 0; i < (*n) * nd; i++) pts[i] = -1.0 + 0.25 * (pts[i] + 1.0);
    return pts;
}

static void check_tree(const or_tree_t *t, int64_t n) {
    const int L = or_tree_nlevels(t);
    CHECK(L >= 1, "no levels");
    for (int l = 0; l < L; l++) {
        const int64_t ns = or_tree_level_size(t, l);
        int64_t total = 0;
        char *seen = calloc((size_t)(n > 0 ? n : 1), 1);
        for (int64_t j = 0; j < ns; j++) {
            const int64_t c = or_tree_node_npts(t, l, j);
            int64_t *ids = malloc(sizeof(int64_t) * (size_t)(c > 0 ? c : 1));
            or_tree_node_points(t, l, j, ids);
            for (int64_t a = 0; a < c; a++) {
                CHECK(ids[a] >= 0 && ids[a] < n, "point id out of range");
                if (ids[a] >= 0 && ids[a] < n) {
                    CHECK(!seen[ids[a]], "point in two nodes of one level");
                    seen[ids[a]] = 1;
                }
                if (a) CHECK(ids[a] > ids[a - 1], "node points not in ascending id order");
            }
            free(ids);
            total += c;
            for (int which = 0; which < 2; which++) {
                const int64_t len = or_tree_list(t, l, j, which, NULL);
                int64_t *lst = malloc(sizeof(int64_t) * (size_t)(len > 0 ? len : 1));
                CHECK(or_tree_list(t, l, j, which, lst) == len, "list length changed");
                for (int64_t a = 0; a < len; a++) {
                    CHECK(lst[a] >= 0 && lst[a] < ns && lst[a] != j, "bad list entry");
                    /* interaction lists are symmetric */
                    if (which == 1 && lst[a] >= 0 && lst[a] < ns) {
                        const int64_t l2 = or_tree_list(t, l, lst[a], 1, NULL);
                        int64_t *back = malloc(sizeof(int64_t) * (size_t)(l2 > 0 ? l2 : 1));
                        or_tree_list(t, l, lst[a], 1, back);
                        int found = 0;
                        for (int64_t q = 0; q < l2; q++) found |= back[q] == j;
                        CHECK(found, "interaction list not symmetric");
                        free(back);
                        (void)or_interaction_type(t, l, j, lst[a]);
                    }
                }
                free(lst);
            }
        }
        CHECK(total == n, "level %d holds %lld of %lld points", l, (long long)total, (long long)n);
        free(seen);
    }
}

static void run_case(const char *name, int nd, int64_t n, int kind, int nmax, int dprime, int family,
                     double eps, int sys_size, double max_err) {
    double *pts = make_points(nd, &n, kind, 1000 + (uint64_t)nd * 17 + (uint64_t)n);
    double lo[OR_MAXD], hi[OR_MAXD];
    for (int d = 0; d < nd; d++) {
        lo[d] = -1.0;
        hi[d] = 1.0;
    }
    or_tree_t *t = or_tree_build(nd, n, pts, lo, hi, nmax, dprime);
    CHECK(t != NULL, "%s: tree build failed", name);
    if (!t) {
        free(pts);
        return;
    }
    check_tree(t, n);
    double yk[1] = {2.0};
    or_kernel_t k = {family, nd, 1.0 / (double)n, family == OR_K_LOG ? 0.0 : 1.0, NULL, NULL, NULL};
    if (family == OR_K_CUSTOM) {
        k.fn = or_f_yukawa;
        k.params = yk;
    }
    double *w = NULL;
    if (kind == 2) { /* weighted columns on the clustered case */
        w = malloc(sizeof(double) * n);
        or_splitmix_uniform(77, n, w);
        for (int64_t i = 0; i < n; i++) w[i] = 1.0 + 0.5 * w[i];
        k.weight = w;
    }
    or_op_t *op = or_op_assemble(t, &k, pts, eps, sys_size, 2);
    CHECK(op != NULL, "%s: assembly failed", name);
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n), *yd = malloc(sizeof(double) * n);
    double *xs = malloc(sizeof(double) * n);
    or_splitmix_uniform(5, n, x);
    or_op_matvec(op, x, y, 2);
    or_dense_matvec(&k, pts, n, x, yd, 2);
    const double e = relerr(y, yd, n);
    CHECK(e < max_err, "%s: matvec error %.3g against the dense product", name, e);
    or_op_stats_t st;
    or_op_get_stats(op, &st);
    const int64_t nbk = or_op_block_ranks(op, NULL, NULL);
    CHECK(nbk == st.n_blocks, "%s: block count %lld vs stats %lld", name, (long long)nbk, (long long)st.n_blocks);
    int32_t *rk = malloc(sizeof(int32_t) * (size_t)(nbk > 0 ? nbk : 1));
    int32_t *dn = malloc(sizeof(int32_t) * (size_t)(nbk > 0 ? nbk : 1));
    or_op_block_ranks(op, rk, dn);
    for (int64_t b = 0; b < nbk; b++) CHECK(rk[b] >= 0 && (dn[b] == 0 || dn[b] == 1), "%s: bad rank/flag", name);
    free(rk);
    free(dn);
    /* GMRES with a restart shorter than the iteration count, and the history */
    memcpy(xs, y, sizeof(double) * n);
    double hist[64];
    or_gmres_info_t gi;
    const int it = or_gmres(op, y, xs, 7, 60, 1e-10, 2, hist, 64, &gi);
    CHECK(it >= 0 && isfinite(gi.resid), "%s: gmres returned %d, resid %g", name, it, gi.resid);
    or_op_free(op);
    /* sampled assembly: every other node on each level */
    const int L = or_tree_nlevels(t);
    int64_t stride[64], off[64];
    for (int l = 0; l < L && l < 64; l++) {
        stride[l] = 2;
        off[l] = l & 1;
    }
    or_op_t *sop = or_op_assemble_sample(t, &k, pts, eps, sys_size, 2, stride, off);
    CHECK(sop != NULL, "%s: sampled assembly failed", name);
    if (sop) {
        or_op_matvec(sop, x, y, 2);
        or_op_free(sop);
    }
    or_tree_free(t);
    free(x);
    free(y);
    free(yd);
    free(xs);
    free(w);
    free(pts);
    fprintf(stderr, "ok %-28s n=%-6lld levels=%d matvec err %.2e gmres %d it\n", name, (long long)n, L, e, it);
}

static void aca_direct(void) {
    int64_t n = 400;
    double *pts = make_points(2, &n, 0, 9);
    int64_t rows[120], cols[150];
    int m = 0, c = 0;
    for (int64_t i = 0; i < n && (m < 120 || c < 150); i++) {
        if (pts[2 * i] < -0.2 && m < 120) rows[m++] = i;
        else if (pts[2 * i] > 0.4 && c < 150) cols[c++] = i;
    }
    or_kernel_t k = {OR_K_INV, 2, 1.0, 0.0, NULL, NULL, NULL};
    const int cap = m < c ? m : c;
    double *U = malloc(sizeof(double) * m * cap), *V = malloc(sizeof(double) * c * cap);
    int32_t rp[150], cp[150];
    or_aca_info_t info;
    const int r = or_aca(&k, pts, rows, m, cols, c, 1e-10, or_block_seed(1, 0, 1), cap, U, V, rp, cp, &info);
    CHECK(r > 0 && r <= cap, "aca rank %d", r);
    /* a block whose partners touch: the dense fallback */
    int64_t rows2[40], cols2[40];
    for (int i = 0; i < 40; i++) {
        rows2[i] = i;
        cols2[i] = 40 + i;
    }
    or_aca_info_t info2;
    const int r2 = or_aca(&k, pts, rows2, 40, cols2, 40, 1e-14, 3, 40, U, V, rp, cp, &info2);
    CHECK(r2 >= 0 && r2 <= 40, "aca rank %d", r2);
    /* scale 0: rank 0 below the floor */
    or_kernel_t k0 = {OR_K_INV, 2, 0.0, 0.0, NULL, NULL, NULL};
    or_aca_info_t info3;
    const int r3 = or_aca(&k0, pts, rows, m, cols, c, 1e-8, 5, cap, U, V, rp, cp, &info3);
    CHECK(r3 == 0, "zero kernel rank %d", r3);
    free(U);
    free(V);
    free(pts);
    fprintf(stderr, "ok aca direct: ranks %d / %d (dense %d) / %d\n", r, r2, info2.dense, r3);
}

static void dense_and_rules(void) {
    /* CPQR: full rank and rank deficient */
    const int n = 24;
    double *A = malloc(sizeof(double) * n * n), *b = malloc(sizeof(double) * n), *x = malloc(sizeof(double) * n);
    or_splitmix_uniform(31, (int64_t)n * n, A);
    or_splitmix_uniform(32, n, b);
    CHECK(or_cpqr_solve(n, A, b, x) >= 0, "cpqr");
    for (int i = 0; i < n; i++) A[(size_t)5 * n + i] = A[(size_t)3 * n + i]; /* repeated column */
    CHECK(or_cpqr_solve(n, A, b, x) >= 0, "cpqr rank deficient");
    /* GMRES on a dense diagonally dominant matrix */
    for (int i = 0; i < n; i++) A[(size_t)i * n + i] += (double)n;
    or_gmres_info_t gi;
    memcpy(x, b, sizeof(double) * n);
    const int it = or_gmres_dense(A, n, b, x, 5, 100, 1e-12, &gi);
    CHECK(it >= 0 && gi.resid < 1e-10, "dense gmres %d %g", it, gi.resid);
    free(A);
    free(b);
    free(x);
    double nodes[32], weights[32];
    for (int q = 1; q <= 32; q *= 2) {
        CHECK(or_gauss_legendre(q, nodes, weights) == 0, "gauss-legendre %d", q);
        double s = 0.0;
        for (int i = 0; i < q; i++) s += weights[i];
        CHECK(fabs(s - 2.0) < 1e-13, "weights sum %g", s);
    }
    const double a4[4] = {-0.1, -0.1, -0.1, -0.1}, b4[4] = {0.1, 0.1, 0.1, 0.1};
    const double I = or_quadruple_integral(a4, b4, 4);
    CHECK(isfinite(I), "quadrature");
    double pm[100];
    or_park_miller(100, pm);
    double w1[4] = {1, 2, 3, 4}, w4[256];
    or_tensor_weights(4, 4, w1, w4);
    CHECK(w4[255] == 256.0, "tensor weights");
    fprintf(stderr, "ok dense solvers and rules (gmres %d it)\n", it);
}

int main(void) {
    run_case("1d random inv", 1, 600, 0, 16, 0, OR_K_INV, 1e-10, -1, 1e-7);
    run_case("2d random log", 2, 1500, 0, 32, 0, OR_K_LOG, 1e-10, -1, 1e-7);
    run_case("2d grid exp d'=-1", 2, 1600, 1, 32, -1, OR_K_EXP, 1e-10, -1, 1e-7);
    run_case("3d random inv skeleton", 3, 2500, 0, 32, 0, OR_K_INV, 1e-8, 64, 1e-4);
    run_case("3d grid inv sys_size 0", 3, 1728, 1, 32, 0, OR_K_INV, 1e-8, 0, 1e-4);
    run_case("3d clustered weighted", 3, 2000, 2, 32, 0, OR_K_INV, 1e-8, -1, 1e-5);
    run_case("3d custom yukawa", 3, 1500, 0, 32, 0, OR_K_CUSTOM, 1e-10, -1, 1e-7);
    run_case("4d random d'=3", 4, 2000, 0, 64, 3, OR_K_EXP, 1e-8, -1, 1e-5);
    run_case("4d grid d'=-1", 4, 1296, 1, 64, -1, OR_K_INV, 1e-8, -1, 1e-5);
    run_case("depth zero", 3, 50, 0, 64, 0, OR_K_INV, 1e-10, -1, 1e-12);
    aca_direct();
    dense_and_rules();
    if (fails) {
        fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    fprintf(stderr, "oracle sanitizer run clean\n");
    return 0;
}
