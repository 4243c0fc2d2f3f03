/*
 * hodlr_oracle.c -- FP64 CPU restatement of HODLRnD's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see hodlr_oracle.h.  Citations
 * are file:line in the reference repository kandapva/HODLRnD.
 */
#include "hodlr_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* generators                                                          */
/* ------------------------------------------------------------------ */

/* myHeaders.hpp:111-120 */
void or_uniform_nodes(double a, double b, int n, double *out) {
    double l = (b - a);
    for (int k = 0; k < n; k++) out[k] = a + l / (double)(n - 1) * k;
}

/* user_kernel.hpp:79-95: tensor grid, dimension 0 outermost */
void or_tensor_grid(int ndim, int n, const double *nodes1d, double *pts) {
    int64_t total = 1;
    for (int d = 0; d < ndim; d++) total *= n;
    for (int64_t p = 0; p < total; p++) {
        int64_t rem = p;
        for (int d = ndim - 1; d >= 0; d--) {
            pts[p * ndim + d] = nodes1d[rem % n];
            rem /= n;
        }
    }
}

void or_tensor_weights(int ndim, int n, const double *w1d, double *w) {
    int64_t total = 1;
    for (int d = 0; d < ndim; d++) total *= n;
    for (int64_t p = 0; p < total; p++) {
        int64_t rem = p;
        int64_t idx[OR_MAXD];
        for (int d = ndim - 1; d >= 0; d--) { idx[d] = rem % n; rem /= n; }
        double v = 1.0;
        for (int d = 0; d < ndim; d++) v *= w1d[idx[d]];
        w[p] = v;
    }
}

/* Eigen's VectorXd::Random on a libc whose rand() is the Park-Miller
 * minimal-standard generator with seed 1 (the generator that wrote
 * data/x_10000.bin, generate_rhs.cpp:37,43): x = -1 + 2*rand()/RAND_MAX. */
void or_park_miller(int64_t n, double *out) {
    uint64_t s = 1;
    const uint64_t M = 2147483647ULL;
    for (int64_t i = 0; i < n; i++) {
        s = (s * 16807ULL) % M;
        out[i] = -1.0 + 2.0 * (double)s / (double)M;
    }
}

static inline uint64_t splitmix64_next(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void or_splitmix_uniform(uint64_t seed, int64_t n, double *out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) {
        double u = (double)(splitmix64_next(&s) >> 11) * (1.0 / 9007199254740992.0);
        out[i] = 2.0 * u - 1.0;
    }
}

/* Gauss-Legendre nodes by Newton's method on P_n (extended precision),
 * emitted in the order of Gauss_Legendre_Nodes_and_Weights.hpp: for odd n
 * the centre node first, then (-x, +x) pairs from the centre outward. */
int or_gauss_legendre(int n, double *nodes, double *weights) {
    if (n < 1 || n > 1024) return -1;
    int half = n / 2;
    long double *pos = (long double *)malloc(sizeof(long double) * (size_t)(half + 1));
    long double *pw = (long double *)malloc(sizeof(long double) * (size_t)(half + 1));
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 1; i <= half; i++) {
        long double x = cosl(pi * ((long double)i - 0.25L) / ((long double)n + 0.5L));
        long double dp = 0;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1, p1 = x;
            for (int k = 2; k <= n; k++) {
                long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
            long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-19L) break;
        }
        {
            long double p0 = 1, p1 = x;
            for (int k = 2; k <= n; k++) {
                long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
        }
        /* i = 1 is the largest root; store ascending */
        pos[half - i] = x;
        pw[half - i] = 2.0L / ((1 - x * x) * dp * dp);
    }
    int o = 0;
    if (n % 2 == 1) {
        long double p0 = 1, p1 = 0;
        for (int k = 2; k <= n; k++) {
            long double p2 = ((2 * k - 1) * 0.0L * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        long double dp = n * (0.0L * p1 - p0) / (0.0L - 1);
        nodes[o] = 0.0;
        weights[o] = (double)(2.0L / (dp * dp));
        o++;
    }
    for (int i = 0; i < half; i++) {
        nodes[o] = (double)(-pos[i]);
        weights[o] = (double)pw[i];
        o++;
        nodes[o] = (double)pos[i];
        weights[o] = (double)pw[i];
        o++;
    }
    free(pos);
    free(pw);
    return 0;
}

/* integral4d.hpp:11-62: tensor Gauss-Legendre quadrature of
 * -exp(-|x|)/(4 pi) over the box [a, b] in four dimensions. */
double or_quadruple_integral(const double *a, const double *b, int ngauss) {
    double *x = (double *)malloc(sizeof(double) * (size_t)ngauss);
    double *w = (double *)malloc(sizeof(double) * (size_t)ngauss);
    or_gauss_legendre(ngauss, x, w);
    const double PI = 4.0 * atan(1.0);
    const double c = -1.0 / (4 * PI);
    double cx = 0.5 * (b[0] - a[0]), cy = 0.5 * (b[1] - a[1]);
    double cz = 0.5 * (b[2] - a[2]), cw = 0.5 * (b[3] - a[3]);
    double Lx = 0.5 * (b[0] + a[0]), Ly = 0.5 * (b[1] + a[1]);
    double Lz = 0.5 * (b[2] + a[2]), Lw = 0.5 * (b[3] + a[3]);
    double result = 0.0;
    for (int i = 0; i < ngauss; i++) {
        double tx = cx * x[i] + Lx;
        for (int j = 0; j < ngauss; j++) {
            double ty = cy * x[j] + Ly;
            for (int k = 0; k < ngauss; k++) {
                double tz = cz * x[k] + Lz;
                for (int l = 0; l < ngauss; l++) {
                    double tw = cw * x[l] + Lw;
                    double f = c * exp(-sqrt(tx * tx + ty * ty + tz * tz + tw * tw));
                    result += w[i] * w[j] * w[k] * w[l] * f;
                }
            }
        }
    }
    result *= (cx * cy * cz * cw);
    free(x);
    free(w);
    return result;
}

/* ------------------------------------------------------------------ */
/* entries                                                             */
/* ------------------------------------------------------------------ */

/* user_kernel.hpp:50-58,105-113,171-177 with points_dt.hpp:31-39.
 * The distance sum is accumulated with separate multiply and add (no
 * contraction); the device kernels use the same sequence. */
static inline double kernel_value(const or_kernel_t *k, const double *pts,
                                  int64_t i, int64_t j) {
    if (i == j) return k->diag;
    const int nd = k->ndim;
    const double *a = pts + i * nd, *b = pts + j * nd;
    double s = 0.0;
    for (int d = 0; d < nd; d++) {
        double t = a[d] - b[d];
        double t2 = t * t;
        s = s + t2;
    }
    double r = sqrt(s);
    double f;
    switch (k->family) {
    case OR_K_LOG: f = log(r); break;
    case OR_K_INV: f = 1.0 / r; break;
    case OR_K_CUSTOM: f = k->fn(r, k->params); break;
    default: f = exp(-r); break;
    }
    double v = k->scale * f;
    if (k->weight) v = v * k->weight[j];
    return v;
}

double or_f_yukawa(double r, const double *p) { return p[0] * exp(-r) / r; }
double or_f_matern32(double r, const double *p) { return (1.0 + p[0] * r) * exp(-p[0] * r); }
double or_f_gauss(double r, const double *p) { return exp(-r * r / p[0]); }

double or_entry(const or_kernel_t *k, const double *pts, int64_t i, int64_t j) {
    return kernel_value(k, pts, i, j);
}

/* main.cpp:42-57 / generate_rhs.cpp:44-51: y_i = K(i, :) . x */
void or_dense_matvec(const or_kernel_t *k, const double *pts, int64_t n,
                     const double *x, double *y, int nthreads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < n; i++) {
        double acc = 0.0;
        for (int64_t j = 0; j < n; j++) acc += kernel_value(k, pts, i, j) * x[j];
        y[i] = acc;
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------ */
/* tree: HDD_tree.hpp:23-74, HDD_clusters.hpp:22-162, HDD_node.hpp:24-31,150-180 */
/* ------------------------------------------------------------------ */

typedef struct {
    double x1[OR_MAXD], x2[OR_MAXD];
    double L;                 /* HDD_clusters.hpp:27,33-34: max box side */
    double c[OR_MAXD];        /* HDD_clusters.hpp:70-79 */
    int64_t n;
    int64_t *idx;             /* ascending point ids */
    int64_t parent;
    int64_t nnbr, nil;
    int64_t *nbr, *il;
} or_node_t;

struct or_tree {
    int ndim, nlevels, dprime;
    int64_t npts;
    const double *pts;
    int64_t *lsize;
    or_node_t **lv;
};

static void node_set_box(or_node_t *nd, int ndim, const double *x1, const double *x2) {
    nd->L = fabs(x1[0] - x2[0]);
    for (int i = 0; i < ndim; i++) {
        nd->x1[i] = x1[i];
        nd->x2[i] = x2[i];
        if (fabs(x1[i] - x2[i]) > nd->L) nd->L = fabs(x1[i] - x2[i]);
    }
    for (int i = 0; i < ndim; i++) nd->c[i] = 0.5 * (nd->x1[i] + nd->x2[i]);
}

/* HDD_clusters.hpp:142-162 */
static int itype(int ndim, const or_node_t *A, const or_node_t *B) {
    double L = A->L > B->L ? A->L : B->L;
    double s = 0.0;
    for (int i = 0; i < ndim; i++) s += (A->c[i] - B->c[i]) * (A->c[i] - B->c[i]);
    double dp = sqrt(s) / L;
    dp *= dp;
    if ((double)ndim + 1 > dp) return (int)((double)ndim - round(dp));
    return -1;
}

/* HDD_node.hpp:24-31 with points_dt.hpp:20-28 */
static int admissible(int ndim, int dprime, const or_node_t *self, const or_node_t *o) {
    double dist = fabs(self->c[0] - o->c[0]);
    for (int i = 0; i < ndim; i++) {
        double t = fabs(self->c[i] - o->c[i]);
        if (t > dist) dist = t;
    }
    if (dist < 1.5 * self->L)
        if (itype(ndim, self, o) > dprime) return 0;
    return 1;
}

static void push_list(int64_t **arr, int64_t *len, int64_t *cap, int64_t v) {
    if (*len == *cap) {
        *cap = *cap ? 2 * *cap : 16;
        *arr = (int64_t *)realloc(*arr, sizeof(int64_t) * (size_t)*cap);
    }
    (*arr)[(*len)++] = v;
}

or_tree_t *or_tree_build(int ndim, int64_t n, const double *pts,
                         const double *lo, const double *hi, int nmax, int dprime) {
    if (ndim < 1 || ndim > OR_MAXD) return NULL;
    or_tree_t *t = (or_tree_t *)calloc(1, sizeof(or_tree_t));
    t->ndim = ndim;
    t->npts = n;
    t->pts = pts;
    t->dprime = dprime;
    int cap_levels = 64;
    t->lsize = (int64_t *)calloc((size_t)cap_levels, sizeof(int64_t));
    t->lv = (or_node_t **)calloc((size_t)cap_levels, sizeof(or_node_t *));
    /* root: HDD_tree.hpp:27-39 */
    t->lv[0] = (or_node_t *)calloc(1, sizeof(or_node_t));
    t->lsize[0] = 1;
    or_node_t *root = &t->lv[0][0];
    node_set_box(root, ndim, lo, hi);
    root->n = n;
    root->idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) root->idx[i] = i;
    root->parent = -1;
    int level = 1;
    int64_t nlevel = n;
    const int64_t nch = (int64_t)1 << ndim;
    while (nlevel > nmax && level < cap_levels) { /* HDD_tree.hpp:43-69 */
        int64_t np = t->lsize[level - 1];
        t->lsize[level] = np * nch;
        t->lv[level] = (or_node_t *)calloc((size_t)(np * nch), sizeof(or_node_t));
        for (int64_t p = 0; p < np; p++) {
            or_node_t *par = &t->lv[level - 1][p];
            /* level_clustering: successive binary splits, dims 0..ndim-1
             * (HDD_clusters.hpp:80-133); child q has bit (ndim-1-d) = side in dim d */
            double mid[OR_MAXD];
            for (int d = 0; d < ndim; d++) mid[d] = 0.5 * (par->x1[d] + par->x2[d]);
            for (int64_t q = 0; q < nch; q++) {
                or_node_t *ch = &t->lv[level][p * nch + q];
                double x1[OR_MAXD], x2[OR_MAXD];
                for (int d = 0; d < ndim; d++) {
                    int bit = (int)((q >> (ndim - 1 - d)) & 1);
                    x1[d] = bit ? mid[d] : par->x1[d];
                    x2[d] = bit ? par->x2[d] : mid[d];
                }
                node_set_box(ch, ndim, x1, x2);
                ch->parent = p;
                ch->idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(par->n > 0 ? par->n : 1));
                ch->n = 0;
            }
            for (int64_t k = 0; k < par->n; k++) {
                int64_t pi = par->idx[k];
                int64_t q = 0;
                for (int d = 0; d < ndim; d++) {
                    int side = (pts[pi * ndim + d] < mid[d]) ? 0 : 1;
                    q = (q << 1) | side;
                }
                or_node_t *ch = &t->lv[level][p * nch + q];
                ch->idx[ch->n++] = pi;
            }
        }
        /* HDD_tree.hpp:62-68 */
        or_node_t *L = t->lv[level];
        nlevel = L[0].n;
        for (int64_t j = 0; j < t->lsize[level]; j++) {
            or_node_t *me = &L[j];
            or_node_t *par = &t->lv[level - 1][me->parent];
            int64_t ncap = 0, icap = 0;
            /* siblings (HDD_node.hpp:156-165) */
            for (int64_t s = 0; s < nch; s++) {
                int64_t sj = me->parent * nch + s;
                if (sj == j) continue;
                if (admissible(ndim, dprime, me, &L[sj])) push_list(&me->il, &me->nil, &icap, sj);
                else push_list(&me->nbr, &me->nnbr, &ncap, sj);
            }
            /* children of the parent's neighbours (HDD_node.hpp:167-176) */
            for (int64_t a = 0; a < par->nnbr; a++) {
                int64_t pn = par->nbr[a];
                for (int64_t s = 0; s < nch; s++) {
                    int64_t sj = pn * nch + s;
                    if (admissible(ndim, dprime, me, &L[sj])) push_list(&me->il, &me->nil, &icap, sj);
                    else push_list(&me->nbr, &me->nnbr, &ncap, sj);
                }
            }
            if (me->n > nlevel) nlevel = me->n;
        }
        level++;
    }
    t->nlevels = level;
    return t;
}

void or_tree_free(or_tree_t *t) {
    if (!t) return;
    for (int l = 0; l < t->nlevels; l++) {
        for (int64_t j = 0; j < t->lsize[l]; j++) {
            free(t->lv[l][j].idx);
            free(t->lv[l][j].nbr);
            free(t->lv[l][j].il);
        }
        free(t->lv[l]);
    }
    free(t->lv);
    free(t->lsize);
    free(t);
}

int or_tree_nlevels(const or_tree_t *t) { return t->nlevels; }
int64_t or_tree_level_size(const or_tree_t *t, int level) { return t->lsize[level]; }
int64_t or_tree_node_npts(const or_tree_t *t, int level, int64_t j) { return t->lv[level][j].n; }
void or_tree_node_points(const or_tree_t *t, int level, int64_t j, int64_t *out) {
    memcpy(out, t->lv[level][j].idx, sizeof(int64_t) * (size_t)t->lv[level][j].n);
}
int64_t or_tree_list(const or_tree_t *t, int level, int64_t j, int which, int64_t *out) {
    const or_node_t *nd = &t->lv[level][j];
    int64_t len = which ? nd->nil : nd->nnbr;
    if (out && len) memcpy(out, which ? nd->il : nd->nbr, sizeof(int64_t) * (size_t)len);
    return len;
}
int or_interaction_type(const or_tree_t *t, int level, int64_t a, int64_t b) {
    return itype(t->ndim, &t->lv[level][a], &t->lv[level][b]);
}

/* ------------------------------------------------------------------ */
/* ACA: LowRank_matrix.hpp:64-377 (and 379-724, whose iteration is identical) */
/* ------------------------------------------------------------------ */

uint64_t or_block_seed(int level, int64_t row_node, int64_t col_node) {
    uint64_t s = 0x243F6A8885A308D3ULL ^ ((uint64_t)(level + 1) * 0x9E3779B97F4A7C15ULL);
    s ^= (uint64_t)row_node * 0xBF58476D1CE4E5B9ULL;
    s ^= (uint64_t)col_node * 0x94D049BB133111EBULL;
    return s;
}

/* kernel_function.hpp:55-68 over the unused indices (ascending, strict >);
 * an empty set yields index 0 as libstdc++'s *begin() of an empty set does. */
static void max_abs(const double *v, const char *used, int len, double *maxv, int *idx) {
    int first = -1;
    for (int i = 0; i < len; i++)
        if (!used[i]) { first = i; break; }
    if (first < 0) { *idx = 0; *maxv = v[0]; return; }
    int best = first;
    double bm = v[first];
    for (int i = first + 1; i < len; i++) {
        if (used[i]) continue;
        if (fabs(v[i]) > fabs(bm)) { best = i; bm = v[i]; }
    }
    *idx = best;
    *maxv = bm;
}

static int kth_unused(const char *used, int len, int64_t k) {
    for (int i = 0; i < len; i++) {
        if (used[i]) continue;
        if (k == 0) return i;
        k--;
    }
    return 0;
}
static int first_unused(const char *used, int len) {
    for (int i = 0; i < len; i++) if (!used[i]) return i;
    return 0;
}
static int last_unused(const char *used, int len) {
    for (int i = len - 1; i >= 0; i--) if (!used[i]) return i;
    return 0;
}
static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}
/* LowRank_matrix.hpp:181-198: midpoint of the widest gap of the used indices */
static int mid_gap(const int *ind, int cnt, int *tmp) {
    memcpy(tmp, ind, sizeof(int) * (size_t)cnt);
    qsort(tmp, (size_t)cnt, sizeof(int), cmp_int);
    unsigned int mx = 0;
    int at = 0;
    for (int i = 0; i + 1 < cnt; i++) {
        unsigned int d = (unsigned int)(tmp[i + 1] - tmp[i]);
        if (d > mx) { at = i; mx = d; }
    }
    return tmp[at] + (int)(mx / 2);
}

/* Storage of the factors grows on demand (*U, *V realloc'd, *cap updated);
 * row_piv / col_piv, when given, must hold min(m, n) entries. */
static int aca_grow(const or_kernel_t *k, const double *pts,
                    const int64_t *rows, int m, const int64_t *cols, int n,
                    double eps, uint64_t seed, int *capp, double **Up, double **Vp,
                    int growable, int32_t *row_piv, int32_t *col_piv, or_aca_info_t *info) {
    const double tol = eps;
    int cap = *capp;
    double *U = *Up, *V = *Vp;
    const int max_tries = 10;
    int mn = m < n ? m : n;
    char *rused = (char *)calloc((size_t)m, 1);
    char *cused = (char *)calloc((size_t)n, 1);
    int rrem = m, crem = n;
    int icap = mn + 4 * max_tries + 8;
    int *row_ind = (int *)malloc(sizeof(int) * (size_t)icap);
    int *col_ind = (int *)malloc(sizeof(int) * (size_t)icap);
    int *tmp = (int *)malloc(sizeof(int) * (size_t)icap);
    double *row = (double *)malloc(sizeof(double) * (size_t)n);
    double *col = (double *)malloc(sizeof(double) * (size_t)m);
    int nri = 0, nci = 0;
    uint64_t rng = seed;
    int retries = 0;
    int64_t rows_ev = 0, cols_ev = 0; /* rows and columns evaluated */

    row_ind[nri++] = 0; /* :96-97 */
    rused[0] = 1;
    rrem--;
    double matrix_norm = 0.0, maxv = 0.0, gamma;
    double row_norm = 0.0, col_norm = 0.0;
    int pivot = 0, r = 0;

#define EVAL_ROW(ri)                                                          \
    do {                                                                      \
        rows_ev++;                                                            \
        for (int j_ = 0; j_ < n; j_++) {                                      \
            double v_ = kernel_value(k, pts, rows[(ri)], cols[j_]);           \
            for (int q_ = 0; q_ < r; q_++)                                    \
                v_ = v_ - U[(size_t)q_ * m + (ri)] * V[(size_t)q_ * n + j_];  \
            row[j_] = v_;                                                     \
        }                                                                     \
    } while (0)
#define EVAL_COL(cj)                                                          \
    do {                                                                      \
        cols_ev++;                                                            \
        for (int i_ = 0; i_ < m; i_++) {                                      \
            double v_ = kernel_value(k, pts, rows[i_], cols[(cj)]);           \
            for (int q_ = 0; q_ < r; q_++)                                    \
                v_ = v_ - V[(size_t)q_ * n + (cj)] * U[(size_t)q_ * m + i_];  \
            col[i_] = v_;                                                     \
        }                                                                     \
    } while (0)

    do {
        int ri = row_ind[nri - 1];
        EVAL_ROW(ri); /* :126-131 */
        max_abs(row, cused, n, &maxv, &pivot);
        int count = 0;
        int at_end = 0, use_rand = 1;
        while (fabs(maxv) < tol && count < max_tries && crem > 0 && rrem > 0) { /* :144-216 */
            nri--;
            int nr;
            if (r < 3) {
                nr = at_end ? last_unused(rused, m) : first_unused(rused, m);
                at_end = !at_end;
            } else {
                if (use_rand) nr = kth_unused(rused, m, (int64_t)(splitmix64_next(&rng) % (uint64_t)rrem));
                else nr = mid_gap(row_ind, nri, tmp);
                use_rand = !use_rand;
            }
            row_ind[nri++] = nr;
            if (!rused[nr]) { rused[nr] = 1; rrem--; }
            EVAL_ROW(nr);
            max_abs(row, cused, n, &maxv, &pivot);
            count++;
            retries++;
        }
        if (count == max_tries || crem == 0 || rrem == 0) break; /* :220-225 */
        count = 0;
        col_ind[nci++] = pivot;
        if (!cused[pivot]) { cused[pivot] = 1; crem--; }
        gamma = 1.0 / maxv; /* :233 */
        int cj = col_ind[nci - 1];
        EVAL_COL(cj); /* :237-241 */
        max_abs(col, rused, m, &maxv, &pivot);
        while (fabs(maxv) < tol && count < max_tries && crem > 0 && rrem > 0) { /* :245-316 */
            nci--;
            int nc;
            if (nci < 3) {
                /* :256-258: *remaining_col_ind.end() reads the set's size */
                nc = at_end ? crem : first_unused(cused, n);
                at_end = !at_end;
            } else {
                if (use_rand) nc = kth_unused(cused, n, (int64_t)(splitmix64_next(&rng) % (uint64_t)crem));
                else nc = mid_gap(col_ind, nci, tmp);
                use_rand = !use_rand;
            }
            col_ind[nci++] = nc;
            if (!cused[nc]) { cused[nc] = 1; crem--; }
            EVAL_COL(nc);
            max_abs(col, rused, m, &maxv, &pivot);
            count++;
            retries++;
        }
        row_ind[nri++] = pivot; /* :318-319 */
        if (!rused[pivot]) { rused[pivot] = 1; rrem--; }
        if (r >= cap) {
            if (!growable) break; /* caller's storage is full (never for cap >= min(m,n)) */
            int nc = 2 * cap < mn ? 2 * cap : mn;
            if (nc < 8) nc = mn < 8 ? mn : 8;
            U = (double *)realloc(U, sizeof(double) * (size_t)m * (size_t)nc);
            V = (double *)realloc(V, sizeof(double) * (size_t)n * (size_t)nc);
            cap = nc;
            *Up = U;
            *Vp = V;
            *capp = cap;
        }
        double rs = 0.0, cs = 0.0;
        for (int j = 0; j < n; j++) { V[(size_t)r * n + j] = row[j]; rs += row[j] * row[j]; }
        for (int i = 0; i < m; i++) { U[(size_t)r * m + i] = gamma * col[i]; cs += col[i] * col[i]; }
        row_norm = sqrt(rs);
        col_norm = sqrt(cs);
        matrix_norm += fabs(gamma * gamma * rs * cs); /* :333 */
        for (int q = 0; q < r; q++) { /* :335-338 */
            double du = 0.0, dv = 0.0;
            for (int i = 0; i < m; i++) du += U[(size_t)q * m + i] * U[(size_t)r * m + i];
            for (int j = 0; j < n; j++) dv += V[(size_t)q * n + j] * V[(size_t)r * n + j];
            matrix_norm += 2.0 * fabs(du) * fabs(dv);
        }
        if (row_piv) row_piv[r] = row_ind[r];
        if (col_piv) col_piv[r] = col_ind[r];
        r++;
    } while ((double)(r * (m + n)) * row_norm * col_norm > fabs(maxv) * tol * matrix_norm &&
             r < mn); /* :341-344 */
#undef EVAL_ROW
#undef EVAL_COL
    int dense = (r >= mn - 1); /* :348 */
    if (info) {
        info->steps = r;
        info->dense = dense;
        info->rank = dense ? mn : r;
        info->retries = retries;
        info->entries = rows_ev * n + cols_ev * m;
    }
    free(rused); free(cused); free(row_ind); free(col_ind); free(tmp); free(row); free(col);
    return dense ? mn : r;
}

int or_aca(const or_kernel_t *k, const double *pts,
           const int64_t *rows, int m, const int64_t *cols, int n,
           double eps, uint64_t seed, int cap, double *U, double *V,
           int32_t *row_piv, int32_t *col_piv, or_aca_info_t *info) {
    return aca_grow(k, pts, rows, m, cols, n, eps, seed, &cap, &U, &V, 0, row_piv, col_piv, info);
}

/* ------------------------------------------------------------------ */
/* column-pivoted Householder QR solve                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int n, rank;
    double *qr;   /* n x n column-major: R above the diagonal, reflectors below */
    double *tau;
    int *perm;
} or_cpqr_t;

static void cpqr_factor(or_cpqr_t *f, int n, const double *A) {
    f->n = n;
    f->qr = (double *)malloc(sizeof(double) * (size_t)n * n);
    f->tau = (double *)calloc((size_t)n, sizeof(double));
    f->perm = (int *)malloc(sizeof(int) * (size_t)n);
    memcpy(f->qr, A, sizeof(double) * (size_t)n * n);
    double *a = f->qr;
    for (int j = 0; j < n; j++) f->perm[j] = j;
    double maxpiv = 0.0;
    for (int kk = 0; kk < n; kk++) {
        int best = kk;
        double bn = -1.0;
        for (int j = kk; j < n; j++) {
            double s = 0.0;
            for (int i = kk; i < n; i++) s += a[(size_t)j * n + i] * a[(size_t)j * n + i];
            if (s > bn) { bn = s; best = j; }
        }
        if (best != kk) {
            for (int i = 0; i < n; i++) {
                double t = a[(size_t)kk * n + i];
                a[(size_t)kk * n + i] = a[(size_t)best * n + i];
                a[(size_t)best * n + i] = t;
            }
            int t = f->perm[kk]; f->perm[kk] = f->perm[best]; f->perm[best] = t;
        }
        double *x = a + (size_t)kk * n;
        double alpha = x[kk], s = 0.0;
        for (int i = kk + 1; i < n; i++) s += x[i] * x[i];
        double beta, tau = 0.0;
        if (s == 0.0) {
            beta = alpha;
        } else {
            beta = -copysign(sqrt(alpha * alpha + s), alpha);
            tau = (beta - alpha) / beta;
            double sc = 1.0 / (alpha - beta);
            for (int i = kk + 1; i < n; i++) x[i] *= sc;
        }
        x[kk] = beta;
        f->tau[kk] = tau;
        if (kk == 0) maxpiv = fabs(beta);
        for (int j = kk + 1; j < n && tau != 0.0; j++) {
            double *y = a + (size_t)j * n;
            double d = y[kk];
            for (int i = kk + 1; i < n; i++) d += x[i] * y[i];
            d *= tau;
            y[kk] -= d;
            for (int i = kk + 1; i < n; i++) y[i] -= d * x[i];
        }
    }
    double thr = DBL_EPSILON * (double)n * maxpiv;
    f->rank = 0;
    for (int kk = 0; kk < n; kk++)
        if (fabs(a[(size_t)kk * n + kk]) > thr) f->rank = kk + 1;
        else break;
}

static void cpqr_solve(const or_cpqr_t *f, const double *b, double *x, double *work) {
    int n = f->n;
    const double *a = f->qr;
    memcpy(work, b, sizeof(double) * (size_t)n);
    for (int kk = 0; kk < n; kk++) {
        if (f->tau[kk] == 0.0) continue;
        const double *v = a + (size_t)kk * n;
        double d = work[kk];
        for (int i = kk + 1; i < n; i++) d += v[i] * work[i];
        d *= f->tau[kk];
        work[kk] -= d;
        for (int i = kk + 1; i < n; i++) work[i] -= d * v[i];
    }
    for (int i = f->rank; i < n; i++) work[i] = 0.0;
    for (int i = f->rank - 1; i >= 0; i--) {
        double s = work[i];
        for (int j = i + 1; j < f->rank; j++) s -= a[(size_t)j * n + i] * work[j];
        work[i] = s / a[(size_t)i * n + i];
    }
    for (int i = 0; i < n; i++) x[f->perm[i]] = work[i];
}

static void cpqr_free(or_cpqr_t *f) { free(f->qr); free(f->tau); free(f->perm); }

int or_cpqr_solve(int n, const double *A, const double *b, double *x) {
    or_cpqr_t f;
    cpqr_factor(&f, n, A);
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    cpqr_solve(&f, b, x, w);
    free(w);
    int r = f.rank;
    cpqr_free(&f);
    return r;
}

/* ------------------------------------------------------------------ */
/* operator: HDD_node.hpp:118-147 (assembly), :223-254 (node matvec)    */
/* ------------------------------------------------------------------ */

enum { BLK_LOWRANK = 0, BLK_DENSE = 1, BLK_SKEL = 2 };

typedef struct {
    int level;
    int64_t row_node, col_node;
    int m, n, rank, kind;
    double *L, *R;        /* lowrank: U (m x r), V (n x r) col-major; dense: K (m x n) col-major */
    int64_t *rb, *cb;     /* skeleton bases (point ids) */
    or_cpqr_t qr;
    int built;
    int64_t aca_entries;  /* entries its ACA evaluated */
} or_block_t;

typedef struct {
    int level;
    int64_t node;
    double *self;         /* leaf: n x n */
    double **nbr;         /* leaf: one dense block per neighbour (NULL if empty) */
    int64_t blk0, nblk;   /* far blocks of this node in op->blk */
    int built;
} or_opnode_t;

struct or_op {
    or_tree_t *t;
    or_kernel_t k;
    const double *pts;
    double eps;
    int sys_size;
    int64_t nblk;
    or_block_t *blk;
    int64_t nnodes;
    or_opnode_t *nodes;
    double init_seconds;
};

static double *dense_block(const or_kernel_t *k, const double *pts,
                           const int64_t *rows, int m, const int64_t *cols, int n) {
    double *D = (double *)malloc(sizeof(double) * (size_t)m * (size_t)n + 8);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) D[(size_t)j * m + i] = kernel_value(k, pts, rows[i], cols[j]);
    return D;
}

/* LowRank_matrix.hpp:22-36 (dispatch), :64-377 / :379-724 (ACA), :348-363
 * and :662-680 (dense fallback), :685-698 (skeleton bases + QR). */
static void build_block(or_op_t *op, or_block_t *b) {
    or_tree_t *t = op->t;
    const or_node_t *rn = &t->lv[b->level][b->row_node];
    const or_node_t *cn = &t->lv[b->level][b->col_node];
    int m = (int)rn->n, n = (int)cn->n;
    int mn = m < n ? m : n;
    int cap = mn < 32 ? mn : 32;
    double *U = (double *)malloc(sizeof(double) * (size_t)m * (size_t)cap + 8);
    double *V = (double *)malloc(sizeof(double) * (size_t)n * (size_t)cap + 8);
    int32_t *rp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(mn + 1));
    int32_t *cp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(mn + 1));
    or_aca_info_t info;
    uint64_t seed = or_block_seed(b->level, b->row_node, b->col_node);
    aca_grow(&op->k, op->pts, rn->idx, m, cn->idx, n, op->eps, seed, &cap, &U, &V, 1, rp, cp, &info);
    b->m = m;
    b->n = n;
    b->rank = info.rank;
    b->aca_entries = info.entries;
    b->L = b->R = NULL;
    b->rb = b->cb = NULL;
    if (info.dense) {
        b->kind = BLK_DENSE;
        b->L = dense_block(&op->k, op->pts, rn->idx, m, cn->idx, n);
        free(U);
        free(V);
    } else if (op->sys_size >= 0 && n > op->sys_size) {
        b->kind = BLK_SKEL;
        int r = info.rank;
        b->rb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(r + 1));
        b->cb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(r + 1));
        for (int i = 0; i < r; i++) {
            b->rb[i] = rn->idx[rp[i]];
            b->cb[i] = cn->idx[cp[i]];
        }
        if (r > 0) {
            double *A = dense_block(&op->k, op->pts, b->rb, r, b->cb, r);
            cpqr_factor(&b->qr, r, A);
            free(A);
        }
        free(U);
        free(V);
    } else {
        /* keep the r factor columns only (U, V are column-major, m x cap and
         * n x cap, cap >= r: the first r columns are contiguous) */
        b->kind = BLK_LOWRANK;
        const int r = info.rank;
        double *L = (double *)realloc(U, sizeof(double) * (size_t)m * (size_t)r + 8);
        double *R = (double *)realloc(V, sizeof(double) * (size_t)n * (size_t)r + 8);
        b->L = L ? L : U;
        b->R = R ? R : V;
    }
    free(rp);
    free(cp);
    b->built = 1;
}

static void build_node(or_op_t *op, or_opnode_t *nd) {
    or_tree_t *t = op->t;
    const or_node_t *me = &t->lv[nd->level][nd->node];
    nd->self = NULL;
    nd->nbr = NULL;
    if (me->n == 0 || nd->level != t->nlevels - 1) { nd->built = 1; return; }
    nd->self = dense_block(&op->k, op->pts, me->idx, (int)me->n, me->idx, (int)me->n);
    nd->nbr = (double **)calloc((size_t)(me->nnbr + 1), sizeof(double *));
    for (int64_t a = 0; a < me->nnbr; a++) {
        const or_node_t *o = &t->lv[nd->level][me->nbr[a]];
        if (o->n) nd->nbr[a] = dense_block(&op->k, op->pts, me->idx, (int)me->n, o->idx, (int)o->n);
    }
    nd->built = 1;
}

or_op_t *or_op_assemble_sample(or_tree_t *t, const or_kernel_t *k, const double *pts,
                               double eps, int sys_size, int nthreads,
                               const int64_t *stride, const int64_t *offset) {
    or_op_t *op = (or_op_t *)calloc(1, sizeof(or_op_t));
    op->t = t;
    op->k = *k;
    op->pts = pts;
    op->eps = eps;
    op->sys_size = sys_size;
    int64_t nn = 0, nb = 0;
    for (int l = 0; l < t->nlevels; l++) nn += t->lsize[l];
    op->nnodes = nn;
    op->nodes = (or_opnode_t *)calloc((size_t)nn, sizeof(or_opnode_t));
    /* block order: level, node, interaction-list index (HDD_tree.hpp:76-83,
     * HDD_node.hpp:127-130); only blocks whose two nodes are non-empty */
    for (int l = 0; l < t->nlevels; l++)
        for (int64_t j = 0; j < t->lsize[l]; j++) {
            const or_node_t *me = &t->lv[l][j];
            if (!me->n) continue;
            for (int64_t a = 0; a < me->nil; a++)
                if (t->lv[l][me->il[a]].n) nb++;
        }
    op->nblk = nb;
    op->blk = (or_block_t *)calloc((size_t)(nb + 1), sizeof(or_block_t));
    int64_t bi = 0, ni = 0;
    for (int l = 0; l < t->nlevels; l++)
        for (int64_t j = 0; j < t->lsize[l]; j++, ni++) {
            or_opnode_t *nd = &op->nodes[ni];
            nd->level = l;
            nd->node = j;
            nd->blk0 = bi;
            int sel = 1;
            if (stride) sel = stride[l] > 0 && (j % stride[l]) == offset[l];
            nd->built = sel ? 0 : -1;
            const or_node_t *me = &t->lv[l][j];
            if (me->n) {
                for (int64_t a = 0; a < me->nil; a++) {
                    if (!t->lv[l][me->il[a]].n) continue;
                    or_block_t *b = &op->blk[bi++];
                    b->level = l;
                    b->row_node = j;
                    b->col_node = me->il[a];
                    b->m = (int)me->n;
                    b->n = (int)t->lv[l][me->il[a]].n;
                    b->built = sel ? 0 : -1;
                }
            }
            nd->nblk = bi - nd->blk0;
        }
    double t0 = now_s();
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
#pragma omp for schedule(dynamic, 1) nowait
        for (int64_t i = 0; i < nb; i++)
            if (op->blk[i].built == 0) build_block(op, &op->blk[i]);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < nn; i++)
            if (op->nodes[i].built == 0) build_node(op, &op->nodes[i]);
    }
#else
    for (int64_t i = 0; i < nb; i++)
        if (op->blk[i].built == 0) build_block(op, &op->blk[i]);
    for (int64_t i = 0; i < nn; i++)
        if (op->nodes[i].built == 0) build_node(op, &op->nodes[i]);
#endif
    (void)nthreads;
    op->init_seconds = now_s() - t0;
    return op;
}

or_op_t *or_op_assemble(or_tree_t *t, const or_kernel_t *k, const double *pts,
                        double eps, int sys_size, int nthreads) {
    return or_op_assemble_sample(t, k, pts, eps, sys_size, nthreads, NULL, NULL);
}

void or_op_free(or_op_t *op) {
    if (!op) return;
    for (int64_t i = 0; i < op->nblk; i++) {
        or_block_t *b = &op->blk[i];
        free(b->L); free(b->R); free(b->rb); free(b->cb);
        if (b->kind == BLK_SKEL && b->rank > 0 && b->built == 1) cpqr_free(&b->qr);
    }
    for (int64_t i = 0; i < op->nnodes; i++) {
        or_opnode_t *nd = &op->nodes[i];
        if (nd->nbr) {
            const or_node_t *me = &op->t->lv[nd->level][nd->node];
            for (int64_t a = 0; a < me->nnbr; a++) free(nd->nbr[a]);
            free(nd->nbr);
        }
        free(nd->self);
    }
    free(op->blk);
    free(op->nodes);
    free(op);
}

/* y += A x for column-major A (m x n) */
static void gemv_acc(const double *A, int m, int n, const double *x, double *y) {
    for (int j = 0; j < n; j++) {
        double xj = x[j];
        const double *a = A + (size_t)j * m;
        for (int i = 0; i < m; i++) y[i] += a[i] * xj;
    }
}

/* LowRank_matrix.hpp:40-63 */
static void block_apply(const or_op_t *op, const or_block_t *b, const double *q, double *pot) {
    int m = b->m, n = b->n, r = b->rank;
    if (b->built != 1) return;
    if (b->kind == BLK_DENSE) { gemv_acc(b->L, m, n, q, pot); return; }
    if (r == 0) return;
    double *tv = (double *)calloc((size_t)r, sizeof(double));
    if (b->kind == BLK_LOWRANK) {
        for (int kk = 0; kk < r; kk++) {
            const double *v = b->R + (size_t)kk * n;
            double s = 0.0;
            for (int j = 0; j < n; j++) s += v[j] * q[j];
            tv[kk] = s;
        }
        gemv_acc(b->L, m, r, tv, pot);
    } else {
        const or_tree_t *t = op->t;
        const or_node_t *rn = &t->lv[b->level][b->row_node];
        const or_node_t *cn = &t->lv[b->level][b->col_node];
        double *Ar = dense_block(&op->k, op->pts, b->rb, r, cn->idx, n);
        double *Ac = dense_block(&op->k, op->pts, rn->idx, m, b->cb, r);
        double *ax = (double *)calloc((size_t)r, sizeof(double));
        double *w = (double *)malloc(sizeof(double) * (size_t)r);
        gemv_acc(Ar, r, n, q, ax);
        cpqr_solve(&b->qr, ax, tv, w);
        gemv_acc(Ac, m, r, tv, pot);
        free(Ar); free(Ac); free(ax); free(w);
    }
    free(tv);
}

/* HDD_tree.hpp:86-107 with HDD_node.hpp:82-99,223-254.  Empty nodes give 0. */
void or_op_matvec(or_op_t *op, const double *x, double *y, int nthreads) {
    or_tree_t *t = op->t;
    int64_t nn = op->nnodes;
    double **pot = (double **)calloc((size_t)nn, sizeof(double *));
    int64_t *base = (int64_t *)malloc(sizeof(int64_t) * (size_t)(t->nlevels + 1));
    base[0] = 0;
    for (int l = 0; l < t->nlevels; l++) base[l + 1] = base[l] + t->lsize[l];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < nn; i++) {
        or_opnode_t *nd = &op->nodes[i];
        const or_node_t *me = &t->lv[nd->level][nd->node];
        if (!me->n || nd->built != 1) continue;
        double *p = (double *)calloc((size_t)me->n, sizeof(double));
        double *q = (double *)malloc(sizeof(double) * (size_t)me->n);
        if (nd->self) {
            for (int64_t a = 0; a < me->n; a++) q[a] = x[me->idx[a]];
            gemv_acc(nd->self, (int)me->n, (int)me->n, q, p);
            for (int64_t a = 0; a < me->nnbr; a++) {
                const or_node_t *o = &t->lv[nd->level][me->nbr[a]];
                if (!o->n) continue;
                double *qo = (double *)malloc(sizeof(double) * (size_t)o->n);
                for (int64_t c = 0; c < o->n; c++) qo[c] = x[o->idx[c]];
                gemv_acc(nd->nbr[a], (int)me->n, (int)o->n, qo, p);
                free(qo);
            }
        }
        for (int64_t bi = nd->blk0; bi < nd->blk0 + nd->nblk; bi++) {
            const or_block_t *b = &op->blk[bi];
            const or_node_t *o = &t->lv[b->level][b->col_node];
            double *qo = (double *)malloc(sizeof(double) * (size_t)o->n);
            for (int64_t c = 0; c < o->n; c++) qo[c] = x[o->idx[c]];
            block_apply(op, b, qo, p);
            free(qo);
        }
        free(q);
        pot[i] = p;
    }
    (void)nthreads;
    memset(y, 0, sizeof(double) * (size_t)t->npts);
    for (int64_t i = 0; i < nn; i++) { /* collect in node order (HDD_tree.hpp:104-105) */
        if (!pot[i]) continue;
        const or_node_t *me = &t->lv[op->nodes[i].level][op->nodes[i].node];
        for (int64_t a = 0; a < me->n; a++) y[me->idx[a]] += pot[i][a];
        free(pot[i]);
    }
    free(pot);
    free(base);
}

/* HDD_node.hpp:183-220, HDD_matrix.hpp:69-85 */
void or_op_get_stats(const or_op_t *op, or_op_stats_t *s) {
    memset(s, 0, sizeof(*s));
    const or_tree_t *t = op->t;
    int leaf = t->nlevels - 1;
    for (int64_t j = 0; j < t->lsize[leaf]; j++) {
        const or_node_t *me = &t->lv[leaf][j];
        if (!me->n) continue;
        double w = (double)me->n;
        for (int64_t a = 0; a < me->nnbr; a++) w += (double)t->lv[leaf][me->nbr[a]].n;
        s->e_near += w * (double)me->n;
    }
    s->n_flop_ref = s->e_near;
    for (int64_t i = 0; i < op->nblk; i++) {
        const or_block_t *b = &op->blk[i];
        if (b->built != 1) continue;
        s->n_blocks++;
        s->aca_entries += (double)b->aca_entries;
        s->n_flop_ref += (double)b->rank * (double)(b->m + b->n);
        if ((int64_t)b->rank > s->max_rank) s->max_rank = b->rank;
        if (b->kind == BLK_DENSE) { s->n_dense++; s->e_far += (double)b->m * (double)b->n; }
        else if (b->kind == BLK_SKEL) { s->n_skeleton++; s->e_far += (double)b->rank * (double)(b->m + b->n); }
        else s->e_far += (double)b->rank * (double)(b->m + b->n);
    }
    s->init_seconds = op->init_seconds;
}

int64_t or_op_block_ranks(const or_op_t *op, int32_t *rank, int32_t *dense) {
    for (int64_t i = 0; i < op->nblk; i++) {
        if (rank) rank[i] = op->blk[i].rank;
        if (dense) dense[i] = op->blk[i].kind == BLK_DENSE;
    }
    return op->nblk;
}

/* ------------------------------------------------------------------ */
/* GMRES.hpp:83-146,160-303 (real branch, -DUSE_DOUBLE)                 */
/* ------------------------------------------------------------------ */

typedef void (*or_apply_fn)(void *ctx, const double *x, double *y);

static void gen_rot(double dx, double dy, double *cs, double *sn) {
    if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
    else if (fabs(dy) > fabs(dx)) {
        double temp = dx / dy;
        *sn = 1.0 / sqrt(1.0 + temp * temp);
        *cs = temp * *sn;
    } else {
        double temp = dy / dx;
        *cs = 1.0 / sqrt(1.0 + temp * temp);
        *sn = temp * *cs;
    }
}
static void app_rot(double *dx, double *dy, double cs, double sn) {
    double temp = cs * *dx + sn * *dy;
    *dy = (-sn) * *dx + cs * *dy;
    *dx = temp;
}
static double vdot(const double *a, const double *b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static void gmres_update(double *x, int64_t N, int k, const double *H, int ldh,
                         const double *s, double **v) {
    double *y = (double *)malloc(sizeof(double) * (size_t)(k + 1));
    memcpy(y, s, sizeof(double) * (size_t)(k + 1));
    for (int i = k; i >= 0; i--) {
        y[i] /= H[(size_t)i * ldh + i];
        for (int j = i - 1; j >= 0; j--) y[j] -= H[(size_t)i * ldh + j] * y[i];
    }
    for (int j = 0; j <= k; j++)
        for (int64_t a = 0; a < N; a++) x[a] += v[j][a] * y[j];
    free(y);
}

static int gmres_core(or_apply_fn A, void *ctx, int64_t N, const double *b, double *x,
                      int restart, int max_iter, double tol,
                      double *hist, int hist_cap, or_gmres_info_t *info) {
    int m = restart, i, j = 1, kk, nmv = 0;
    int ldh = restart + 1;  /* H stored column-major (restart+1) x restart */
    double *H = (double *)calloc((size_t)ldh * (size_t)restart, sizeof(double));
    double *s = (double *)calloc((size_t)(restart + 1), sizeof(double));
    double *cs = (double *)calloc((size_t)(restart + 1), sizeof(double));
    double *sn = (double *)calloc((size_t)(restart + 1), sizeof(double));
    double *w = (double *)malloc(sizeof(double) * (size_t)N);
    double *r = (double *)malloc(sizeof(double) * (size_t)N);
    int nh = 0;
    double normb = sqrt(vdot(b, b, N));
    A(ctx, x, w); nmv++;
    for (int64_t a = 0; a < N; a++) r[a] = b[a] - w[a];
    double beta = sqrt(vdot(r, r, N));
    if (normb == 0.0) normb = 1;
    double resid = beta / normb;
    int ret = -1;
    double **v = NULL;
    if (resid <= tol) { ret = 0; goto done; }
    v = (double **)calloc((size_t)(restart + 1), sizeof(double *));
    for (kk = 0; kk <= restart; kk++) v[kk] = (double *)malloc(sizeof(double) * (size_t)N);
    while (j <= max_iter) {
        for (int64_t a = 0; a < N; a++) v[0][a] = r[a] * (1.0 / beta);
        memset(s, 0, sizeof(double) * (size_t)(restart + 1));
        s[0] = beta;
        for (i = 0; i < m && j <= max_iter; i++, j++) {
            if (hist && nh < hist_cap) hist[nh++] = resid;
            A(ctx, v[i], w); nmv++;
            for (kk = 0; kk <= i; kk++) {
                double h = vdot(w, v[kk], N);
                H[(size_t)i * ldh + kk] = h;
                for (int64_t a = 0; a < N; a++) w[a] -= h * v[kk][a];
            }
            double hn = sqrt(vdot(w, w, N));
            H[(size_t)i * ldh + i + 1] = hn;
            if (hn != 0.0)
                for (int64_t a = 0; a < N; a++) v[i + 1][a] = (1.0 / hn) * w[a];
            else
                memset(v[i + 1], 0, sizeof(double) * (size_t)N);
            for (kk = 0; kk < i; kk++)
                app_rot(&H[(size_t)i * ldh + kk], &H[(size_t)i * ldh + kk + 1], cs[kk], sn[kk]);
            gen_rot(H[(size_t)i * ldh + i], H[(size_t)i * ldh + i + 1], &cs[i], &sn[i]);
            app_rot(&H[(size_t)i * ldh + i], &H[(size_t)i * ldh + i + 1], cs[i], sn[i]);
            app_rot(&s[i], &s[i + 1], cs[i], sn[i]);
            resid = fabs(s[i + 1]) / normb;
            if (resid < tol) {
                gmres_update(x, N, i, H, ldh, s, v);
                ret = j;
                goto done;
            }
        }
        gmres_update(x, N, m - 1, H, ldh, s, v);
        A(ctx, x, w); nmv++;
        for (int64_t a = 0; a < N; a++) r[a] = b[a] - w[a];
        beta = sqrt(vdot(r, r, N));
        resid = beta / normb;
        if (resid < tol) { ret = j; goto done; }
    }
    ret = max_iter;
done:
    if (hist && nh < hist_cap) hist[nh++] = resid;
    if (info) { info->iterations = ret; info->matvecs = nmv; info->resid = resid; }
    if (v) { for (kk = 0; kk <= restart; kk++) free(v[kk]); free(v); }
    free(H); free(s); free(cs); free(sn); free(w); free(r);
    return ret;
}

typedef struct { or_op_t *op; int nthreads; } op_ctx_t;
static void op_apply(void *c, const double *x, double *y) {
    op_ctx_t *oc = (op_ctx_t *)c;
    or_op_matvec(oc->op, x, y, oc->nthreads);
}
int or_gmres(or_op_t *op, const double *b, double *x, int restart, int max_iter,
             double tol, int nthreads, double *hist, int hist_cap, or_gmres_info_t *info) {
    op_ctx_t c = {op, nthreads};
    return gmres_core(op_apply, &c, op->t->npts, b, x, restart, max_iter, tol, hist, hist_cap, info);
}

typedef struct { const double *A; int64_t n; } dense_ctx_t;
static void dense_apply(void *c, const double *x, double *y) {
    dense_ctx_t *d = (dense_ctx_t *)c;
    for (int64_t i = 0; i < d->n; i++) y[i] = vdot(d->A + i * d->n, x, d->n);
}
int or_gmres_dense(const double *A, int64_t n, const double *b, double *x,
                   int restart, int max_iter, double tol, or_gmres_info_t *info) {
    dense_ctx_t c = {A, n};
    return gmres_core(dense_apply, &c, n, b, x, restart, max_iter, tol, NULL, 0, info);
}
