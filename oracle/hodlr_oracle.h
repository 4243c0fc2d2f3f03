/*
 * hodlr_oracle.h -- FP64 CPU restatement of HODLRnD's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and nothing else.  The product (libhodlrnd.so) never links, loads or
 * calls it.
 *
 * Each function restates one part of the reference (kandapva/HODLRnD) and
 * cites the file:line it follows.  Where the reference is undefined or
 * non-deterministic the oracle makes a documented choice:
 *   - rand() in the ACA retry paths (LowRank_matrix.hpp:175,274) is replaced
 *     by a SplitMix64 stream seeded per block (or_block_seed);
 *   - *std::set::end() (LowRank_matrix.hpp:258) is taken as the set's size,
 *     which is what libstdc++ returns;
 *   - empty nodes contribute nothing to the matvec (HDD_node.hpp:121-126 vs
 *     :243-250 dereference unallocated factors);
 *   - a rank-0 skeleton block contributes nothing (LowRank_matrix.hpp:694-698
 *     leaves the QR pointer unset).
 * Residual updates in ACA round the product, then the difference (no fused
 * multiply-add: the reference's Eigen row - u * v, LowRank_matrix.hpp:130,
 * 240, compiles without -mfma, and this file is built with
 * -ffp-contract=off); the GPU kernels round the same way, so both sides
 * agree bit for bit.
 */
#ifndef HODLR_ORACLE_H
#define HODLR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXD 8

/* kernel families: entry(i,j) = diag if i == j else scale * f(|x_i - x_j|) * w_j */
enum { OR_K_LOG = 0, OR_K_INV = 1, OR_K_EXP = 2, OR_K_CUSTOM = 3 };

/* f(r) of a custom family (a user kernel's Kernel_Fun, user_kernel.hpp:46,101,167) */
typedef double (*or_fn_t)(double r, const double *params);

typedef struct {
    int32_t family;
    int32_t ndim;
    double scale;
    double diag;
    const double *weight; /* nullable, indexed by point id */
    or_fn_t fn;           /* OR_K_CUSTOM: f(r) = fn(r, params) */
    const double *params;
} or_kernel_t;

/* Kernel_Fun of the device-compiled kernel classes the custom-family tests
 * use, written as those classes write them */
double or_f_yukawa(double r, const double *p);   /* p[0] exp(-r) / r */
double or_f_matern32(double r, const double *p); /* (1 + p[0] r) exp(-p[0] r) */
double or_f_gauss(double r, const double *p);    /* exp(-r r / p[0]) */

/* ---- generators and quadrature ---- */
void or_uniform_nodes(double a, double b, int n, double *out);
void or_tensor_grid(int ndim, int n, const double *nodes1d, double *pts);
void or_tensor_weights(int ndim, int n, const double *w1d, double *w);
void or_park_miller(int64_t n, double *out);
void or_splitmix_uniform(uint64_t seed, int64_t n, double *out);
int or_gauss_legendre(int n, double *nodes, double *weights);
double or_quadruple_integral(const double *a, const double *b, int ngauss);

/* ---- entries and dense products ---- */
double or_entry(const or_kernel_t *k, const double *pts, int64_t i, int64_t j);
void or_dense_matvec(const or_kernel_t *k, const double *pts, int64_t n,
                     const double *x, double *y, int nthreads);

/* ---- tree and lists ---- */
typedef struct or_tree or_tree_t;
or_tree_t *or_tree_build(int ndim, int64_t n, const double *pts,
                         const double *lo, const double *hi, int nmax, int dprime);
void or_tree_free(or_tree_t *t);
int or_tree_nlevels(const or_tree_t *t);
int64_t or_tree_level_size(const or_tree_t *t, int level);
int64_t or_tree_node_npts(const or_tree_t *t, int level, int64_t j);
void or_tree_node_points(const or_tree_t *t, int level, int64_t j, int64_t *out);
/* which = 0: neighbour list, 1: interaction list.  Returns the length and,
 * if out is non-NULL, writes the partner node indices in reference order. */
int64_t or_tree_list(const or_tree_t *t, int level, int64_t j, int which, int64_t *out);
int or_interaction_type(const or_tree_t *t, int level, int64_t a, int64_t b);

/* ---- one ACA block ---- */
typedef struct {
    int32_t rank;       /* columns of the stored factors */
    int32_t dense;      /* 1 if the block fell back to dense storage */
    int32_t steps;      /* computed_rank before the fallback decision */
    int32_t retries;    /* row + column retry iterations taken */
    int64_t entries;    /* kernel entries evaluated: rows evaluated x n + columns x m */
} or_aca_info_t;

uint64_t or_block_seed(int level, int64_t row_node, int64_t col_node);
/* Runs ACA on K(rows, cols).  U (m x cap) and V (n x cap), column-major, must
 * hold cap >= min(m, n) columns.  Returns the rank; U,V hold the factors for a
 * non-dense block.  row_piv/col_piv (nullable, cap entries) get the pivots. */
int or_aca(const or_kernel_t *k, const double *pts,
           const int64_t *rows, int m, const int64_t *cols, int n,
           double eps, uint64_t seed, int cap, double *U, double *V,
           int32_t *row_piv, int32_t *col_piv, or_aca_info_t *info);

/* ---- operator ---- */
typedef struct or_op or_op_t;
typedef struct {
    double n_flop_ref;     /* HDD_node.hpp:183-202 */
    int64_t max_rank;      /* HDD_node.hpp:205-220 */
    double e_near;         /* dense near-field entries */
    double e_far;          /* stored far-field entries (r(m+n) or m n) */
    int64_t n_blocks;      /* admissible blocks with both sides non-empty */
    int64_t n_dense;       /* of which dense fallback */
    int64_t n_skeleton;    /* of which skeleton (sys_size) blocks */
    double init_seconds;
    double aca_entries;    /* kernel entries the blocks' ACA evaluated (or_aca_info_t.entries summed) */
} or_op_stats_t;

/* sys_size < 0: store every factor (the device path's form);
 * sys_size >= 0: blocks with |cols| > sys_size keep the skeleton form of
 * LowRank_matrix.hpp:379-724 and re-evaluate entries on every matvec. */
or_op_t *or_op_assemble(or_tree_t *t, const or_kernel_t *k, const double *pts,
                        double eps, int sys_size, int nthreads);
/* Assemble only the blocks whose row node is selected: sel[level] gives a
 * stride s and offset o; node j is built iff j % s == o (s = 0: none). */
or_op_t *or_op_assemble_sample(or_tree_t *t, const or_kernel_t *k, const double *pts,
                               double eps, int sys_size, int nthreads,
                               const int64_t *stride, const int64_t *offset);
void or_op_free(or_op_t *op);
void or_op_matvec(or_op_t *op, const double *x, double *y, int nthreads);
void or_op_get_stats(const or_op_t *op, or_op_stats_t *s);
/* ranks / dense flags in block order (level, node, IL index; both non-empty) */
int64_t or_op_block_ranks(const or_op_t *op, int32_t *rank, int32_t *dense);

/* ---- GMRES (GMRES.hpp:160-303) ---- */
typedef struct {
    int32_t iterations;  /* return value of iterSolver::GMRES */
    int32_t matvecs;
    double resid;        /* last residual estimate */
} or_gmres_info_t;

int or_gmres(or_op_t *op, const double *b, double *x, int restart, int max_iter,
             double tol, int nthreads, double *hist, int hist_cap, or_gmres_info_t *info);
/* the same solver over a dense row-major matrix (test harness) */
int or_gmres_dense(const double *A, int64_t n, const double *b, double *x,
                   int restart, int max_iter, double tol, or_gmres_info_t *info);

/* ---- column-pivoted Householder QR solve (square, LowRank_matrix.hpp:57,698) ---- */
int or_cpqr_solve(int n, const double *A, const double *b, double *x);

#ifdef __cplusplus
}
#endif
#endif
