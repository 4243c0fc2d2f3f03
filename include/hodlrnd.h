/*
 * hodlrnd.h -- C ABI of the MI355X (gfx950) HODLRnD hot path.
 *
 * The reference (kandapva/HODLRnD) exposes its operator as C++ templates
 * (include/HDD_matrix.hpp, include/GMRES.hpp); it has no FFI.  Each entry
 * point below replaces one of those calls, cited as file:line in the
 * reference.  The reference-shaped C++ API over this ABI is
 * include/hodlrnd/HODLRnD.hpp.
 *
 * Conventions: every function returns HODLR_OK (0) or a negative status;
 * the message of the last failure on a context is hodlr_last_error(ctx).
 * No exception crosses the ABI.  Vectors passed as host pointers are in the
 * caller's point order (ptsnD order); device pointers are in tree order
 * (see hodlr_tree_perm).  One operator may be used from one host thread at
 * a time.
 */
#ifndef HODLRND_H
#define HODLRND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HODLR_OK = 0,
    HODLR_E_ARG = -1,        /* invalid argument */
    HODLR_E_DEVICE = -2,     /* HIP runtime error */
    HODLR_E_NOMEM = -3,      /* device or host allocation failed */
    HODLR_E_RANK = -4,       /* ACA exceeded its rank budget */
    HODLR_E_NOCONV = -5,     /* solver stopped at max_iter */
    HODLR_E_BREAKDOWN = -6,  /* non-finite value in the solver */
    HODLR_E_STATE = -7       /* call not valid in the object's state */
};

/* kernel families: A_ij = diag if i == j, else scale * f(|x_i - x_j|) * w_j */
enum {
    HODLR_KERNEL_LOG = 0,  /* f(r) = log r     (user_kernel.hpp:46-49)   */
    HODLR_KERNEL_INV_R = 1,/* f(r) = 1 / r     (user_kernel.hpp:101-104) */
    HODLR_KERNEL_EXP = 2,  /* f(r) = exp(-r)   (user_kernel.hpp:167-170) */
    HODLR_KERNEL_CUSTOM = 3 /* f(r) = the caller's Kernel_Fun (user_kernel.hpp:46,101,167),
                               run on the device by caller-compiled kernels:
                               hodlr_op_assemble_custom only */
};

/* operator precisions built by hodlr_op_assemble */
enum {
    HODLR_PREC_F64 = 1,    /* FP64 factors (always built) */
    HODLR_PREC_F32 = 2     /* an FP32 copy of every factor as well */
};

/* solver modes */
enum {
    HODLR_SOLVE_F64 = 0,   /* FP64 GMRES on the FP64 operator (GMRES.hpp:160-303) */
    HODLR_SOLVE_MIXED = 1  /* FP64 iterative refinement around FP32-operator GMRES */
};

typedef struct hodlr_ctx hodlr_ctx;
typedef struct hodlr_tree hodlr_tree;
typedef struct hodlr_op hodlr_op;

/* Replaces the user kernel's getMatrixEntry (user_kernel.hpp:50-58,105-113,
 * 171-177) for device evaluation.  weight is nullable; when given it holds
 * one value per point in point order. */
typedef struct {
    int32_t family;
    int32_t ndim;
    double scale;
    double diag;
    const double *weight;
} hodlr_kernel_desc;

typedef struct {
    int32_t nlevels;       /* levels including the root (HDD_tree.hpp:73) */
    int32_t ndim;
    int64_t npoints;
    int64_t nnodes;        /* all levels */
    int64_t nleaves;
    int64_t max_leaf;      /* largest leaf */
    int64_t n_admissible;  /* interaction-list entries with both sides non-empty */
    int64_t n_near;        /* neighbour-list entries at the leaves, both non-empty */
} hodlr_tree_info;

typedef struct {
    double n_flop_ref;     /* HDD_node.hpp:183-202 */
    int64_t max_rank;      /* HDD_node.hpp:205-220 */
    double e_near;         /* near-field entries */
    double e_far;          /* stored far-field entries */
    int64_t n_blocks;      /* compressed blocks */
    int64_t n_dense;       /* blocks stored dense (LowRank_matrix.hpp:348-363) */
    double bytes_f64;      /* HBM bytes of the FP64 operator */
    double bytes_f32;      /* HBM bytes of the FP32 copy (0 if not built) */
    double matvec_bytes_f64; /* algorithmic bytes of one FP64 matvec, 1 RHS */
    double matvec_bytes_f32;
    double assemble_seconds;
    double rank_seconds;   /* ACA rank pass */
    double fill_seconds;   /* ACA fill pass + near field */
    int64_t n_skeleton;    /* skeleton blocks (hodlr_op_assemble_skeleton) */
    double e_skel;         /* entries of skeleton blocks evaluated per matvec: sum r (m + n) */
    double bytes_core;     /* HBM bytes of the skeleton cores and pivot positions */
    double aca_entries;    /* kernel entries the blocks' ACA evaluates (rows x n + columns x m per
                              block, pivot retries included; 0 for an operator read from a file) */
} hodlr_op_stats;

typedef struct {
    int32_t restart;       /* GMRES.hpp:44 (reference solve uses 100) */
    int32_t max_iter;      /* reference solve uses N */
    double tol;            /* relative residual (reference solve uses eps_ACA) */
    int32_t mode;          /* HODLR_SOLVE_F64 or HODLR_SOLVE_MIXED */
    int32_t x0_is_b;       /* 1: initial guess x0 = b (HDD_matrix.hpp:52-53);
                              2: x0 = 0 (the first residual is b itself);
                              0: x holds the initial guess */
    double inner_tol;      /* mixed mode: tightest relative tolerance an FP32 sweep may be
                              asked for (each sweep targets 0.3 tol ||b|| / ||r||);
                              0 = no floor */
    int32_t max_outer;     /* mixed mode: refinement sweeps */
    int32_t verify;        /* F64 mode: recompute ||b - A x|| / ||b|| at exit */
    int32_t hist_cap;      /* length of hist (may be 0) */
    double *hist;          /* residual history, one entry per iteration */
    int32_t ortho;         /* Arnoldi orthogonalisation of single-problem solves: HODLR_ORTHO_CGS2
                              (default: two classical Gram-Schmidt passes, two multi-dot kernels per
                              step) or HODLR_ORTHO_MGS (modified Gram-Schmidt, one basis vector at a
                              time in the reference's order, GMRES.hpp:236-240); batched solves
                              accept CGS2 only */
} hodlr_gmres_params;

#define HODLR_ORTHO_CGS2 0
#define HODLR_ORTHO_MGS 1

typedef struct {
    int32_t iterations;    /* return value of iterSolver::GMRES (GMRES.hpp:212,267,294,302) */
    int32_t matvecs_f64;
    int32_t matvecs_f32;
    int32_t outer;         /* refinement sweeps (mixed mode) */
    int32_t converged;
    int32_t hist_len;
    double resid;          /* relative residual estimate at exit */
    double true_resid;     /* ||b - A64 x|| / ||b|| recomputed in FP64 */
    double seconds;
} hodlr_gmres_info;

/* ---- context ---- */
/* device -1 gives a host-only context: tree construction and ordering helpers
 * work, everything that needs the GPU returns HODLR_E_STATE. */
int hodlr_ctx_create(int device, hodlr_ctx **out);
void hodlr_ctx_destroy(hodlr_ctx *ctx);
const char *hodlr_last_error(hodlr_ctx *ctx);
const char *hodlr_version(void);
/* No reference counterpart.  Assembly staging: the ACA rank pass keeps each
 * compressed block's factors in a staging buffer and the fill step copies
 * them into the operator instead of repeating the ACA; blocks that do not
 * fit are recomputed.  Both ways give bit-identical operators.  max_bytes
 * caps the buffer for the operators this context assembles afterwards:
 * -1 (default) = a quarter of free device memory (1/16 in a server's
 * builds), 0 = no staging, > 0 = at most that many bytes. */
int hodlr_ctx_set_assembly_staging(hodlr_ctx *ctx, int64_t max_bytes);

/* HDD_matrix.hpp:20-33 -> HDD_tree.hpp:23-74: 2^d tree and lists.
 * pts is n x ndim row-major (ptsnD order); lo/hi the bounding box. */
int hodlr_tree_build(hodlr_ctx *ctx, int ndim, int64_t n, const double *pts,
                     const double *lo, const double *hi, int nmax, int dprime,
                     hodlr_tree **out);
int hodlr_tree_get_info(hodlr_tree *t, hodlr_tree_info *info);
/* tree position -> point id (leaf order) */
int hodlr_tree_perm(hodlr_tree *t, int64_t *perm);
/* per node (level-major numbering): start in perm and point count */
int hodlr_tree_nodes(hodlr_tree *t, int64_t *level_offset /* nlevels+1 */,
                     int64_t *start, int64_t *count);
/* HDD_node.hpp:150-180: which = 0 neighbours, 1 interaction list; lists are
 * same-level node indices in reference order, CSR over the level-major nodes */
int hodlr_tree_lists(hodlr_tree *t, int which, int64_t *ptr /* nnodes+1 */,
                     int64_t *idx /* nullable: query the size via ptr */);
void hodlr_tree_destroy(hodlr_tree *t);

/* HDD_matrix.hpp:39-41 -> HDD_tree.hpp:76-83 -> HDD_node.hpp:118-147:
 * batched ACA of every admissible block (LowRank_matrix.hpp:64-377) and
 * the dense near field.  The tree must outlive the operator. */
int hodlr_op_assemble(hodlr_ctx *ctx, hodlr_tree *t, const hodlr_kernel_desc *k,
                      double eps_aca, int prec_flags, hodlr_op **out);
/* HDD_matrix.hpp:39-41 (Assemble_matrix_operators) for a user kernel with
 * any Kernel_Fun (user_kernel.hpp:46-58,101-113,167-177):
 *   A_ij = diag (i == j),  scale * f(|x_i - x_j|) [* weight_j] (i != j),
 * f evaluated on the device by kernels the caller compiled with hipcc for its
 * own f.  `entries` is that launcher table: hodlr::entry_launchers<Fn>() of
 * include/hodlrnd/device/entry_kernels.hpp, which the C++ API instantiates
 * for a kernel class whose Kernel_Fun is device code; a table built against
 * another layout of the descriptors is refused.  kernel_obj_bytes bytes at
 * kernel_obj (at most 1 MiB; the caller's kernel object) are copied to device
 * memory and handed to f.  k->family must be HODLR_KERNEL_CUSTOM.  Every
 * factor is stored (no skeleton form) and the operator cannot be saved;
 * otherwise it is an operator like hodlr_op_assemble's. */
int hodlr_op_assemble_custom(hodlr_ctx *ctx, hodlr_tree *t, const hodlr_kernel_desc *k,
                             const void *kernel_obj, int64_t kernel_obj_bytes, const void *entries,
                             double eps_aca, int prec_flags, hodlr_op **out);
/* A batch of independent problems assembled together (no reference
 * counterpart; the reference assembles one operator at a time through
 * HDD_matrix.hpp:39-41).  The result is their block-diagonal operator: one
 * ACA launch per pass and one launch per matvec kernel serve every problem.
 * Device vectors concatenate the problems in order, each in its own tree
 * order; hodlr_op_problem_offsets gives the nprob + 1 boundaries (and
 * returns nprob).  Each problem's blocks and ranks are the same as when it
 * is assembled alone. */
int hodlr_op_assemble_batch(hodlr_ctx *ctx, int nprob, hodlr_tree *const *trees,
                            const hodlr_kernel_desc *k, double eps_aca, int prec_flags,
                            hodlr_op **out);
int hodlr_op_problem_offsets(hodlr_op *op, int64_t *offsets, int cap);
/* LowRank_matrix.hpp:22-36 (SYS_SIZE < |cols| selects ACA_MEM_EFF),
 * :55-58 and :685-698: hodlr_op_assemble_batch (nprob = 1 for one problem)
 * in the reference's memory-efficient form.  A compressed block whose
 * column node holds more than sys_size points keeps no factors: only the
 * tree positions of its r pivot rows Is and columns Js and G = C^+, the
 * pseudo-inverse of its core C = A(Is, Js) (column-pivoted QR, rank cut at
 * eps r |R_00|).  Every matvec evaluates A(Is, J) and A(I, Js) again:
 * y_I += A(I, Js) G A(Is, J) x_J, in FP64 for both precisions.  The
 * reference uses sys_size = 2 Nmax (myHeaders.hpp:35-48); sys_size < 0
 * stores every factor, as hodlr_op_assemble_batch does.  Ranks, dense
 * fallbacks and the near field are those of the stored form. */
int hodlr_op_assemble_skeleton(hodlr_ctx *ctx, int nprob, hodlr_tree *const *trees,
                               const hodlr_kernel_desc *k, double eps_aca, int sys_size,
                               int prec_flags, hodlr_op **out);
/* HDD_matrix.hpp:69-85 */
int hodlr_op_get_stats(hodlr_op *op, hodlr_op_stats *s);
/* ranks and dense flags in block order (level, node, interaction index) */
int hodlr_op_block_ranks(hodlr_op *op, int32_t *rank, int32_t *dense, int64_t cap);
void hodlr_op_destroy(hodlr_op *op);

/* HDD_matrix.hpp:43-47 -> HDD_tree.hpp:86-107: y = A x for nrhs columns
 * (column-major n x nrhs), host buffers in point order. prec = HODLR_PREC_F64
 * or HODLR_PREC_F32.  nrhs <= 32 (at most 16 in FP64 on skeleton
 * operators); FP64 with 17-32 columns uses the matrix-core kernel forms. */
int hodlr_matvec(hodlr_op *op, int prec, int nrhs, const double *x, double *y);
/* Same on device buffers in tree order, stream-ordered on `stream` (a
 * hipStream_t, NULL for the default stream).  F64: x,y are double n x nrhs
 * with the right-hand sides interleaved (element (i, c) at i*nrhs + c);
 * F32: float, same layout. */
int hodlr_matvec_device(hodlr_op *op, int prec, int nrhs, const void *x, void *y,
                        void *stream);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The product of hodlr_matvec_device -- same arguments, same
 * kernels, bit-identical results -- with its intermediates (the V^T x
 * values, partial sums and skeleton slot values) taken from one fixed
 * 64 MiB device array per GPU instead of buffers the operator allocates;
 * the call allocates nothing.  An operator and RHS count whose
 * intermediates need more than 64 MiB are refused with HODLR_E_NOMEM and
 * a message stating the size.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race (serialise them). */
int hodlr_matvec_device_noalloc(hodlr_op *op, int prec, int nrhs, const void *x, void *y,
                                void *stream);

/* Rows of the dense operator applied to x, evaluated on the device in FP64
 * by direct summation (the dense self-check of main.cpp:42-57, on chosen
 * rows): y[r] = sum_j A(rows[r], j) x[j].  rows are point ids (problems
 * concatenated), x and y host buffers in point order / row order. */
int hodlr_dense_rows(hodlr_op *op, int64_t nrows, const int64_t *rows, const double *x, double *y);

/* HDD_matrix.hpp:51-62 with GMRES.hpp:160-303: solve A x = b, host buffers
 * in point order. x holds the initial guess unless params->x0_is_b. */
int hodlr_gmres(hodlr_op *op, const double *b, double *x,
                const hodlr_gmres_params *params, hodlr_gmres_info *info);
/* Same with b and x on the device (tree order, double). */
int hodlr_gmres_device(hodlr_op *op, const double *b, double *x,
                       const hodlr_gmres_params *params, hodlr_gmres_info *info);

/* Every (problem, right-hand side) of a batch operator solved as its own
 * GMRES instance with the semantics of GMRES.hpp:160-303 (no reference
 * counterpart for the batching itself).  The instances advance in lockstep
 * so each Arnoldi step is one nrhs-column matvec over the whole batch; an
 * instance's arithmetic is the same alone or in any batch.  Device form:
 * b, x are N x nrhs interleaved (element (i, c) at i * nrhs + c), tree order,
 * nrhs in {1, 2, 4, 8, 16}.  Host form: problems concatenated, each in point
 * order, column-major N x nrhs (nrhs <= 16).  infos holds nprob * nrhs
 * entries, instance p * nrhs + c. */
int hodlr_gmres_batch_device(hodlr_op *op, int nrhs, const double *b, double *x,
                             const hodlr_gmres_params *params, hodlr_gmres_info *infos);
int hodlr_gmres_batch(hodlr_op *op, int nrhs, const double *b, double *x,
                      const hodlr_gmres_params *params, hodlr_gmres_info *infos);
/* No reference counterpart: the batched solver captures the device work of
 * each Arnoldi step (matvec, CGS2 passes, norm, copies of the Hessenberg
 * column) into a hipGraph the first time the step runs and replays it as
 * one launch afterwards (default on; off while live kernel timing is on).
 * enable = 0 launches every kernel separately; results are identical. */
int hodlr_op_set_solver_graphs(hodlr_op *op, int enable);

/* ---- operator files and vector files ----
 * save writes the operator's problems (points, box, tree knobs, kernel,
 * weights), its ACA ranks and its FP64 arenas; load rebuilds the trees
 * and lists from the points and takes ranks and arenas from the file (no
 * ACA), building the FP32 copy when prec_flags asks for it.  A loaded
 * operator owns its trees (hodlr_op_get_tree; destroyed with the operator).
 * No reference counterpart: the reference keeps operators in memory. */
int hodlr_op_save(hodlr_op *op, const char *path);
int hodlr_op_load(hodlr_ctx *ctx, const char *path, int prec_flags, hodlr_op **out);
int hodlr_op_get_tree(hodlr_op *op, int problem, hodlr_tree **tree);
/* cereal portable-binary std::vector<double>, the reference's vector files
 * (myHeaders.hpp:146-170, the data/ fixtures): byte 0x01, u64 count, doubles.
 * load with v = NULL only reports the count. */
int hodlr_vec_save(const char *path, int64_t n, const double *v);
int hodlr_vec_load(const char *path, int64_t cap, double *v, int64_t *n);

/* ---- kernel selection (no reference counterpart) ----
 * Each matvec pass has several kernel forms.  variant bits: 1 = matrix-core
 * (MFMA) projection V^T x, 2 = MFMA panel product (both only with 4 or more
 * RHS), 4 = vector-ALU panel product walking each leaf's segments as one
 * flattened column list (else one segment at a time; suits many short
 * segments), 8 = vector-ALU panel product reading 16 bytes per lane and
 * column (two FP64 / four FP32 rows; up to 8 RHS; MFMA takes precedence
 * when also set; 8 | 4 = the same 16-byte loads over the flattened walk of
 * 4, for leaves with many short segments), 16 = vector-ALU projection packing several
 * rows of a narrow V-panel tile into one wave load (1-2 RHS), 32 = with 2,
 * the matrix-core panel product staging each group of columns through LDS
 * with 16-byte loads (4-32 RHS), 64 = vector-ALU projection taking items of
 * at most 64 rows one wave each, four per workgroup (1-32 RHS, FP64 to 16;
 * MFMA takes precedence).  tune times the forms of each pass on this operator for
 * nrhs = 1 .. 32 and the precisions in prec_mask and keeps the fastest;
 * results agree up to rounding order.  The 1-4 RHS default is 8, or 8 | 4
 * where the leaves' segment lists are short enough that one segment at a
 * time would leave most of each load round idle; each operator decides from
 * its own lists, so two operators (a batch and its problems alone) give
 * bit-identical rows only under the same forms.  get/set read or force the
 * choice for one (prec, nrhs). */
#define HODLR_MV_PROJ_MFMA 1
#define HODLR_MV_PANEL_MFMA 2
#define HODLR_MV_PANEL_FLAT 4
#define HODLR_MV_PANEL_VEC 8
#define HODLR_MV_PROJ_PACK 16
#define HODLR_MV_PANEL_WIDE 32
#define HODLR_MV_PROJ_WAVE 64
int hodlr_op_tune_matvec(hodlr_op *op, int prec_mask, int reps);
int hodlr_op_set_matvec_variant(hodlr_op *op, int prec, int nrhs, int variant);
int hodlr_op_get_matvec_variant(hodlr_op *op, int prec, int nrhs, int *variant);

/* ---- skeleton schedule (no reference counterpart) ----
 * Where a matvec of an operator with skeleton blocks (sys_size >= 0)
 * evaluates them: 1 = on the calling stream after the stored blocks'
 * passes; 2 = on a second stream of the operator, beside the stored passes
 * (compute-bound work beside memory-bound work); 3 = as 2 with each
 * skeleton workgroup holding 24 KiB of LDS, so at most 6 of them share a CU
 * with the stored passes' workgroups; 0 (default) = the tuner's pick for
 * that (prec, nrhs), else a cost model of stored bytes against skeleton
 * entries.  Every schedule gives bit-identical results.  Products under HIP
 * graph capture and hodlr_matvec_device_noalloc use 1.  get returns the
 * schedule a (prec, nrhs) product takes; hodlr_op_tune_matvec times 1-3. */
#define HODLR_SKEL_AUTO 0
#define HODLR_SKEL_SERIAL 1
#define HODLR_SKEL_OVERLAP 2
#define HODLR_SKEL_CAPPED 3
int hodlr_op_set_skeleton_schedule(hodlr_op *op, int mode);
int hodlr_op_get_skeleton_schedule(hodlr_op *op, int prec, int nrhs, int *mode);

/* ---- measurement ---- */
/* Times `reps` back-to-back matvecs on the op's stream with HIP events per
 * kernel; ms[0] = whole matvec, ms[1] = projection kernel (V^T x),
 * ms[2] = panel kernel (near field + U t), ms[3] = skeleton blocks' on-the-fly
 * product from its first to its last kernel, on whichever stream it ran (0
 * without skeleton blocks; beside the stored passes ms[0] < the sum),
 * averages per launch. */
int hodlr_matvec_timing(hodlr_op *op, int prec, int nrhs, int reps, double *ms);
/* Live timing: while enabled, every matvec records HIP events around its
 * projection, panel and skeleton launches on the streams they run on.  read
 * returns the summed milliseconds ms[6] = {f64 proj, f64 panel, f32 proj,
 * f32 panel, f64 skeleton blocks, f32 skeleton blocks} and counts[2] =
 * matvecs timed {f64, f32} since the last reset. */
int hodlr_op_profile(hodlr_op *op, int enable);
int hodlr_op_profile_read(hodlr_op *op, double *ms, int64_t *counts, int reset);
/* algorithmic HBM bytes of one launch of each matvec kernel (same order) */
int hodlr_matvec_kernel_bytes(hodlr_op *op, int prec, int nrhs, double *bytes);

/* Measurement helper with no counterpart in the reference: copies `bytes`
 * between two device buffers `reps` times with a plain streaming kernel and
 * returns the achieved (read + write) GB/s -- the measured ceiling that the
 * matvec's HBM figures are compared with, and the known-size kernel used to
 * scale the FETCH_SIZE / WRITE_SIZE hardware counters. */
int hodlr_stream_copy(hodlr_ctx *ctx, int64_t bytes, int reps, double *gbps);
/* Measurement helper with no counterpart in the reference: reads `bytes` of
 * one device buffer `reps` times with a plain streaming kernel (16-byte
 * loads, eight in flight per lane) and returns the achieved read GB/s -- the
 * measured ceiling of kernels that mostly read, such as the matvec's. */
int hodlr_stream_read(hodlr_ctx *ctx, int64_t bytes, int reps, double *gbps);

/* device memory helpers for callers that keep vectors resident */
int hodlr_device_alloc(hodlr_ctx *ctx, int64_t bytes, void **ptr);
int hodlr_device_free(hodlr_ctx *ctx, void *ptr);
int hodlr_copy_to_device(hodlr_ctx *ctx, void *dst, const void *src, int64_t bytes);
int hodlr_copy_to_host(hodlr_ctx *ctx, void *dst, const void *src, int64_t bytes);
/* point order <-> tree order for n x nrhs double vectors (host) */
int hodlr_to_tree_order(hodlr_tree *t, int nrhs, const double *src, double *dst);
int hodlr_from_tree_order(hodlr_tree *t, int nrhs, const double *src, double *dst);

/* ---- solve server (no reference counterpart) ----
 * A continuous-batching request loop on one GPU.  Operators are built by a
 * builder thread (tree + assembly) and admitted by their exact HBM footprint
 * (known after the ACA rank pass, before any arena is allocated); idle
 * resident operators are evicted least-recently-used when memory is needed
 * and rebuilt from their kept points when asked for again.  Solve requests
 * are split into columns; every column is an independent FP32-inner /
 * FP64-refined GMRES instance (restarted GMRES cycles of the reference's
 * semantics, GMRES.hpp:160-303, on the FP32 operator, each followed by an
 * FP64 residual) that takes a free slot of its operator's engine as soon as
 * one opens, at any iteration of the others, so one matvec per iteration
 * serves every column of an operator.  Engines of different operators run
 * on their own streams and overlap on the device.  A column's arithmetic
 * does not depend on the other columns: a result is the same alone and in
 * any mix. */
typedef struct hodlr_server hodlr_server;

typedef struct {
    double hbm_budget;     /* bytes for operators and solver space; <= 0: 90% of free memory */
    int32_t slots;         /* solve columns per operator engine: 1, 2, 4, 8, 16 or 32 */
    int32_t restart;       /* inner GMRES cycle length (<= 127) */
    int32_t max_outer;     /* FP64 refinement steps per column */
    double tol;            /* relative residual target, FP64-certified */
    double inner_floor;    /* tightest relative tolerance asked of one inner cycle */
    double eps_aca;
    int32_t nmax;
    int32_t dprime;
    int32_t builders;      /* concurrent operator builds (builder threads, each with its own stream) */
} hodlr_server_config;

typedef struct {
    int64_t op_id;
    int32_t nrhs;
    int32_t status;        /* 0 queued, 1 running, 2 done, < 0 a HODLR_E_* code */
    int32_t converged;     /* every column reached tol */
    int32_t iterations;    /* largest inner GMRES iteration count of the columns */
    int32_t outer;         /* largest FP64 refinement count of the columns */
    int32_t built;         /* the request waited for a build of its operator */
    double max_resid;      /* largest FP64 relative residual of the columns */
    double t_submit, t_start, t_done;  /* server clock, seconds */
} hodlr_request_info;

typedef struct {
    int64_t builds, evictions, requests_done, columns_done, ticks;
    int64_t resident_ops;
    double used_bytes, peak_bytes, budget_bytes;
    double build_seconds;  /* summed assembly wall time */
    double tick_seconds;   /* scheduler time spent in ticks (device work and collection) */
} hodlr_server_stats;

int hodlr_server_create(int device, const hodlr_server_config *cfg, hodlr_server **out);
/* stops the threads after the queued work is finished or failed */
void hodlr_server_destroy(hodlr_server *srv);
/* queue a new operator: points n x ndim (point order, copied), bounding box,
 * kernel (weights copied); the id is valid at once */
int hodlr_server_submit_build(hodlr_server *srv, int ndim, int64_t n, const double *pts, const double *lo,
                              const double *hi, const hodlr_kernel_desc *k, int64_t *op_id);
/* queue nrhs right-hand sides (n x nrhs column-major, point order, copied) */
int hodlr_server_submit_solve(hodlr_server *srv, int64_t op_id, int nrhs, const double *b, int64_t *req_id);
/* wait up to timeout_s (< 0: forever) for a request; x (n x nrhs, point
 * order) may be NULL.  Returns HODLR_E_STATE on time-out. */
int hodlr_server_wait(hodlr_server *srv, int64_t req_id, double timeout_s, double *x, hodlr_request_info *info);
int hodlr_server_get_stats(hodlr_server *srv, hodlr_server_stats *st);
const char *hodlr_server_last_error(hodlr_server *srv);

#ifdef __cplusplus
}
#endif
#endif
