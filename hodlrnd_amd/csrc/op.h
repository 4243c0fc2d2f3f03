// op.h -- the assembled HODLR operator (host view of device arenas).
#pragma once

#include <functional>
#include <memory>

#include "../../include/hodlrnd/device/aca_layout.h"
#include "../../include/hodlrnd/device/device_common.h"
#include "internal.h"

namespace hodlr {

// the entry kernels' launchers for the built-in kernel families
const EntryLaunchers &builtin_entry_launchers();
hipError_t launch_stage_copy(const StageCopy *items, int64_t n, const double *buf, const int64_t *leaf_start,
                             const int32_t *leaf_cnt, double *UP, double *VP, hipStream_t s);
hipError_t launch_stage_pivots(const PivotCopy *items, int64_t n, const double *buf, int32_t *SK, hipStream_t s);
hipError_t launch_convert(const double *a, float *b, int64_t n, hipStream_t s);
hipError_t launch_zero_self_diag(const NearLeaf *l, int64_t nl, float *NP, hipStream_t s);
hipError_t launch_f32_result(const float *y, const double *x, const double *diag, double *out, int64_t n, int k,
                             hipStream_t s);
hipError_t launch_add_diag_f32(float *y, const float *x, const double *diag, int64_t n, int k, hipStream_t s);
hipError_t launch_stream_copy(const void *a, void *b, int64_t bytes, hipStream_t s);
hipError_t launch_stream_read(const void *a, int64_t bytes, float *out, hipStream_t s);
template <typename T>
hipError_t launch_matvec(int nrhs, const ProjItem *pi, int64_t npi, const ProjItem *pw, int64_t npw_big,
                         const ProjReduce *pr, int64_t npr,
                         const PanelItem *pa, int64_t npa,
                         const Segment *segs, const QuadDesc *quads, const T *VP, const T *NP, const T *UP,
                         const T *x, T *t, T *part, const int32_t *tidx, T *y, int variant, hipEvent_t *ev,
                         hipStream_t s);

// skeleton blocks' on-the-fly product into ys (FP64, nrows x nrhs, zeroed
// here; t1, t2: FP64 slot values, part: partial projections; all
// nrhs-interleaved); launch_skeleton_add then adds ys into y
template <typename T>
hipError_t launch_skeleton(int nrhs, const SkeletonPlan &p, const T *x, double *t1, double *part, double *t2,
                           double *ys, int64_t nrows, hipStream_t s);
template <typename T>
hipError_t launch_skeleton_add(T *y, const double *ys, int64_t n, hipStream_t s);

hipError_t launch_bmdot(const double *const *V, int nv, const double *w, const Chunk *chunks, int nchunk,
                        double *partial, const int32_t *pchunk, int nprob, int K, double *out, hipStream_t s);
hipError_t launch_bgemv(const double *const *V, int nv, const double *coef, double sign, double *w,
                        const Chunk *chunks, int nchunk, int K, hipStream_t s);
hipError_t launch_bscale(const double *src, const double *coef, double *dst, const Chunk *chunks, int nchunk, int K,
                         hipStream_t s);

hipError_t launch_gather_cols(const double *const *src, double *out, int64_t n, int K, hipStream_t s);
hipError_t launch_scatter_cols(const double *w, const double *scale, double *const *dst, int64_t n, int K,
                               hipStream_t s);
hipError_t launch_mask_coef(double *cf, int nv, int K, const int32_t *lim, hipStream_t s);
hipError_t launch_put_col(double *dst, const double *src, int64_t n, int K, int c, hipStream_t s);
hipError_t launch_get_col(const double *src, double *dst, int64_t n, int K, int c, hipStream_t s);

int64_t mdot_chunks(int64_t n);
hipError_t launch_mdot(const double *V, int64_t ldv, int nv, const double *w, int64_t n, double *partial,
                       double *out, int accumulate, hipStream_t s);
hipError_t launch_gemv(const double *V, int64_t ldv, int nv, const double *h, double sign, double *w,
                       int64_t n, hipStream_t s);
hipError_t launch_scale(const double *src, double alpha, double *dst, int64_t n, hipStream_t s);
hipError_t launch_sub(const double *a, const double *b, double *c, int64_t n, hipStream_t s);
hipError_t launch_add(double *a, const double *b, int64_t n, hipStream_t s);
hipError_t launch_to_f32(const double *a, float *b, int64_t n, hipStream_t s);
hipError_t launch_to_f64(const float *a, double *b, int64_t n, hipStream_t s);
// Givens rotations and the residual test on the device (blas.hip)
hipError_t launch_gmres_init(const double *b2, double tol, int budget, GmresDevState *st, hipStream_t s);
hipError_t launch_gmres_cycle(const double *r2, double *sv, int ld, int first, GmresDevState *st, hipStream_t s);
hipError_t launch_gmres_step(double *H, double *cs, double *sn, double *sv, const double *hd, int i, int ld,
                             GmresDevState *st, hipStream_t s);
hipError_t launch_gmres_backsolve(const double *H, const double *sv, int ld, int k, double *y, hipStream_t s);
hipError_t launch_scale_by(const double *src, const double *alpha, double *dst, int64_t n, hipStream_t s);
hipError_t launch_bgmres_init(int NI, const double *b2, const double *tol, const int32_t *budget, GmresDevState *st,
                              hipStream_t s);
hipError_t launch_bgmres_cycle(int NI, const double *r2, int ld, int first, double *sv, GmresDevState *st,
                               double *scale, int32_t *count, hipStream_t s);
hipError_t launch_bgmres_step(int NI, int K, const double *h1, const double *h2, const double *n2, int i, int ld,
                              double *H, double *cs, double *sn, double *sv, GmresDevState *st, double *scale,
                              int32_t *count, hipStream_t s);
hipError_t launch_bgmres_backsolve(int NI, int K, int kmax, int ld, const double *H, const double *sv,
                                   const GmresDevState *st, double *y, double *coef, hipStream_t s);

// device buffers of one matvec's intermediates: the projection's V^T x
// values and partial-sum slots (operator precision), the skeleton phase's
// slot values and partial sums (FP64)
// events of one timed matvec (hodlr_op::matvec_dev)
constexpr int kMatvecEvents = 6;

struct MatvecWork {
    void *t = nullptr, *part = nullptr;
    double *skt1 = nullptr, *skt2 = nullptr, *skpart = nullptr, *skys = nullptr;
    // HODLR_SKEL_SERIAL: the skeleton phases on the calling stream; _OVERLAP
    // / _CAPPED: on the operator's second stream (which must exist)
    int skel = HODLR_SKEL_SERIAL;
};

// dynamic LDS of a skeleton workgroup under HODLR_SKEL_CAPPED: at most 6
// of them fit in a CU's 160 KiB, leaving 16 KiB for the stored blocks'
// workgroups (measured on the N = 10^6 grid operator: 20-24 KiB best for
// FP32, 24-26 KiB for FP64; 16 KiB leaves the projection short of slots,
// 32 KiB starves the skeleton phases)
constexpr int kSkelCapLds = 24576;

}  // namespace hodlr

// The operator of one problem, or the block-diagonal operator of a batch of
// independent problems (their vectors concatenated in problem order, each
// in its own tree order).
struct hodlr_op {
    hodlr_ctx *ctx = nullptr;
    hodlr_tree *tree = nullptr;           // first problem's tree
    std::vector<hodlr_tree *> trees;      // every problem's tree
    std::vector<hodlr_kernel_desc> kds;   // every problem's kernel (weights not kept)
    std::vector<int64_t> w_off;           // per problem offset of its weights in W, -1 = none
    std::vector<int64_t> pos_base;        // nprob + 1: first global position of each problem
    hodlr_kernel_desc kd{};
    double eps = 0.0;
    int prec_flags = 0;
    int64_t N = 0;

    // tree-order coordinates, weights, per-position leaf start / size
    hodlr::DevBuf X, W, leaf_start, leaf_cnt;
    hodlr::DevBuf ord;   // per problem and level: global positions in ascending point id per node
    hodlr::DevBuf kdev;  // per problem KernelDev
    // entry kernels: the built-in families', or a caller's (custom family,
    // whose kernel object is copied to kobj)
    const hodlr::EntryLaunchers *entries = nullptr;
    hodlr::DevBuf kobj;
    hodlr::DevBuf diag;  // per global position: A_ii (the FP32 copy stores a zero diagonal)
    bool uniform_diag = true;

    // blocks in reference order (level, node, interaction index)
    int64_t nblk = 0;
    std::vector<int32_t> b_level, b_m, b_n, b_prob, rank, flags;
    std::vector<int64_t> b_row, b_col;   // global node ids (problem base + level-major id)
    std::vector<int64_t> b_rowj, b_colj; // node index within its level (seeds)
    std::vector<int64_t> node_start;     // global node id -> first global position
    std::vector<int64_t> ord_off;        // (problem, level) -> offset of its order map

    // arenas: near field, U panels, V panels
    int64_t n_np = 0, n_up = 0, n_vp = 0, n_t = 0, n_slots = 0;
    hodlr::DevBuf NP64, UP64, VP64, NP32, UP32, VP32;

    // matvec work lists
    int64_t n_proj = 0, n_red = 0, n_panel = 0, n_seg = 0, n_quad = 0;
    hodlr::DevBuf proj, red, panel, segs, quads, tidx;
    // the projection items again for the one-wave-per-item form: the first
    // n_projw_big have more than kProjWaveRows rows (one workgroup each),
    // the rest are taken four per workgroup
    int64_t n_projw_big = 0;
    hodlr::DevBuf projw;
    double e_near = 0, e_far = 0, e_v = 0, e_u = 0, t_lowrank = 0;

    // per-precision work buffers (grown on demand)
    int cap_rhs64 = 0, cap_rhs32 = 0;
    // bumped whenever a work buffer is reallocated or a kernel form changes:
    // graphs captured from matvecs are valid only for the generation they
    // were captured at
    uint64_t work_gen = 1;
    // solvers replay captured Arnoldi steps (hipGraph) unless disabled or
    // while live kernel timing is on
    bool solver_graphs = true;
    hodlr::DevBuf t64, part64, x64, y64, t32, part32, x32, y32;

    // skeleton blocks (LowRank_matrix.hpp:22-36): compressed blocks whose
    // column node holds more than sys_size points keep only their pivot
    // positions (SK: n_sk row slots, then n_sk column slots) and the
    // pseudo-inverse of their core (GA, n_g FP64 entries); their entries are
    // evaluated in every matvec.  sys_size < 0: none.
    int sys_size = -1;
    int64_t n_skel = 0, n_sk = 0, n_g = 0, n_skslots = 0;
    hodlr::DevBuf SK, GA, skproj, skred, skcore, skex;
    int64_t n_skproj = 0, n_skred = 0, n_skcore = 0;
    std::vector<hodlr::SkelLaunch> skproj_l, skex_l;
    double e_skel = 0;
    int cap_rhs_sk = 0;
    hodlr::DevBuf skt1, skt2, skpart, skys;
    hodlr::SkeletonPlan skel_plan(int lds_min = 0) const;
    // where a matvec's skeleton phases run (HODLR_SKEL_*): on the calling
    // stream after the stored blocks' passes, or on `side` (created on first
    // use) beside them, with all the occupancy they can get or capped so the
    // stored passes keep most of every CU.  Every schedule gives the same
    // bits.  skel_mode is the caller's choice (auto: skel_pick, else the
    // cost model), skel_pick[prec f64/f32][log2 nrhs] the tuner's (0: none)
    int skel_mode = HODLR_SKEL_AUTO;
    int skel_pick[2][6] = {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}};
    int skel_schedule(int prec, int nrhs) const;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    hodlr_op_stats stats{};

    // solver state kept between solves (Krylov basis, mixed-mode vectors)
    std::shared_ptr<void> solver_ws;
    std::shared_ptr<void> batch_ws;

    // live kernel timing: kMatvecEvents events per matvec (before the
    // projection, before the panel product, after the panel product, after
    // the whole product, before and after the skeleton phases) recorded on
    // the streams they run on while profiling is on
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, size_t>> ev_pending;  // (precision, first event index)
    double prof_ms[2][3] = {{0, 0, 0}, {0, 0, 0}};   // [prec f64/f32][proj/panel/skeleton]
    int64_t prof_count[2] = {0, 0};
    void profile_collect();
    ~hodlr_op();

    // FP32: with add_diag false the result omits the diagonal (the FP32
    // copy stores the self blocks with a zero diagonal; callers add diag x).
    // ev (nullable): kMatvecEvents events, see profiling
    void matvec_dev(int prec, int nrhs, const void *x, void *y, hipStream_t s, hipEvent_t *ev = nullptr,
                    bool add_diag = true);
    void ensure_work(int prec, int nrhs);
    // the same product with its intermediates in caller-given device
    // buffers (sizes from matvec_work_sizes): allocates nothing
    void matvec_dev_on(int prec, int nrhs, const void *x, void *y, hipStream_t s, const hodlr::MatvecWork &w,
                       hipEvent_t *ev = nullptr, bool add_diag = true);
    // bytes of the intermediates of one (prec, nrhs) product, in the order
    // of MatvecWork: t, part, skt1, skt2, skpart, skys (zero where unused)
    void matvec_work_sizes(int prec, int nrhs, size_t sz[6]) const;

    // called by assemble once the exact arena footprint (bytes, all
    // requested precisions) is known and before it is allocated; may block
    // until memory is available or throw to abandon the build
    std::function<void(double)> admit;

    // operator files: the stored ranks / flags replace the ACA rank pass and
    // fill_from writes the FP64 arenas in place of the fill pass
    std::vector<int32_t> preset_rank, preset_flags;
    std::function<void(hodlr_op &)> fill_from;
    std::vector<std::unique_ptr<hodlr_tree>> owned_trees;  // trees of a loaded operator

    // kernel form per (precision, RHS count): kMatvec* bits, indexed
    // [prec f64/f32][log2 nrhs]; set by tune_matvec or the caller.  The
    // defaults: the vector-ALU 16-byte panel up to 4 RHS, the LDS-staged
    // matrix-core panel from 8 on; short projection items one wave each
    // for FP64 at 8-16 RHS and FP32 at 16-32; FP64 at 32 RHS has only the
    // matrix-core projection
    int mv_variant[2][6] = {{8, 8, 8, 98, 98, 35}, {8, 8, 8, 34, 98, 98}};
    int variant_for(int prec, int nrhs) const;
    // time every kernel form per phase on this operator and keep the fastest
    void tune_matvec(int prec_mask, int reps);
};

namespace hodlr {
// op.trees and op.kds must be set; weights in kds are read during assembly only
void assemble(hodlr_op &op);
// host threads used by assemblies run from the calling thread (0 = up to 16)
void set_assembly_threads(int n);
// keep freed stream-ordered memory of the device's default pool mapped
void keep_pool_memory(int device);
// device address of this device's scratch array of the allocation-free
// matvec, looked up once per device (noalloc.hip; the device must be current)
void *noalloc_scratch_setup(int device);
// operator files (opio.cpp)
void save_op(hodlr_op &op, const char *path);
hodlr_op *load_op(hodlr_ctx *ctx, const char *path, int prec_flags);
}
