// aca_layout.h -- work descriptors shared by host planning code and kernels.
#pragma once

#include <cstdint>

namespace hodlr {

constexpr int kMaxTries = 10;     // LowRank_matrix.hpp:117
constexpr int kProjTile = 256;    // columns of one V-panel tile (projection work item)
constexpr int kProjChunk = 2048;  // rows of one projection work item
// matvec kernel forms (bit set): matrix-core projection / panel product
// (only used with 4 or more right-hand sides); vector-ALU panel product
// walking the leaf's segments as one flattened column list (otherwise one
// segment at a time)
constexpr int kMatvecProjMfma = 1;
constexpr int kMatvecPanelMfma = 2;
constexpr int kMatvecPanelFlat = 4;
// vector-ALU panel product with one 16-byte load per lane and column (up to
// 8 right-hand sides); the near-field and U arenas carry kArenaPad elements
// of tail padding for its reads past a tile's last row
constexpr int kMatvecPanelVec = 8;
// row-packed vector-ALU projection (1-2 right-hand sides): narrow V-panel
// tiles put several rows in one wave load
constexpr int kMatvecProjPack = 16;
// matrix-core panel product loading its operands with 16-byte loads through
// a wave-private LDS stage (with kMatvecPanelMfma)
constexpr int kMatvecPanelWide = 32;
// projection of short items (at most kProjWaveRows rows) one wave per item,
// four per workgroup, with no LDS reduction (else every item is a workgroup)
constexpr int kMatvecProjWave = 64;
constexpr int kProjWaveRows = 64;
constexpr int64_t kArenaPad = 16;

// one admissible block for the ACA kernel
// Row i of the block is the i-th smallest point id of the row node (the
// order of index_of_points, HDD_clusters.hpp:112-124), found at global tree
// position ord[row0l + i], where ord is the block's (problem, level) order
// map indexed by problem-local position; columns likewise.
struct AcaBlock {
    int64_t row0, col0;   // first global tree position of the row / column node
    int64_t row0l, col0l; // the same, local to the block's problem
    int64_t scratch;      // offset (doubles) of this block's scratch in the pool
    const int32_t *ord;   // ascending-id order map of the block's problem and level
    int32_t m, n;
    int32_t cap;          // rank capacity of the scratch
    int32_t id;           // block id
    uint64_t seed;        // retry stream (replaces rand(), LowRank_matrix.hpp:175,274)
    int32_t prob;         // problem of the block (kernel descriptor index)
    int32_t stage;        // rank pass: keep 1 = the factors, 2 = the pivot positions (skeleton
                          // candidate) in the staging buffer, 0 = nothing
};

// Factor staging of the rank pass.  A compressed block whose factors fit
// takes r(m + n) doubles of buf at an offset handed out by an atomic
// cursor: U (m x r, column-major, rows in ACA order) then V (n x r); a
// skeleton candidate takes r doubles holding its 2r pivot positions (int32:
// rows then columns, in step order).  off[id] records the offset, -1 when
// the block was not staged (dense, out of room).  The fill step then copies
// staged factors into their panels and staged pivots into the operator's
// pivot slots instead of repeating the ACA.
struct AcaStage {
    double *buf;                  // nullptr: no staging
    unsigned long long *cursor;   // doubles handed out so far (may exceed cap)
    int64_t cap;                  // doubles in buf
    int64_t *off;                 // per block id
    int64_t *evals;               // per block id: entries its last ACA run evaluated (nullptr: not counted)
    int32_t pts_lds;              // workgroup ACA: dynamic LDS bytes for the block's points (0: none)
};

// one piece of a staged block's factors to copy into the panels: elements
// [e0, e1) of its r(m + n) staged values (U first, then V)
struct StageCopy {
    int64_t src;           // staging offset of the block's U
    int64_t row0, col0, row0l, col0l;
    const int32_t *ord;    // ascending-id order map (as AcaBlock)
    int64_t up_base, vp_base;
    int32_t m, n, r, R, coff, T, voff;
    int32_t e0, e1, pad;
};

// a skeleton block's staged pivots: 2r int32 at buf + src into its row and
// column pivot slots
struct PivotCopy {
    int64_t src, sk_row, sk_col;
    int32_t r, pad;
};

// where a block's factors go in the operator (fill pass)
struct FillDest {
    int64_t up_base;      // U panel of the row node (elements)
    int64_t vp_base;      // V panel of the column node (elements, tile-major)
    int32_t R, coff;      // U panel width and this block's first column
    int32_t T, voff;      // V panel width (padded to 4) and this block's first column
    // skeleton blocks (>= 0): the block keeps no factors; the fill pass writes
    // the tree positions of its r pivot rows to SK[sk_row ..) and of its r
    // pivot columns to SK[sk_col ..) (LowRank_matrix.hpp:685-689)
    int64_t sk_row, sk_col;
};

// a block stored dense (LowRank_matrix.hpp:348-363)
struct DenseBlock {
    int64_t row0, col0, up_base;
    int32_t m, n, R, coff;
    int32_t prob, pad;
};

// one leaf's near-field panel (HDD_node.hpp:136-144)
struct NearLeaf {
    int64_t np_base;      // panel base (elements), nc x ncols column-major
    int64_t row0;         // first global tree position of the leaf
    int64_t colsrc;       // offset into the column-source table
    int32_t nc, ncols;
    int32_t prob, pad;
};

__host__ __device__ inline int64_t aca_icap(int cap) { return (int64_t)cap + 8; }

// doubles of scratch one block needs: U, V, row, col, used flags, pivot lists
__host__ __device__ inline int64_t aca_scratch_doubles(int m, int n, int cap) {
    int64_t mn = (int64_t)m + n;
    return mn * cap + mn + (mn + 7) / 8 + (3 * aca_icap(cap) + 1) / 2 + 2;
}

// ---- matvec work ----

// V-panel projection tile: t[cols] (+)= VP[rows j0..j1, c0..c0+w)^T x[b0+j0..]
struct ProjItem {
    int64_t vp;           // element offset of (row j0, column c0) in the V arena
    int64_t x0;           // tree position matching row j0
    int64_t out;          // slot: >= 0 partial-sum slot, < 0 direct (-(tile index) - 1)
    int32_t ld;           // panel width T_b
    int32_t rows;
    int32_t w;            // tile width (<= kProjTile), padding columns excluded
    int32_t tcol;         // first global V column of the tile (into tidx)
};

// reduction of partial projection sums of one tile
struct ProjReduce {
    int64_t slot0;
    int32_t nchunks, w, tcol, pad;
};

// one row tile of one leaf for the panel kernel
struct PanelItem {
    int64_t y0;           // first tree position of the tile's rows
    int64_t seg0;         // first segment
    int32_t nseg;
    int32_t h;            // rows in this tile (<= 64)
    int32_t ld;           // leaf size (column stride of every segment)
    int32_t r0;           // row offset of the tile inside the leaf
    int32_t nq4;          // sum over segments of ceil(len / 4)
    int32_t nqg;          // sum over segments of ceil(len / G), G = 64 / h
    int64_t q0;           // first quad descriptor of the leaf
};

// one group of up to 4 consecutive columns of a leaf's segment list: column
// c < ncol is arena element mat + c*ld + row (arena NP if kind == 0, else
// UP) and multiplies vector element vec + c (x if kind != 1, else t)
struct QuadDesc {
    int64_t mat;
    int32_t vec;
    int32_t info;         // kind | ncol << 8
};

// one contiguous ld x len column-major matrix times a vector slice:
//   arena 0 (near-field block):           column q multiplies x[vec + q];
//   arena 1 (U panel, low-rank blocks):   column q multiplies t[vec + q];
//   arena 2 (U panel, dense-fallback K):  column q multiplies x[vec + q]
struct Segment {
    int64_t mat;          // element offset in the arena (row 0 of the leaf)
    int64_t vec;
    int32_t len;
    int32_t arena;
};

// ---- skeleton blocks (evaluated on the fly) ----
// Slots: every skeleton block b of rank r owns r row slots (its pivot rows,
// grouped by column node) and r column slots (its pivot columns, grouped by
// row node).  SK holds the tree position of every row slot, then of every
// column slot; t1 / t2 hold one value per row / column slot and RHS.
constexpr int kSkelChunk = 1024;  // sources of one projection work item
constexpr int kSkelTile = 128;    // targets of one work item (two per lane)

// t1[t0 + l] (+)= sum over sources j in [src0, src0 + nsrc) of A(SK[t0 + l], j) x_j
struct SkelProj {
    int64_t t0;           // first row slot (<= kSkelTile targets)
    int64_t src0;         // first source tree position (the column node's points)
    int64_t out;          // >= 0: first partial slot (kSkelTile per item), < 0: direct into t1
    int32_t ntgt, nsrc;
    int32_t prob, pad;
};

// sum of the partial projections of one target tile
struct SkelReduce {
    int64_t t0, slot0;
    int32_t ntgt, nch;
};

// t2[c0 .. c0 + r) = G t1[t0 .. t0 + r), G the r x r core pseudo-inverse
// (column-major, at element g of the core arena)
struct SkelCore {
    int64_t g, t0, c0;
    int32_t r, pad;
};

// y[y0 + l] += sum over column slots q in [c0, c0 + nc) of A(y0 + l, SKC[q]) t2[q]
struct SkelExpand {
    int64_t y0, c0;
    int32_t h, nc;        // rows (<= kSkelTile) and column slots
    int32_t prob, pad;
};

struct KernelDev;

// one launch of a specialised proj / expand kernel: items [begin, end) share
// the kernel family and NDIM (key = 16 family + ndim); expand launches are
// ordered by tree level, so each row has one writer per launch
struct SkelLaunch {
    int64_t begin, end;
    int32_t key, level;
};

// the skeleton product's work lists (device pointers into the operator's
// buffers; the launch lists are host arrays)
struct SkeletonPlan {
    const SkelProj *proj;
    const SkelReduce *red;
    const SkelCore *core;
    const SkelExpand *ex;
    int64_t nred, ncore;
    const SkelLaunch *proj_l, *ex_l;
    int nproj_l, nex_l;
    const KernelDev *K;
    const int32_t *SK;    // row slots, then column slots at SK + nrow_slots
    int64_t nrow_slots;
    const double *G;
    int lds_min;          // dynamic LDS of a proj / expand workgroup at least this many bytes
};

// rows of a batched vector that belong to one problem (never crosses a
// problem boundary)
struct Chunk {
    int64_t row0;
    int32_t rows, prob;
};

// state of one GMRES solve (or one batch instance) kept on the device
// between Arnoldi steps; the host reads it to learn when to stop
struct GmresDevState {
    double normb;    // ||b||, 1 when b = 0 (GMRES.hpp:192-193)
    double tol;
    double resid;    // relative residual estimate after the last step
    double scale;    // scale of the next basis vector (0: none)
    int32_t active;  // 1 while the solve goes on
    int32_t conv;    // step of this cycle at which it ended; -1 running or out of budget at a
                     // cycle start; -2 non-finite Arnoldi norm; -3 residual test passed at a cycle start
    int32_t steps;   // Arnoldi steps taken (GMRES.hpp's j - 1)
    int32_t budget;  // max_iter
    int32_t iters;   // the iteration count reported once the solve has ended (GMRES.hpp:212,267,294,302)
    int32_t converged;
};

}  // namespace hodlr
