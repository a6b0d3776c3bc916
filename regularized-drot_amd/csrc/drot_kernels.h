// Host-visible launch interface of the solver kernels (drot_kernels.hip).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "drot_device.h"

namespace drot {

// Arguments of the fused QR pass (drot_qr.hip): one DR iteration over X.
struct QrPassArgs {
    float *x;
    const float *c;
    const float *phi1, *phi2;  // explicit phi (SolverState::phi_mode == 1)
    const float *ahat, *bhat;  // phi numerators: phi1 = (ahat + gamma)/n_cols, ...
    double *prow;              // [n_ct][n_rows] row partial sums
    double *pcol;              // [n_rt][n_cols] column partial sums
    double *pobj;              // [n_rt * n_ct] objective partials (even k)
    int n_rows, n_cols;
    int n_rt, n_ct;            // row tiles, column tiles
    int nstep;                 // column steps per wave; tile columns = nstep * 4 * (64 / lpc)
    int prio;                  // set by launch_qr_pass: 1 = issue priority by remaining steps
    float step, kappa;
    // zero-segment flags ([n_rt][n_cols] bytes; null: off): zin from the
    // previous pass, zout written by this one (see drot_qr_tile.h)
    const unsigned char *zin;
    unsigned char *zout;
    // screening of flagged segments (null: off; needs the flags and the
    // branch-free tile): headrooms [n_rt][n_cols], one G1 per tile
    // [n_rt * n_ct] (tile rt + ct * n_rt), row potentials of the previous
    // pass (read) and of this one (written)
    float *seg_e;
    double *g1;
    const float *f1_prev;
    float *f1_out;
    IterCtl ctl;
};

// Arguments of the fused GL pass (prox + group soft-threshold + sums).
struct GlPassArgs {
    float *x;
    const float *c;
    const float *phi1, *phi2, *ahat, *bhat;
    double *prow;       // [n_ct][n_rows]
    double *pcol;       // [n_cols]
    double *pobj;       // [n_ct]
    const int *bounds;  // [n_groups + 1] group row bounds
    const int *row_group; // [n_rows] group of each row
    int n_rows, n_cols, n_groups;
    int n_ct, tc;       // column tiles, columns per tile
    int lane_per_group; // 1: one lane sums a group serially (small groups)
    float step, lambda;
    IterCtl ctl;
    // long columns (GlTile::split): row chunks of kBlock*VEC*RCH rows; group g
    // owns the partial-norm rows gseg[g] .. gseg[g+1]-1 of pnorm
    // ([n_seg][n_cols]), one per chunk it touches from chunk gclo[g] on
    int n_rt;           // row chunks (split) / 1
    double *pnorm;
    const int *gseg;    // [n_groups + 1]
    const int *gclo;    // [n_groups]
};

// Arguments of the dual / aux update (shared by QR and GL).
struct FinalizeArgs {
    const double *prow, *pcol, *pobj;
    const float *p, *q;
    float *a, *b, *ahat, *bhat;
    double *pfin;       // [rb + cb][3]: residual part, sum-X part, objective part
    const SolverState *st;
    int n_rows, n_cols, n_rt, n_ct;
    int n_obj;          // number of objective partials
    int rb, cb;         // row blocks, column blocks (kFinRows each)
    int with_obj;       // this iteration produced objective partials
};

struct InitArgs {
    float *x;
    const float *p, *q;
    float *phi1, *phi2, *a, *b, *ahat, *bhat;
    SolverState *st;
    int n_rows, n_cols;
    int warm;
    float v_phi1, v_phi2, v_a, v_b, v_alpha;
    unsigned char *zf;  // QR zero-segment flags, cleared here (null: none)
    size_t zf_bytes;
    float *seg_e;       // QR screening headrooms, set to +inf here (null: none)
    size_t seg_e_count;
    double *g1;         // QR screening drift bounds, zeroed here
    size_t g1_count;
};

struct QrTile {
    int vec;     // rows of a column per lane: 1 or 4 (one access), 8 (two 4-row chunks)
    int lpc;     // lanes per column: 16, 32 or 64 (vec 1, 4); 8 or 16 (vec 8)
    int nt = 0;  // X store policy: 0 default (L2 write-back), 2 write-through (sc1)
    int u = 4;   // column steps per load batch (2 or 4; 1 for vec 8)
};

struct GlTile {
    int vec;        // 1 or 4
    int rch;        // row chunks per thread (per lane for the wave-per-column kernel)
    int wave;       // 1: one wave per column (gl_wave_kernel), n_rows <= 2048;
                    // 2: threads over rows, four columns per round (gl_rows_kernel)
    int split = 0;  // 1: columns split into row chunks over blocks (any n_rows):
                    // gl_split_norm_kernel + gl_split_apply_kernel per iteration
};

// Floats per column slice of gl_rows_kernel's LDS (prox values, then the
// group scales), rounded up to a 16-byte multiple for its vector stores.
__host__ __device__ inline int gl_rows_slice(int n_rows, int n_groups) {
    return (n_rows + n_groups + 3) & ~3;
}

hipError_t launch_init(const InitArgs &a, hipStream_t s);
// e0/e1 (optional): events recorded at the pass kernel's start and end.
hipError_t launch_qr_pass(const QrTile &t, bool even, int grid, const QrPassArgs &a,
                          hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
int qr_pass_grid(int n_rt, int n_ct);
hipError_t launch_gl_pass(const GlTile &t, int grid, size_t lds_bytes,
                          const GlPassArgs &a, hipStream_t s, hipEvent_t e0 = nullptr,
                          hipEvent_t e1 = nullptr);
hipError_t launch_finalize(int grid, const FinalizeArgs &a, hipStream_t s);
hipError_t launch_fix_x(float *x, const float *c, float step, size_t total,
                        const SolverState *st, hipStream_t s);
// Writes the stored value into the segments the last executed QR pass left
// flagged (zf: both flag buffers, zf_bytes each; tr rows per segment).
hipError_t launch_qr_unflag(float *x, const float *c, float step, const unsigned char *zf,
                            size_t zf_bytes, int n_rows, int n_cols, int tr,
                            const SolverState *st, hipStream_t s);
hipError_t launch_export_phi(const float *ahat, const float *bhat, const float *phi1,
                             const float *phi2, const SolverState *st, int n_rows,
                             int n_cols, float *out_phi1, float *out_phi2,
                             hipStream_t s);

bool qr_tile_supported(const QrTile &t);

// ------------------------------------------------------------------ batches

// One problem of a batch as the device sees it.  Matrices are column-major
// with leading dimension ld (ld % 4 == 0, mat_off % 4 == 0, padding rows are
// scratch); row vectors (p, phi1, a, ahat) live at row_off, column vectors at
// col_off of the packed arrays.
struct BatchProblem {
    long long mat_off;
    long long row_off, col_off;      // packed p / q (and dual vector) offsets
    long long prow_off, pcol_off;    // offsets into the partial buffers
    int n_rows, n_cols, ld;
    int n_rt, n_ct, nstep;
    int tile_off;                    // objective partials: pobj[tile_off + rt + ct * n_rt]
    int fin_off, n_fin, rb;          // finalize blocks: rb row blocks, then column blocks
    int max_iters;
    float step, kappa, eps;
    float v_phi1, v_phi2, v_a, v_b, v_alpha;  // warm-start constants
    int warm;
    long long zf_off;                // zero-segment flags: [n_rt][n_cols] bytes at zf + zf_off
};

struct BatchTile {
    int prob, rt, ct, pad;
};

struct BatchFin {
    int prob, blk;
};

struct QrBatchArgs {
    float *x;
    const float *c, *p, *q;
    float *phi1, *phi2, *a, *b, *ahat, *bhat;  // packed dual vectors
    double *prow, *pcol, *pobj, *pfin;
    SolverState *st;                           // one per problem
    const BatchProblem *probs;
    const BatchTile *tiles;
    int n_tiles;
    const BatchFin *fins;
    int n_fins;
    int n_probs;
    int k;                                     // iteration of this launch
    // zero-segment flags (null: off): two buffers of zf_len bytes; the pass of
    // iteration k reads buffer (k + 1) & 1 and writes buffer k & 1
    unsigned char *zf;
    long long zf_len;
    // screening (null: off): headrooms packed like the flags ([n_rt][n_cols]
    // floats per problem at zf_off), one G1 per tile (at tile_off), and two
    // buffers of row potentials (packed like p, row_len floats each; the pass
    // of iteration k reads buffer (k + 1) & 1 and writes buffer k & 1)
    float *seg_e;
    double *g1;
    float *f1pb;
    long long row_len;
};

// Batched tiles are 64 rows by 256 columns: 16 lanes per column holding 4
// rows each, 4 columns per wave step, 16 column steps.  (Two 4-row chunks per
// lane on 8 lanes, the same tile, made the mixed 256-problem forward 4%
// slower: 0.874 vs 0.837 s; its occupancy is 4 waves per SIMD against 5.)
constexpr int kBatchVec = 4;
constexpr int kBatchLpc = 16;
constexpr int kBatchTileRows = kBatchLpc * kBatchVec;
constexpr int kBatchTileCols = 256;          // 512 in batches of 50M elements and more
constexpr int kBatchTileColsScreened = 512;

int batch_pass_grid(int n_tiles);
hipError_t launch_batch_init(const QrBatchArgs &a, hipStream_t s);
hipError_t launch_batch_pass(bool even, const QrBatchArgs &a, hipStream_t s);
hipError_t launch_batch_finalize(const QrBatchArgs &a, hipStream_t s);
hipError_t launch_batch_fix_x(const QrBatchArgs &a, hipStream_t s);

// rows (or columns) reduced per finalize block
constexpr int kFinRows = 32;
bool gl_tile_supported(const GlTile &t);

// ------------------------------------------------------- implicit backward
//
// dL/dC of a converged QR-DROT plan by implicit differentiation of its KKT
// conditions on the support S (see drot_backward.hip).  One problem or a
// batch; every matrix is column-major n_rows x n_cols with leading dimension
// ld at float offset mat_off of the packed buffers.

struct BwdProblem {
    long long mat_off;
    long long iptr_off;   // colptr[n_cols + 1], rowptr[n_rows + 1], rowcur[n_rows]
    long long idx_off;    // rowidx[cap] (CSC order), colidx[cap] (CSR order)
    long long vec_off;    // 5 fp64 vectors of n_rows + n_cols: w, r, z, p, q
    long long cap;        // index capacity (< 2^31)
    int n_rows, n_cols, ld;
    int tile_begin, n_tiles;
    int max_cg;
    float thr;            // support: x > 0 and x > thr * |c|  (thr = tol * step)
    float pad;
    double reg;           // quadratic weight of the problem: (kappa - 1) / step
    double cg_tol;        // relative residual target of the CG solve
};

struct BwdTile {
    int prob, c0, c1, pad;
};

struct alignas(64) BwdState {
    long long nnz;        // support size (may exceed cap: overflow)
    int cg_iters;
    int converged;
    double res;           // final ||r|| / ||g||
    double gnorm;         // ||g||, g = (row sums, column sums) of G on S
};

struct BwdArgs {
    const float *c, *x, *g;
    const float *gobj;    // [n_probs] device scalars (grad of the objective) or null
    float *grad_c;
    const BwdProblem *probs;
    const BwdTile *tiles;
    BwdState *st;
    int *tile_cnt;
    int *iptr;
    int *idx;
    double *vec;
    int n_tiles, n_probs;
};

hipError_t launch_bwd_count(const BwdArgs &a, hipStream_t s);
hipError_t launch_bwd_fill(const BwdArgs &a, hipStream_t s);
hipError_t launch_bwd_cg(const BwdArgs &a, hipStream_t s);

// ----------------------------------------------------------- cost matrices
//
// C_ij = ||src_i - tgt_j||^2 (optionally / max C) for point clouds stored
// row-major [n_points, dim]; C in the solver layout (i contiguous).  Tiles of
// kCostTile source x kCostTile target points (drot_cost.hip).

constexpr int kCostTile = 64;
constexpr int kCostMaxDim = 64;

struct CostProblem {
    long long mat_off;           // float offset of C (leading dimension ld)
    long long src_off, tgt_off;  // first point of the problem in src / tgt
    long long xpart_off;         // fp64 partials: n_jt * n_rows * (dim + 1)
    long long ypart_off;         // fp64 partials: n_it * n_cols * (dim + 1)
    int n_rows, n_cols, ld;      // n_rows source points, n_cols target points
    int tile_begin, n_it, n_jt;
};

struct CostTile {
    int prob, it, jt, pad;
};

struct CostArgs {
    const float *src, *tgt;
    float *c;
    const float *grad_c;
    float *grad_src, *grad_tgt;
    const CostProblem *probs;
    const CostTile *tiles;
    float *tmax;                 // per tile: largest D
    int *targ;                   // per tile: its index j * n_rows + i (smallest on ties)
    float *scale;                // per problem: max D (1 when normalisation is off or max D = 0)
    int *argmax;                 // per problem: i*, j*
    double *xpart, *ypart;
    double *tdot;                // per tile: sum of grad_c * C
    int dim, normalize, n_tiles, n_probs, max_points;
};

hipError_t launch_cost_forward(const CostArgs &a, hipStream_t s);
hipError_t launch_cost_backward(const CostArgs &a, hipStream_t s);

}  // namespace drot
