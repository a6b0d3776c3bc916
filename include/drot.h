/*
 * drot.h -- C ABI of the MI355X (gfx950) Douglas-Rachford regularized optimal
 * transport solver, libdrot.so.
 *
 * Matrix convention (same as the reference, core/kernel_qr.hpp:3): C and X are
 * column-major n_rows x n_cols, element (i, j) at [j * n_rows + i], with
 * n_rows = len(p) and n_cols = len(q).  A row-major PyTorch tensor of shape
 * [n_cols, n_rows] (= [M, N] in pytorch/src/qr_drot_cuda_kernel.cu:89-93) has
 * exactly this memory.
 *
 * Two families of entry points:
 *   *_solve_f32    device pointers + caller-owned workspace + stream; what a
 *                  framework binding calls (replaces the reference's L1 loops
 *                  quadratic_regularizer_drot / group_lasso_regularizer_drot).
 *   *_wrapper_f32  host pointers, allocation, copies and timers inside; the
 *                  reference's core wrappers with the same argument order.
 * All functions are thread-safe for distinct workspaces.  Solves and backwards
 * block until they have finished on the given stream (as the reference's loops
 * do); the cost-matrix functions are asynchronous on their stream.
 */
#ifndef DROT_H_
#define DROT_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DROT_ABI_VERSION 2

typedef enum drot_status {
    DROT_OK = 0,
    DROT_ERR_INVALID_ARGUMENT = 1,   /* bad size, null pointer, bad parameter */
    DROT_ERR_WORKSPACE_TOO_SMALL = 2,
    DROT_ERR_UNSUPPORTED = 3,        /* shape outside what the kernels handle */
    DROT_ERR_HIP = 4,                /* a HIP runtime call failed */
    DROT_ERR_ALLOC = 5               /* host or device allocation failed */
} drot_status;

/* QR-DROT parameters.  Mirrors the scalar arguments of
 * quadratic_regularizer_drot<T> (core/drot_qr.hpp:159-185). */
typedef struct drot_qr_params {
    float step_size;  /* s = rho / (n_rows + n_cols)                           */
    float scale;      /* kappa = 1 + s (n_rows + n_cols) r_weight              */
    float r_weight;   /* kept for reporting; the iteration uses `scale`       */
    int max_iters;
    float eps;        /* stop when the primal residual < eps (even iterations) */
    int warm_start;   /* 1: closed-form warm start (always used by the PyTorch
                         binding); 0: X = p q^T (core/drot_qr.hpp:308-320)     */
} drot_qr_params;

/* GL-DROT parameters.  Mirrors group_lasso_regularizer_drot<T>
 * (core/drot_glr.hpp:242-273). */
typedef struct drot_glr_params {
    float step_size;  /* s = rho / (n_rows + n_cols)                           */
    float lambda;     /* group soft-threshold (already includes the sqrt factor) */
    int n_groups;     /* groups are row ranges of each column                  */
    int max_iters;
    float eps;        /* stop when the primal residual <= eps (every iteration) */
    int warm_start;
} drot_glr_params;

typedef struct drot_result {
    int n_iter;            /* iteration count in the reference's convention:
                              QR: k+1 at convergence, max_iters+1 otherwise
                              (core/drot_qr.hpp:233; INT_MAX when max_iters is
                              INT_MAX); GL: executed iterations */
    int converged;         /* 1 if the stopping test passed                    */
    int iterations_run;    /* iterations actually executed                     */
    float objective;       /* <C, X'> of the last iteration that computed it   */
    double objective_f64;  /* same, fp64 accumulation                          */
    double residual;       /* sqrt(|rowsum - p|^2 + |colsum - q|^2), fp64, of
                              the last iteration that ran the stopping test    */
    float run_ms;          /* wrappers only: solve time (as dur_in_ms)         */
    float prep_ms;         /* wrappers only: alloc + copies + init             */
    double pass_ms[2];     /* timing mode only (drot_set_timing): summed
                              execution time of sampled fused-pass launches that
                              did work ([0] = QR even iterations / GL, [1] = QR
                              odd): per launch, earliest block start to latest
                              block end on the device's constant-rate clock;
                              launches k with k % 8 < 2 are sampled             */
    int pass_count[2];     /* number of launches summed in pass_ms             */
    double pass_ms_events[2]; /* every working pass launch timed by HIP events
                              bracketing it on the solve's stream (these
                              include the dispatch of the launch)              */
    int pass_count_events[2];
} drot_result;

/* ---------------------------------------------------------------- params */

/* PyTorch binding convention (pytorch/src/qr_drot_cuda_kernel.cu:38-40):
 * s = rho / (float(n_rows) + float(n_cols)), kappa = 1 + s*r_weight*(n_rows+n_cols),
 * warm start. */
drot_status drot_qr_params_from_rho(float rho, float r_weight, int n_rows, int n_cols,
                                    int max_iters, float eps, drot_qr_params *out);

/* Core wrapper convention (core/drot_qr.hpp:262): s given,
 * kappa = 1 + s*(n_rows+n_cols)*r_weight. */
drot_status drot_qr_params_from_step(float step_size, float r_weight, int n_rows,
                                     int n_cols, int max_iters, float eps,
                                     int warm_start, drot_qr_params *out);

/* PyTorch binding convention (pytorch/src/glr_drot_cuda_kernel.cu:45-47):
 * s = rho/(n_rows+n_cols), lambda = r_weight * sqrt(n_rows / n_groups). */
drot_status drot_glr_params_from_rho(float rho, float r_weight, int n_rows, int n_cols,
                                     int n_groups, int max_iters, float eps,
                                     drot_glr_params *out);

/* Core wrapper convention (core/drot_glr.hpp:385):
 * lambda = r_weight / sqrt(n_rows / n_groups). */
drot_status drot_glr_params_from_step(float step_size, float r_weight, int n_rows,
                                      int n_cols, int n_groups, int max_iters,
                                      float eps, int warm_start, drot_glr_params *out);

/* ---------------------------------------------------------------- device API */

/* Bytes of device workspace a solve of this shape needs (256-byte aligned
 * start recommended).  Replaces the 11 (QR) / 17 (GL) workspace tensors the
 * reference binding allocates (pytorch/src/qr_drot_cuda_kernel.cu:107-119,
 * pytorch/src/glr_drot_cuda_kernel.cu:118-135). */
drot_status drot_qr_workspace_size(int n_rows, int n_cols, size_t *bytes);
drot_status drot_glr_workspace_size(int n_rows, int n_cols, int n_groups, size_t *bytes);

/* QR-DROT solve on device memory.  Replaces quadratic_regularizer_drot<float>
 * (core/drot_qr.hpp:159-234) as called by quadratic_regularizer_drot_torch_float32
 * (pytorch/src/qr_drot_cuda_kernel.cu:12-78), including its initialisation.
 * c, p, q, x: device pointers; x receives the plan.  stream: a hipStream_t
 * (NULL = default stream).  result may be NULL. */
drot_status drot_qr_solve_f32(const float *c, const float *p, const float *q, int n_rows,
                              int n_cols, const drot_qr_params *params, float *x,
                              void *workspace, size_t workspace_bytes, void *stream,
                              drot_result *result);

/* Allocation-free variant of drot_qr_solve_f32.  It has no counterpart in the
 * reference; it is for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It calls no
 * hipMalloc, hipFree, hipHostMalloc or other allocating routine: the solver's
 * workspace is a static device array of DROT_NOALLOC_SCRATCH_BYTES that every
 * device holds from the library's first use there until the process ends, and
 * the host polls the done flag through pageable memory (timing mode is
 * ignored).  Shapes whose workspace (drot_qr_workspace_size) exceeds that size
 * return DROT_ERR_WORKSPACE_TOO_SMALL.  All calls on one device share that
 * device's array and it keeps the last call's contents (each solve
 * initialises what it reads): overlapping calls on one device would race, so
 * the library makes them wait for each other within a process.  Same results
 * as drot_qr_solve_f32, bit for bit. */
#define DROT_NOALLOC_SCRATCH_BYTES (32u << 20)
drot_status drot_qr_solve_noalloc_f32(const float *c, const float *p, const float *q,
                                      int n_rows, int n_cols, const drot_qr_params *params,
                                      float *x, void *stream, drot_result *result);

/* GL-DROT solve on device memory.  Replaces group_lasso_regularizer_drot<float>
 * (core/drot_glr.hpp:242-355) as called by group_lasso_regularizer_drot_torch_float32
 * (pytorch/src/glr_drot_cuda_kernel.cu:13-86). */
drot_status drot_glr_solve_f32(const float *c, const float *p, const float *q, int n_rows,
                               int n_cols, const drot_glr_params *params, float *x,
                               void *workspace, size_t workspace_bytes, void *stream,
                               drot_result *result);

/* Export the dual state left in `workspace` by the last solve of this shape:
 * phi1[n_rows], phi2[n_cols] (device pointers, may be NULL), and on the host
 * alpha (may be NULL).  kind: 0 = QR, 1 = GL.  The reference keeps this state
 * internal (core/drot_qr.hpp:74-76); the potentials are u = phi1/s, v = phi2/s. */
drot_status drot_export_duals_f32(int kind, int n_rows, int n_cols, int n_groups,
                                  const void *workspace, size_t workspace_bytes,
                                  float *phi1, float *phi2, float *alpha, void *stream);

/* ---------------------------------------------------------------- batches */

/* One problem of a batch.  All problems' C (and X) live in one packed device
 * buffer: problem i is column-major n_rows x n_cols with leading dimension ld
 * (ld >= n_rows, ld % 4 == 0; rows n_rows..ld-1 are scratch) starting at float
 * offset mat_offset (a multiple of 4).  p and q are packed the same way at
 * p_offset / q_offset.  params are per problem (drot_qr_params_from_*). */
typedef struct drot_batch_problem {
    int n_rows, n_cols;
    int ld;
    long long mat_offset;
    long long p_offset, q_offset;
    drot_qr_params params;
} drot_batch_problem;

/* Fill shapes and a dense packed layout for a batch (ld = n_rows rounded up
 * to 32 floats, so every column starts on a 128-byte cache line and the
 * pass's stores write whole lines); params are left to the caller.  Outputs
 * the packed lengths. */
drot_status drot_batch_layout(int n_problems, const int *n_rows, const int *n_cols,
                              drot_batch_problem *out, long long *mat_elems,
                              long long *p_elems, long long *q_elems);

drot_status drot_qr_batch_workspace_size(int n_problems, const drot_batch_problem *problems,
                                         size_t *bytes);

/* Solve many QR-DROT problems of uneven shape together: one pass launch and
 * one finalize launch per iteration over a work-list of the still-active
 * problems (compacted as problems converge), each problem with its own
 * stopping test.  Replaces the reference's sequential loop over problems
 * (core/multiexp_qr.cu:35-55 calling quadratic_regularizer_drot_wrapper per
 * problem); each problem's result is that of drot_qr_solve_f32 on it up to
 * the grouping of the fp64 partial sums by the tile shape (same n_iter).
 * results: host array of n_problems (may be NULL). */
drot_status drot_qr_batch_solve_f32(int n_problems, const drot_batch_problem *problems,
                                    const float *c, const float *p, const float *q, float *x,
                                    void *workspace, size_t workspace_bytes, void *stream,
                                    drot_result *results);

/* ---------------------------------------------------------------- backward */

/* Options of the implicit-differentiation backward (all fields 0 = default). */
typedef struct drot_backward_params {
    float support_tol;      /* (i,j) is in the support when x_ij > 0 and
                               x_ij > support_tol * step_size * |c_ij|; default
                               2^-22, which rejects the rounding residue that
                               adding step*C back to the stored X - step*C
                               leaves where the plan is zero */
    int max_cg_iters;       /* default 4 * (n_rows + n_cols) + 100 */
    double cg_tol;          /* relative CG residual target; default 1e-10 */
    long long nnz_capacity; /* index capacity per problem; default n_rows*n_cols
                               up to 2^22, else max(2^22, 64 (n_rows + n_cols)) */
} drot_backward_params;

typedef struct drot_backward_info {
    long long nnz;          /* support size */
    int cg_iters;
    int converged;          /* CG reached cg_tol */
    double cg_residual;     /* final ||r|| / ||g|| */
} drot_backward_info;

/* Gradient of a loss L(X*, F) of a QR-DROT plan X* and of its regularized
 * objective F = <C,X*> + (reg/2)||X*||^2, reg = (scale - 1) / step_size, with
 * respect to C:
 *   dL/dC = S o (w 1^T + 1 w'^T - G) / reg + g_F X*,  G = dL/dX*,  g_F = dL/dF,
 * where S is the support of X* and [w; w'] solves the bipartite-Laplacian system
 * of the KKT conditions on S (Jacobi-preconditioned CG in fp64).  The reference
 * has no backward: its binding returns a tensor without gradient
 * (pytorch/src/qr_drot_cuda_kernel.cu:99), so this entry point has no
 * counterpart there.  c, x (the plan returned by drot_qr_solve_f32), grad_x and
 * grad_c are device matrices in the solver layout; grad_obj is a device pointer
 * to one float (NULL = 0).  params: the ones the plan was solved with.
 * bparams and info may be NULL.  Synchronous on `stream`. */
drot_status drot_qr_backward_workspace_size(int n_rows, int n_cols, const drot_qr_params *params,
                                            const drot_backward_params *bparams, size_t *bytes);
drot_status drot_qr_backward_f32(const float *c, const float *x, const float *grad_x,
                                 const float *grad_obj, int n_rows, int n_cols,
                                 const drot_qr_params *params, const drot_backward_params *bparams,
                                 float *grad_c, void *workspace, size_t workspace_bytes,
                                 void *stream, drot_backward_info *info);

/* Batched backward over packed problems (layout of drot_qr_batch_solve_f32;
 * ld needs no rounding here).  grad_obj: device array of n_problems floats or
 * NULL.  Padding rows of grad_c are set to 0.  One count and one fill launch
 * over all problems' column tiles, then one CG block per problem. */
drot_status drot_qr_batch_backward_workspace_size(int n_problems,
                                                  const drot_batch_problem *problems,
                                                  const drot_backward_params *bparams,
                                                  size_t *bytes);
drot_status drot_qr_batch_backward_f32(int n_problems, const drot_batch_problem *problems,
                                       const float *c, const float *x, const float *grad_x,
                                       const float *grad_obj, const drot_backward_params *bparams,
                                       float *grad_c, void *workspace, size_t workspace_bytes,
                                       void *stream, drot_backward_info *infos);

/* ---------------------------------------------------------------- tuning */

/* Tile of the QR pass kernel: vec = rows of a column per lane (1 or 4: one
 * access; 8: two 16-byte accesses, for n_rows a multiple of 4 and 16-byte
 * aligned buffers), lanes_per_col lanes share a column (16, 32, 64; 8 or 16
 * for vec 8), col_steps column steps per block, store_policy 2 = write-through
 * X stores, 0 = plain stores, unroll column steps loaded per batch (2 or 4;
 * 1 for vec 8).  Replaces the reference's launch-size
 * heuristic _q_get_work_size_update_x (core/drot_qr.hpp:236-239) and its
 * launch macros (core/param_qr.hpp:4-24). */
typedef struct drot_qr_tile {
    int vec, lanes_per_col, col_steps, store_policy, unroll;
} drot_qr_tile;

/* The tile a solve of this shape will use (tuned, else the built-in choice). */
drot_status drot_qr_tile_for(int n_rows, int n_cols, int vec4_ok, drot_qr_tile *out);

/* Autotuner: times every candidate tile on this problem (iterations 64 + k/5
 * .. 64 + k from the warm start, k = max(iters, about 4 ms of work), so the
 * plan's zero segments have formed; x is scratch)
 * and records the fastest for (n_rows, n_cols) in the process-wide tile cache,
 * which later solves of that shape use.  The tuned tile can change
 * drot_qr_workspace_size: query it after tuning.  Workspace:
 * drot_qr_autotune_workspace_size.  Synchronous on `stream`. */
drot_status drot_qr_autotune_workspace_size(int n_rows, int n_cols, size_t *bytes);
drot_status drot_qr_autotune_f32(const float *c, const float *p, const float *q, int n_rows,
                                 int n_cols, const drot_qr_params *params, float *x,
                                 void *workspace, size_t workspace_bytes, void *stream,
                                 int iters, drot_qr_tile *best, double *best_us_per_iter);

/* The cache as text, one "n_rows n_cols vec lanes_per_col col_steps
 * store_policy unroll us_per_iter" line per shape (load merges). */
drot_status drot_qr_tile_cache_set(int n_rows, int n_cols, const drot_qr_tile *tile);
drot_status drot_qr_tile_cache_save(const char *path);
drot_status drot_qr_tile_cache_load(const char *path);
void drot_qr_tile_cache_clear(void);

/* ------------------------------------------------------- cost matrices */

/* Squared-Euclidean cost between point clouds, written in the solver layout:
 * C[j * ld + i] = ||src_i - tgt_j||^2 for source point i < n_rows (the p side)
 * and target point j < n_cols (the q side), optionally divided by max C.  The
 * reference computes this on the host with POT before the solve
 * (benchmark/generate_ot.py:22-23, benchmark/generate_da.py:40-41: ot.dist
 * and C / C.max()); here it is a device step that can write straight into a
 * packed batch (mat_offset, ld as in drot_batch_problem). */
typedef struct drot_cost_problem {
    int n_rows, n_cols, ld;
    long long mat_offset;               /* float offset of C */
    long long src_offset, tgt_offset;   /* first point (row of the [*, dim] buffers) */
} drot_cost_problem;

/* dim <= 64. */
drot_status drot_sqdist_workspace_size(int n_problems, const drot_cost_problem *problems, int dim,
                                       size_t *bytes);

/* src, tgt: row-major [points, dim] device buffers.  normalize != 0 divides by
 * max C (1 if every distance is 0) and then needs the device outputs
 * scale[n_problems] (max C) and argmax[2 n_problems] (i*, j*; smallest
 * j*n_rows+i on ties), which the backward reads; they may be NULL otherwise.
 * Padding rows n_rows..ld-1 of a problem's last row tile are set to 0.
 * Asynchronous on `stream`. */
drot_status drot_sqdist_f32(int n_problems, const drot_cost_problem *problems, int dim,
                            const float *src, const float *tgt, int normalize, float *c,
                            float *scale, int *argmax, void *workspace, size_t workspace_bytes,
                            void *stream);

/* Gradient of a loss through C to the points: grad_src, grad_tgt receive
 * dL/dsrc, dL/dtgt (same layout as src, tgt) for grad_c = dL/dC, including
 * the derivative of max C through (i*, j*) when normalize != 0.  Asynchronous. */
drot_status drot_sqdist_backward_f32(int n_problems, const drot_cost_problem *problems, int dim,
                                     const float *src, const float *tgt, const float *grad_c,
                                     int normalize, const float *scale, const int *argmax,
                                     float *grad_src, float *grad_tgt, void *workspace,
                                     size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------- host API */

/* Host-pointer wrapper with the argument order of
 * quadratic_regularizer_drot_wrapper<float> (core/drot_qr.hpp:241-257).
 * step_size is already normalised (rho/(n_rows+n_cols)).  Returns a malloc'd
 * host copy of X when return_x != 0 (free it with drot_free), else NULL; NULL
 * also on error (see drot_last_error). */
float *drot_qr_wrapper_f32(const float *c, const float *p, const float *q, int n_rows,
                           int n_cols, float step_size, float r_weight, int max_iters,
                           float eps, float *dur_in_ms, float *prep_dur_in_ms,
                           int *n_iter, float *objective, int use_warmup_init,
                           int return_x);

/* Host-pointer wrapper with the argument order of
 * group_lasso_regularizer_drot_wrapper<float> (core/drot_glr.hpp:362-379);
 * lambda = r_weight / sqrt(n_rows / n_groups) as in core/drot_glr.hpp:385. */
float *drot_glr_wrapper_f32(const float *c, const float *p, const float *q, int n_rows,
                            int n_cols, int n_groups, float step_size, float r_weight,
                            int max_iters, float eps, float *dur_in_ms,
                            float *prep_dur_in_ms, int *n_iter, float *objective,
                            int use_warmup_init, int return_x);

void drot_free(void *ptr);

/* ---------------------------------------------------------------- misc */

const char *drot_status_string(drot_status status);
/* Message of the last failing call on this thread ("" if none). */
const char *drot_last_error(void);
int drot_abi_version(void);

/* Timing mode for the calling thread: when enabled, solves bracket every fused
 * pass launch with HIP events on the solve's stream and report the per-pass
 * device time in drot_result.pass_ms.  Off by default. */
void drot_set_timing(int enable);

#ifdef __cplusplus
}
#endif

#endif /* DROT_H_ */
