// Batched QR-DROT (gfx950): many independent problems of uneven shape advance
// together, one pass launch and one finalize launch per iteration over a tile
// work-list of the still-active problems.  Replaces the reference's sequential
// host loop over problems (core/multiexp_qr.cu:35-55, which reallocates and
// re-uploads per problem and runs each problem's 9+ launches per iteration).
// Every tile runs the same qr_tile body as the single-problem solver (64-row,
// 256-column tiles, plain X stores, zero-segment flags per problem) and each
// problem keeps its own SolverState, dual vectors, partials and stopping test,
// so a problem's result is that of solving it alone up to the grouping of the
// fp64 partial sums by the tile shape: the same n_iter, plans within 1e-6.
#include "drot_finalize.h"
#include "drot_qr_tile.h"

#include <math.h>

namespace drot {

namespace {

__device__ inline IterCtl batch_ctl(const QrBatchArgs &A, const BatchProblem &P, int prob) {
    IterCtl c;
    c.st = A.st + prob;
    c.pfin = A.pfin + 3 * (size_t)P.fin_off;
    c.n_fin = P.n_fin;
    c.k = A.k;
    c.max_iters = P.max_iters;
    c.eps = P.eps;
    c.stop_rule = kStopEvenStrict;
    c.obj_even_only = 1;
    c.n_rows = P.n_rows;
    c.n_cols = P.n_cols;
    c.tspan = nullptr;
    return c;
}

}  // namespace

// Warm start (pytorch/src/qr_drot_cuda_kernel.cu:45-71) or cold start
// (core/drot_qr.hpp:308-320) of every problem; grid.y = problem.
__global__ __launch_bounds__(kBlock) void batch_init_kernel(QrBatchArgs A) {
    const int prob = blockIdx.y;
    const BatchProblem P = A.probs[prob];
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nthr = (size_t)gridDim.x * blockDim.x;
    const size_t total = (size_t)P.ld * (size_t)P.n_cols;
    float *x = A.x + P.mat_off;
    const float *p = A.p + P.row_off;
    const float *q = A.q + P.col_off;
    for (size_t e = tid; e < total; e += nthr) {
        float v = 0.0f;
        if (!P.warm) {
            const int j = (int)(e / (size_t)P.ld);
            const int i = (int)(e - (size_t)j * (size_t)P.ld);
            v = (i < P.n_rows) ? p[i] * q[j] : 0.0f;
        }
        x[e] = v;
    }
    for (size_t i = tid; i < (size_t)P.n_rows; i += nthr) {
        A.phi1[P.row_off + i] = P.warm ? P.v_phi1 : 0.0f;
        A.a[P.row_off + i] = P.warm ? P.v_a : 0.0f;
        A.ahat[P.row_off + i] = 0.0f;
    }
    for (size_t j = tid; j < (size_t)P.n_cols; j += nthr) {
        A.phi2[P.col_off + j] = P.warm ? P.v_phi2 : 0.0f;
        A.b[P.col_off + j] = P.warm ? P.v_b : 0.0f;
        A.bhat[P.col_off + j] = 0.0f;
    }
    if (A.zf) {  // no segment starts out flagged
        const size_t nz = (size_t)P.n_rt * (size_t)P.n_cols;
        for (size_t i = tid; i < nz; i += nthr) {
            A.zf[P.zf_off + i] = 0;
            A.zf[A.zf_len + P.zf_off + i] = 0;
        }
    }
    if (A.seg_e) {  // no headroom known yet: nothing is screened before its C is read
        const size_t nz = (size_t)P.n_rt * (size_t)P.n_cols;
        for (size_t i = tid; i < nz; i += nthr) A.seg_e[P.zf_off + i] = INFINITY;
        const size_t nt = (size_t)P.n_rt * (size_t)P.n_ct;
        for (size_t i = tid; i < nt; i += nthr) A.g1[P.tile_off + i] = 0.0;
    }
    if (tid == 0) {
        SolverState s;
        s.done = 0;
        s.converged = 0;
        s.n_iter = 0;
        s.k_run = 0;
        s.gamma = 0.0f;
        s.alpha = P.warm ? P.v_alpha : 0.0f;
        s.alpha_slot[0] = 0.0f;
        s.alpha_slot[1] = s.alpha;
        s.phi_mode = 1;
        s.last_k = -1;
        s.obj = 0.0;
        s.res = 1e9;
        s.xsum = 0.0;
        s.res_last = 1e9;
        A.st[prob] = s;
    }
}

// One pass over every listed tile; the list holds the tiles of the active
// problems in row-tile-major order per problem, and blocks take contiguous
// runs of it per XCD (see qr_pass_kernel).
template <bool EVEN, bool ZS, bool SC, int VEC, int LPC>
__global__ __launch_bounds__(kBlock) void batch_pass_kernel(QrBatchArgs A) {
    constexpr int CPW = kWave / LPC;
    static_assert(VEC * LPC == kBatchTileRows, "64-row batch tiles");
    __shared__ double s_row[kWavesPerBlock * CPW][kBatchTileRows];
    __shared__ double s_red[3 * kWavesPerBlock];
    const int per_xcd = (A.n_tiles + kXcds - 1) / kXcds;
    const int t = (int)(blockIdx.x % kXcds) * per_xcd + (int)(blockIdx.x / kXcds);
    if (t >= A.n_tiles) return;
    const BatchTile bt = A.tiles[t];
    const int prob = bt.prob;
    if (A.st[prob].done) return;
    const BatchProblem P = A.probs[prob];
    QrTileJob J;
    J.x = A.x + P.mat_off;
    J.c = A.c + P.mat_off;
    J.phi1 = A.phi1 + P.row_off;
    J.phi2 = A.phi2 + P.col_off;
    J.ahat = A.ahat + P.row_off;
    J.bhat = A.bhat + P.col_off;
    J.prow = A.prow + P.prow_off;
    J.pcol = A.pcol + P.pcol_off;
    J.pobj = A.pobj + P.tile_off + bt.rt + bt.ct * P.n_rt;
    J.n_rows = P.n_rows;
    J.n_cols = P.n_cols;
    J.ld = P.ld;
    J.rt = bt.rt;
    J.ct = bt.ct;
    J.nstep = P.nstep;
    J.prio = 0;
    J.tc = P.nstep * kWavesPerBlock * CPW;
    J.step = P.step;
    J.kappa = P.kappa;
    J.zin = ZS ? A.zf + (size_t)((A.k + 1) & 1) * A.zf_len + P.zf_off : nullptr;
    J.zout = ZS ? A.zf + (size_t)(A.k & 1) * A.zf_len + P.zf_off : nullptr;
    J.seg_e = SC ? A.seg_e + P.zf_off : nullptr;
    J.g1 = SC ? A.g1 + P.tile_off + bt.rt + bt.ct * P.n_rt : nullptr;
    J.f1_prev = SC ? A.f1pb + (size_t)((A.k + 1) & 1) * A.row_len + P.row_off : nullptr;
    J.f1_out = SC ? A.f1pb + (size_t)(A.k & 1) * A.row_len + P.row_off : nullptr;
    const IterCtl ctl = batch_ctl(A, P, prob);
    // plain X stores: a batch's footprint lives in HBM, and write-through stores
    // made the mixed 256-problem forward 5% slower (0.99 vs 0.94 s).  One
    // column step per load batch: with screening the loop then visits only the
    // steps a wave does not screen whole (mixed forward 722 vs 732 ms against
    // two steps per batch; 64 x 512^2 unchanged)
    qr_tile<VEC, LPC, EVEN, 0, 1, ZS, SC>(J, ctl, bt.rt == 0 && bt.ct == 0, s_row, s_red);
}

__global__ __launch_bounds__(kBlock) void batch_finalize_kernel(QrBatchArgs A) {
    __shared__ double s_red[3 * kWavesPerBlock];
    const BatchFin f = A.fins[blockIdx.x];
    const BatchProblem P = A.probs[f.prob];
    FinJob F;
    F.done = &A.st[f.prob].done;
    F.prow = A.prow + P.prow_off;
    F.pcol = A.pcol + P.pcol_off;
    F.pobj = A.pobj + P.tile_off;
    F.p = A.p + P.row_off;
    F.q = A.q + P.col_off;
    F.a = A.a + P.row_off;
    F.b = A.b + P.col_off;
    F.ahat = A.ahat + P.row_off;
    F.bhat = A.bhat + P.col_off;
    F.pfin = A.pfin + 3 * ((size_t)P.fin_off + f.blk);
    F.n_rows = P.n_rows;
    F.n_cols = P.n_cols;
    F.n_rt = P.n_rt;
    F.n_ct = P.n_ct;
    F.n_obj = P.n_rt * P.n_ct;
    F.rb = P.rb;
    F.n_fin = P.n_fin;
    F.blk = f.blk;
    F.with_obj = (A.k % 2) == 0;
    finalize_rows_block(F, s_red);
}

// Per problem, after the loop; grid.y = problem.  First the segments that the
// problem's last executed pass (k_run - 1) left flagged get the value that pass
// would have stored (fmaf(-s, c, +0) after an even pass, +0 after an odd one;
// see drot_qr_tile.h), then core/kernel_qr.hpp:469-482 if it converged.
__global__ __launch_bounds__(kBlock) void batch_fix_x_kernel(QrBatchArgs A) {
    const int prob = blockIdx.y;
    const SolverState st = A.st[prob];
    const BatchProblem P = A.probs[prob];
    const int k_last = st.k_run - 1;
    const unsigned char *z =
        (A.zf && k_last >= 0) ? A.zf + (size_t)(k_last & 1) * A.zf_len + P.zf_off : nullptr;
    if (!st.converged && !z) return;
    const bool even = (k_last & 1) == 0;
    const size_t total = (size_t)P.ld * (size_t)P.n_cols;
    float *x = A.x + P.mat_off;
    const float *c = A.c + P.mat_off;
    const size_t nthr = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += nthr) {
        const int j = (int)(e / (size_t)P.ld);
        const int i = (int)(e - (size_t)j * (size_t)P.ld);
        const bool flagged =
            z && i < P.n_rows && z[(size_t)(i / kBatchTileRows) * (size_t)P.n_cols + j];
        if (!flagged && !st.converged) continue;
        float v = flagged ? (even ? __builtin_fmaf(-P.step, c[e], 0.0f) : 0.0f) : x[e];
        if (st.converged) v = fmaxf(__builtin_fmaf(P.step, c[e], v), 0.0f);
        x[e] = v;
    }
}

int batch_pass_grid(int n_tiles) {
    const int per = (n_tiles + kXcds - 1) / kXcds;
    return per * kXcds;
}

hipError_t launch_batch_init(const QrBatchArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(batch_init_kernel, dim3(64, a.n_probs), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_batch_pass(bool even, const QrBatchArgs &a, hipStream_t s) {
    if (a.n_tiles <= 0) return hipSuccess;
    const int grid = batch_pass_grid(a.n_tiles);
    auto k = a.zf ? (a.seg_e ? (even ? batch_pass_kernel<true, true, true, kBatchVec, kBatchLpc>
                                     : batch_pass_kernel<false, true, true, kBatchVec, kBatchLpc>)
                             : (even ? batch_pass_kernel<true, true, false, kBatchVec, kBatchLpc>
                                     : batch_pass_kernel<false, true, false, kBatchVec, kBatchLpc>))
                  : (even ? batch_pass_kernel<true, false, false, kBatchVec, kBatchLpc>
                          : batch_pass_kernel<false, false, false, kBatchVec, kBatchLpc>);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_batch_finalize(const QrBatchArgs &a, hipStream_t s) {
    if (a.n_fins <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_finalize_kernel, dim3(a.n_fins), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_batch_fix_x(const QrBatchArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(batch_fix_x_kernel, dim3(64, a.n_probs), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace drot
