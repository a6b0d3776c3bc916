// Cost matrices from point clouds (gfx950): C_ij = ||src_i - tgt_j||^2,
// optionally divided by its maximum, and the gradient of a loss through C back
// to both point sets.
//
// The reference builds C on the host with POT before calling the solver
// (benchmark/generate_ot.py:22-23, benchmark/generate_da.py:40-41: ot.dist,
// then C / C.max()).  Here it is a device step in front of the solver that
// writes C straight into the solver's (packed) layout, so differentiable
// pipelines (points -> C -> plan -> loss) stay on the GPU.
//
// Work unit: a tile of kCostTile source x kCostTile target points; both point
// sets of the tile are staged in LDS.  Thread t owns source point i0 + (t & 63)
// and target points j0 + (t >> 6) + 4 r, r < 16, so every store of C is a
// coalesced 256-byte row segment.  D is accumulated over dimensions in a fixed
// order with fused multiply-adds; the normalised C is D / max D with a
// correctly rounded division.  The max is found in a first pass (no stores),
// reduced per problem with a fixed tie-break (smallest index j*n_rows+i), and
// the second pass writes D / max D: C is written once.
//
// Backward with G = dL/dC (mu = max D, T = sum G o C, (i*, j*) the argmax):
//   dL/dD = G / mu - [ij = i*j*] T / mu
//   dL/dsrc_i = 2/mu (src_i sum_j G_ij - sum_j G_ij tgt_j)  (+ argmax term)
//   dL/dtgt_j = 2/mu (tgt_j sum_i G_ij - sum_i G_ij src_i)  (+ argmax term)
// One pass reads G once per tile, stages it in LDS and produces per-tile fp64
// partial sums for both point sets; a final kernel sums the partials of each
// point in tile order (deterministic) and adds the argmax term.
#include "drot_kernels.h"

#include <limits.h>

namespace drot {

namespace {

constexpr int kRowsPerThread = kCostTile / kWavesPerBlock;  // 16
constexpr int kLdsPitch = kCostMaxDim + 1;

struct MaxPair {
    float v;
    int idx;
};

__device__ __forceinline__ MaxPair better(MaxPair a, MaxPair b) {
    if (b.v > a.v || (b.v == a.v && b.idx < a.idx)) return b;
    return a;
}

__device__ __forceinline__ MaxPair wave_best(MaxPair m) {
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) {
        MaxPair o;
        o.v = __shfl_xor(m.v, off, kWave);
        o.idx = __shfl_xor(m.idx, off, kWave);
        m = better(m, o);
    }
    return m;
}

// Stage the tile's source and target points (zeros outside the problem).
__device__ __forceinline__ void stage_points(const CostArgs &A, const CostProblem &P, int i0,
                                             int j0, float (*sx)[kLdsPitch],
                                             float (*sy)[kLdsPitch]) {
    const int dim = A.dim;
    for (int f = threadIdx.x; f < kCostTile * dim; f += kBlock) {
        const int pt = f / dim, k = f - pt * dim;
        const int i = i0 + pt, j = j0 + pt;
        sx[pt][k] = i < P.n_rows ? A.src[(P.src_off + i) * dim + k] : 0.0f;
        sy[pt][k] = j < P.n_cols ? A.tgt[(P.tgt_off + j) * dim + k] : 0.0f;
    }
}

// D of the thread's 16 elements of the tile.
__device__ __forceinline__ void tile_dist(int dim, const float (*sx)[kLdsPitch],
                                          const float (*sy)[kLdsPitch], int il, int w,
                                          float (&acc)[kRowsPerThread]) {
#pragma unroll
    for (int r = 0; r < kRowsPerThread; r++) acc[r] = 0.0f;
    for (int k = 0; k < dim; k++) {
        const float xv = sx[il][k];
#pragma unroll
        for (int r = 0; r < kRowsPerThread; r++) {
            const float d = sy[w + kWavesPerBlock * r][k] - xv;
            acc[r] = fmaf(d, d, acc[r]);
        }
    }
}

}  // namespace

// MODE 0: write D.  MODE 1: per-tile max only.  MODE 2: write D / scale.
template <int MODE>
__global__ __launch_bounds__(kBlock) void cost_tile_kernel(CostArgs A) {
    __shared__ float sx[kCostTile][kLdsPitch];
    __shared__ float sy[kCostTile][kLdsPitch];
    __shared__ MaxPair s_best[kWavesPerBlock];
    const CostTile T = A.tiles[blockIdx.x];
    const CostProblem &P = A.probs[T.prob];
    const int i0 = T.it * kCostTile, j0 = T.jt * kCostTile;
    stage_points(A, P, i0, j0, sx, sy);
    __syncthreads();
    const int il = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    float acc[kRowsPerThread];
    tile_dist(A.dim, sx, sy, il, w, acc);
    const int i = i0 + il;
    if (MODE == 1) {
        MaxPair m{-1.0f, INT_MAX};
#pragma unroll
        for (int r = 0; r < kRowsPerThread; r++) {
            const int j = j0 + w + kWavesPerBlock * r;
            if (i < P.n_rows && j < P.n_cols) m = better(m, MaxPair{acc[r], j * P.n_rows + i});
        }
        m = wave_best(m);
        if (il == 0) s_best[w] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            MaxPair b = s_best[0];
            for (int k = 1; k < kWavesPerBlock; k++) b = better(b, s_best[k]);
            A.tmax[blockIdx.x] = b.v;
            A.targ[blockIdx.x] = b.idx;
        }
        return;
    }
    const float mu = MODE == 2 ? A.scale[T.prob] : 1.0f;
    const bool store_row = i < P.ld && i < i0 + kCostTile;
#pragma unroll
    for (int r = 0; r < kRowsPerThread; r++) {
        const int j = j0 + w + kWavesPerBlock * r;
        if (j < P.n_cols && store_row) {
            float v = 0.0f;  // padding rows n_rows..ld-1 are zeroed
            if (i < P.n_rows) v = MODE == 2 ? acc[r] / mu : acc[r];
            A.c[P.mat_off + (long long)j * P.ld + i] = v;
        }
    }
}

// One block per problem: max over the problem's tiles -> scale, argmax.
__global__ __launch_bounds__(kBlock) void cost_max_kernel(CostArgs A) {
    __shared__ MaxPair s_best[kWavesPerBlock];
    const CostProblem &P = A.probs[blockIdx.x];
    MaxPair m{-1.0f, INT_MAX};
    for (int t = threadIdx.x; t < P.n_it * P.n_jt; t += kBlock)
        m = better(m, MaxPair{A.tmax[P.tile_begin + t], A.targ[P.tile_begin + t]});
    m = wave_best(m);
    if ((threadIdx.x & (kWave - 1)) == 0) s_best[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        MaxPair b = s_best[0];
        for (int k = 1; k < kWavesPerBlock; k++) b = better(b, s_best[k]);
        // all points coincident: keep C = D = 0 instead of 0/0
        A.scale[blockIdx.x] = b.v > 0.0f ? b.v : 1.0f;
        A.argmax[2 * blockIdx.x] = b.idx % P.n_rows;
        A.argmax[2 * blockIdx.x + 1] = b.idx / P.n_rows;
    }
}

// Per-tile partials of the backward (see the file comment).
__global__ __launch_bounds__(kBlock) void cost_bwd_tile_kernel(CostArgs A) {
    __shared__ float sx[kCostTile][kLdsPitch];
    __shared__ float sy[kCostTile][kLdsPitch];
    __shared__ float sg[kCostTile][kCostTile + 1];
    __shared__ double s_red[kWavesPerBlock];
    const CostTile T = A.tiles[blockIdx.x];
    const CostProblem &P = A.probs[T.prob];
    const int i0 = T.it * kCostTile, j0 = T.jt * kCostTile;
    const int dim = A.dim;
    stage_points(A, P, i0, j0, sx, sy);
    __syncthreads();
    const int il = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int i = i0 + il;
    float acc[kRowsPerThread];
    if (A.normalize) tile_dist(dim, sx, sy, il, w, acc);
    const float mu = A.normalize ? A.scale[T.prob] : 1.0f;
    double tdot = 0.0;
#pragma unroll
    for (int r = 0; r < kRowsPerThread; r++) {
        const int jl = w + kWavesPerBlock * r;
        const int j = j0 + jl;
        float g = 0.0f;
        if (i < P.n_rows && j < P.n_cols) {
            g = A.grad_c[P.mat_off + (long long)j * P.ld + i];
            if (A.normalize) tdot += (double)g * (double)(acc[r] / mu);
        }
        sg[jl][il] = g;
    }
    __syncthreads();
    const int pt = threadIdx.x >> 2, q = threadIdx.x & 3;
    constexpr int kPer = kCostMaxDim / 4;
    // source side: sum over the tile's target points
    {
        double s = 0.0, v[kPer];
#pragma unroll
        for (int m = 0; m < kPer; m++) v[m] = 0.0;
        for (int jl = 0; jl < kCostTile; jl++) {
            const double g = (double)sg[jl][pt];
            s += g;
#pragma unroll
            for (int m = 0; m < kPer; m++) {
                const int k = q + 4 * m;
                if (k < dim) v[m] += g * (double)sy[jl][k];
            }
        }
        const int ii = i0 + pt;
        if (ii < P.n_rows) {
            double *o = A.xpart + P.xpart_off + ((long long)T.jt * P.n_rows + ii) * (dim + 1);
            if (q == 0) o[0] = s;
#pragma unroll
            for (int m = 0; m < kPer; m++) {
                const int k = q + 4 * m;
                if (k < dim) o[1 + k] = v[m];
            }
        }
    }
    // target side: sum over the tile's source points
    {
        double s = 0.0, v[kPer];
#pragma unroll
        for (int m = 0; m < kPer; m++) v[m] = 0.0;
        for (int ii = 0; ii < kCostTile; ii++) {
            const double g = (double)sg[pt][ii];
            s += g;
#pragma unroll
            for (int m = 0; m < kPer; m++) {
                const int k = q + 4 * m;
                if (k < dim) v[m] += g * (double)sx[ii][k];
            }
        }
        const int jj = j0 + pt;
        if (jj < P.n_cols) {
            double *o = A.ypart + P.ypart_off + ((long long)T.it * P.n_cols + jj) * (dim + 1);
            if (q == 0) o[0] = s;
#pragma unroll
            for (int m = 0; m < kPer; m++) {
                const int k = q + 4 * m;
                if (k < dim) o[1 + k] = v[m];
            }
        }
    }
    const double t = block_sum(tdot, s_red);
    if (threadIdx.x == 0) A.tdot[blockIdx.x] = t;
}

// grid (chunks of kBlock points, n_probs): point p < n_rows is source p, else
// target p - n_rows.
__global__ __launch_bounds__(kBlock) void cost_bwd_final_kernel(CostArgs A) {
    const CostProblem &P = A.probs[blockIdx.y];
    const int pidx = blockIdx.x * kBlock + threadIdx.x;
    if (pidx >= P.n_rows + P.n_cols) return;
    const int dim = A.dim;
    const bool is_src = pidx < P.n_rows;
    const int idx = is_src ? pidx : pidx - P.n_rows;
    const double mu = A.normalize ? (double)A.scale[blockIdx.y] : 1.0;
    const int n_parts = is_src ? P.n_jt : P.n_it;
    const long long stride = (long long)(is_src ? P.n_rows : P.n_cols) * (dim + 1);
    const double *part = (is_src ? A.xpart + P.xpart_off : A.ypart + P.ypart_off) +
                         (long long)idx * (dim + 1);
    const float *self = is_src ? A.src + (P.src_off + idx) * dim : A.tgt + (P.tgt_off + idx) * dim;
    float *out = is_src ? A.grad_src + (P.src_off + idx) * dim : A.grad_tgt + (P.tgt_off + idx) * dim;
    double s = 0.0;
    for (int t = 0; t < n_parts; t++) s += part[t * stride];
    // argmax term: dL/dD gains -T/mu at (i*, j*)
    double corr = 0.0;
    const float *other = nullptr;
    if (A.normalize) {
        const int is = A.argmax[2 * blockIdx.y], js = A.argmax[2 * blockIdx.y + 1];
        if ((is_src && idx == is) || (!is_src && idx == js)) {
            double tsum = 0.0;
            for (int t = 0; t < P.n_it * P.n_jt; t++) tsum += A.tdot[P.tile_begin + t];
            corr = -tsum / mu;
            other = is_src ? A.tgt + (P.tgt_off + js) * dim : A.src + (P.src_off + is) * dim;
        }
    }
    for (int k = 0; k < dim; k++) {
        double v = 0.0;
        for (int t = 0; t < n_parts; t++) v += part[t * stride + 1 + k];
        const double xk = (double)self[k];
        double gk = 2.0 / mu * (xk * s - v);
        if (other) gk += 2.0 * corr * (xk - (double)other[k]);
        out[k] = (float)gk;
    }
}

// ----------------------------------------------------------------- launch

hipError_t launch_cost_forward(const CostArgs &a, hipStream_t s) {
    if (a.normalize) {
        hipLaunchKernelGGL(cost_tile_kernel<1>, dim3(a.n_tiles), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(cost_max_kernel, dim3(a.n_probs), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(cost_tile_kernel<2>, dim3(a.n_tiles), dim3(kBlock), 0, s, a);
    } else {
        hipLaunchKernelGGL(cost_tile_kernel<0>, dim3(a.n_tiles), dim3(kBlock), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_cost_backward(const CostArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(cost_bwd_tile_kernel, dim3(a.n_tiles), dim3(kBlock), 0, s, a);
    const int chunks = (a.max_points + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(cost_bwd_final_kernel, dim3(chunks, a.n_probs), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace drot
