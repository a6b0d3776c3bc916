// Device-side shared definitions of the DR regularized-OT solver (gfx950).
//
// Matrix layout follows the reference (core/kernel_qr.hpp:3): column-major
// n_rows x n_cols, element (i, j) at x[j * n_rows + i], where n_rows = len(p) is
// the contiguous axis and n_cols = len(q) the strided one.  The PyTorch tensor
// [M, N] = [n_cols, n_rows] has exactly this memory.
//
// Arithmetic contract (shared with the CPU oracle):
//   * per-element updates in fp32 with a fixed operation order and explicit
//     fused multiply-adds where the reference's nvcc build contracts `t - s*c`;
//   * every reduction (row/column sums, total sum, objective, residual, group
//     norms) accumulates in fp64 through fixed-shape trees, so the result is
//     deterministic and equals any other fp64 accumulation after the final
//     rounding to fp32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drot {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // threads per block for every solver kernel
constexpr int kWavesPerBlock = kBlock / kWave;

// Solver state kept in device memory for the whole solve.  The host never
// reads it inside the iteration loop except through asynchronous snapshots.
struct alignas(64) SolverState {
    int k;            // index of the next iteration to run
    int done;         // 1 once stopped (converged or max_iters exhausted)
    int converged;    // 1 if the stopping test passed
    int phi_mode;     // 1: phi1/phi2 arrays are explicit; 0: phi = (hat + gamma)/n
    float gamma;      // shift of the last dual update (core/kernel_qr.hpp:421)
    float alpha;      // aux[4] of the reference (core/drot_qr.hpp:74-76)
    int n_iter;       // iteration count in the reference's convention
    unsigned int counter;  // arrival counter of the finalize kernel
    double obj;       // objective of the last iteration that computed it
    double res;       // primal residual of the last checked iteration
    double xsum;      // sum of X' of the last iteration
    double res_last;  // primal residual of the last iteration (any parity)
    double obj_last;  // objective of the last iteration (0 on QR odd iterations)
    int last_k;       // index of the last executed iteration (-1 before any)
    int pad_[3];
};

// Which stopping rule the finalize kernel applies.
enum StopRule : int {
    kStopEvenStrict = 0,  // QR: test on even k only, residual < eps (core/drot_qr.hpp:217-233)
    kStopEveryLeq = 1,    // GL: test every k, residual <= eps (core/drot_glr.hpp:344-354)
};

// ---------------------------------------------------------------- reductions

// Butterfly sum over aligned segments of `width` lanes (width a power of two,
// <= 64).  Every lane of a segment ends with the same, order-fixed result.
template <int WIDTH>
__device__ __forceinline__ double seg_sum(double v) {
#pragma unroll
    for (int off = WIDTH / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) { return seg_sum<kWave>(v); }

// Fixed-order block-wide sum of one double per thread (kBlock threads).
// `scratch` holds kWavesPerBlock doubles.  Result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) scratch[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; i++) t += scratch[i];
    __syncthreads();
    return t;
}

__device__ __forceinline__ int wave_uniform(int v) {
    return __builtin_amdgcn_readfirstlane(v);
}

}  // namespace drot
