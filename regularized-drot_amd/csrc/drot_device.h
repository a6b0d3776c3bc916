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
constexpr int kXcds = 8;           // accelerator complex dies (separate L2s) on MI355X

// Solver state kept in device memory for the whole solve.  The host never
// reads it inside the iteration loop except through asynchronous snapshots.
struct alignas(64) SolverState {
    int done;             // 1 once stopped (converged or max_iters exhausted)
    int converged;        // 1 if the stopping test passed
    int n_iter;           // iteration count in the reference's convention
    int k_run;            // iterations executed
    float gamma;          // gamma of the last dual update (core/kernel_qr.hpp:421)
    float alpha;          // alpha after the last dual update (aux[4], core/drot_qr.hpp:74-76)
    float alpha_slot[2];  // alpha after iteration j lives in slot j % 2 (slot 1: initial)
    int phi_mode;         // 1 until the first dual update: phi arrays are explicit
    int last_k;           // index of the last finished iteration (-1 before any)
    double obj;           // objective of the last iteration that computed it
    double res;           // primal residual of the last iteration that was tested
    double xsum;          // sum of X' of the last iteration
    double res_last;      // primal residual of the last iteration (any parity)
};

// Which stopping rule the solve applies.
enum StopRule : int {
    kStopEvenStrict = 0,  // QR: test on even k only, residual < eps (core/drot_qr.hpp:217-233)
    kStopEveryLeq = 1,    // GL: test every k, residual <= eps (core/drot_glr.hpp:344-354)
};

// Per-launch iteration control shared by the pass kernels.
struct IterCtl {
    SolverState *st;
    const double *pfin;  // [n_fin][3]: residual part, sum-X part, objective part
    int n_fin;
    int k;               // iteration this launch runs
    int max_iters;
    float eps;
    int stop_rule;       // StopRule
    int obj_even_only;   // QR: objective only on even iterations
    int n_rows, n_cols;
    // timing mode, sampled launches only (else null): block b stores its start
    // and end on the constant-rate device clock at tspan[2b], tspan[2b+1]
    unsigned long long *tspan;
};

// Records this block's execution span (plain stores, one slot pair per block).
__device__ __forceinline__ void record_span(unsigned long long *tspan, unsigned long long t0) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t1 = (unsigned long long)wall_clock64();
        tspan[2 * blockIdx.x] = t0;
        tspan[2 * blockIdx.x + 1] = t1;
    }
}

// ---------------------------------------------------------------- reductions

// A double moved across lanes by a DPP control (two 32-bit DPP moves).  Only
// controls that give every lane a source (row rotations, quad permutations)
// are used, so no lane keeps an old value and none needs initialising.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// v + (v of lane i ^ 32) or (lane i ^ 16) through gfx950's row swaps: with
// both operands equal, v_permlane32_swap leaves {v_i, v_(i^32)} in the two
// results of every lane i, v_permlane16_swap {v_i, v_(i^16)}.
__device__ __forceinline__ double xor32_add(double v) {
    const long long u = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(u >> 32), (unsigned)(u >> 32),
                                                     false, false);
    const double a = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
    const double b = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
    return a + b;
}

__device__ __forceinline__ double xor16_add(double v) {
    const long long u = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(u >> 32), (unsigned)(u >> 32),
                                                     false, false);
    const double a = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
    const double b = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
    return a + b;
}

// Butterfly sum over aligned segments of `width` lanes (16, 32 or 64).  Lane
// i adds lane i ^ off for off = width/2, ..., 2, 1; every lane of a segment
// ends with the same, order-fixed result (the two operands of each addition
// are the same pair in every lane of it, so the order inside a pair does not
// change the bits).  All in registers: the row swaps for offsets 32 and 16,
// row_ror:8 for xor 8 within a 16-lane row, row_ror:4 for xor 4 once the
// values repeat every 8 lanes (after the xor-8 step), quad_perm for 2 and 1.
// An 8-lane segment (WIDTH == 8) starts with row_half_mirror, which pairs lane
// i with lane 7 - i of its half-row: afterwards the lanes i and 7 - i hold the
// same sum (an addition commutes bit for bit), so the quad steps that follow
// leave the same total in all eight lanes.
template <int WIDTH>
__device__ __forceinline__ double seg_sum(double v) {
    static_assert(WIDTH == 8 || WIDTH == 16 || WIDTH == 32 || WIDTH == 64,
                  "segment of 8, 16, 32 or 64 lanes");
    if constexpr (WIDTH == 8) {
        v += dpp_move<0x141>(v);  // row_half_mirror
        v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
        v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
        return v;
    }
    if constexpr (WIDTH >= 64) v = xor32_add(v);
    if constexpr (WIDTH >= 32) v = xor16_add(v);
    v += dpp_move<0x128>(v);  // row_ror:8
    v += dpp_move<0x124>(v);  // row_ror:4
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    return v;
}

__device__ __forceinline__ double wave_sum(double v) { return seg_sum<kWave>(v); }

// Maximum over aligned segments of `width` lanes (8, 16, 32 or 64), in every
// lane of the segment; same lane movements as seg_sum (a maximum does not
// depend on the order).
__device__ __forceinline__ float fmax_xor32(float v) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v),
                                                    false, false);
    return fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

__device__ __forceinline__ float fmax_xor16(float v) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v),
                                                    false, false);
    return fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

__device__ __forceinline__ double fmax_xor32(double v) {
    const long long u = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(u >> 32), (unsigned)(u >> 32),
                                                     false, false);
    const double a = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
    const double b = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
    return fmax(a, b);
}

__device__ __forceinline__ double fmax_xor16(double v) {
    const long long u = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(u >> 32), (unsigned)(u >> 32),
                                                     false, false);
    const double a = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
    const double b = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
    return fmax(a, b);
}

template <int CTRL>
__device__ __forceinline__ float dpp_move_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

template <int WIDTH>
__device__ __forceinline__ float seg_max(float v) {
    static_assert(WIDTH == 8 || WIDTH == 16 || WIDTH == 32 || WIDTH == 64,
                  "segment of 8, 16, 32 or 64 lanes");
    if constexpr (WIDTH == 8) {
        v = fmaxf(v, dpp_move_f<0x141>(v));  // row_half_mirror
        v = fmaxf(v, dpp_move_f<0x4E>(v));   // quad_perm [2,3,0,1]
        v = fmaxf(v, dpp_move_f<0xB1>(v));   // quad_perm [1,0,3,2]
        return v;
    }
    if constexpr (WIDTH >= 64) v = fmax_xor32(v);
    if constexpr (WIDTH >= 32) v = fmax_xor16(v);
    v = fmaxf(v, dpp_move_f<0x128>(v));  // row_ror:8
    v = fmaxf(v, dpp_move_f<0x124>(v));  // row_ror:4
    v = fmaxf(v, dpp_move_f<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fmaxf(v, dpp_move_f<0xB1>(v));   // quad_perm [1,0,3,2]
    return v;
}

template <int WIDTH>
__device__ __forceinline__ double seg_max(double v) {
    static_assert(WIDTH == 8 || WIDTH == 16 || WIDTH == 32 || WIDTH == 64,
                  "segment of 8, 16, 32 or 64 lanes");
    if constexpr (WIDTH == 8) {
        v = fmax(v, dpp_move<0x141>(v));
        v = fmax(v, dpp_move<0x4E>(v));
        v = fmax(v, dpp_move<0xB1>(v));
        return v;
    }
    if constexpr (WIDTH >= 64) v = fmax_xor32(v);
    if constexpr (WIDTH >= 32) v = fmax_xor16(v);
    v = fmax(v, dpp_move<0x128>(v));
    v = fmax(v, dpp_move<0x124>(v));
    v = fmax(v, dpp_move<0x4E>(v));
    v = fmax(v, dpp_move<0xB1>(v));
    return v;
}

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

// Fixed-order block-wide sums of three doubles at once (one LDS round trip).
// `scratch` holds 3 * kWavesPerBlock doubles.  Results valid in every thread.
// reuse = false drops the closing barrier, for a last use of `scratch`.
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *scratch,
                                           bool reuse = true) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        scratch[w] = a;
        scratch[kWavesPerBlock + w] = b;
        scratch[2 * kWavesPerBlock + w] = c;
    }
    __syncthreads();
    double ta = 0.0, tb = 0.0, tc = 0.0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; i++) {
        ta += scratch[i];
        tb += scratch[kWavesPerBlock + i];
        tc += scratch[2 * kWavesPerBlock + i];
    }
    if (reuse) __syncthreads();
    a = ta;
    b = tb;
    c = tc;
}

__device__ __forceinline__ int wave_uniform(int v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// n_iter of a QR solve that never passed its stopping test: max_iters + 1
// (core/drot_qr.hpp:233), saturated at INT_MAX for max_iters == INT_MAX.
__host__ __device__ __forceinline__ int qr_unconverged_n_iter(int max_iters) {
    return max_iters < 0x7fffffff ? max_iters + 1 : 0x7fffffff;
}

struct IterStart {
    bool finish;        // the solve ended with iteration k-1: do not touch X
    bool explicit_phi;  // k == 0: phi1/phi2 arrays hold the initial phi
    float gamma;        // gamma of iteration k-1: phi = (hat + gamma) / n
};

// The first 4 kWave finalize entries of the prologue below, loaded by wave 0
// ahead of time: lane l holds entries l, l + 64, l + 128, l + 192 (0 past
// n_fin, and in the other waves).  Buffer loads with out-of-range offsets for
// the lanes with nothing to read, so that no lane condition turns them into a
// branch that waits.  Issued before a tile's first batch of X and C, the
// prologue's wait for them does not also wait for that batch (a wave's loads
// complete in order).
struct PrologueLoads {
    double v[4][3];
    float alpha_prev;  // SolverState::alpha_slot[k & 1]
};

__device__ __forceinline__ void prologue_issue(const IterCtl &C, PrologueLoads &L) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double *>(C.pfin), 0, C.k > 0 ? C.n_fin * 24 : 0, 0x00020000);
    const bool w0 = threadIdx.x < kWave;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int t = (int)threadIdx.x + u * kWave;
        const unsigned off = (w0 && t < C.n_fin) ? (unsigned)t * 24u : 0x80000000u;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            typedef unsigned int u2_t __attribute__((ext_vector_type(2)));
            const u2_t d = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off + 8u * c), 0, 0);
            L.v[u][c] = __longlong_as_double((long long)(((unsigned long long)d.y << 32) | d.x));
        }
    }
    const __amdgpu_buffer_rsrc_t ar =
        __builtin_amdgcn_make_buffer_rsrc(C.st->alpha_slot, 0, 8, 0x00020000);
    L.alpha_prev = __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(ar, (int)((C.k & 1) * 4), 0, 0));
}

// Prologue of the pass of iteration k, called by every thread of every block;
// s_red holds 3 * kWavesPerBlock doubles.  pre: the first entries, issued
// by prologue_issue (null: loaded here).
// Finishes iteration k-1 from the per-block partials of its finalize launch,
// identically in every block (same fixed-order sums):
//   S = sum X', residual = sqrt(sum rs^2 + sum cs^2), objective,
//   gamma = (2 (S - 1) - alpha) / (n_rows + n_cols)   (alpha of iteration k-2)
//   alpha <- (alpha + 1) - S                          (core/kernel_qr.hpp:421,463-467)
// and applies the stopping test of iteration k-1 (core/drot_qr.hpp:217-229 or
// core/drot_glr.hpp:344-354).  The problem's leader block records the result.  Because the test
// runs before any block writes X, a converged iterate is never overwritten.
__device__ inline IterStart iteration_prologue(const IterCtl &C, double *s_red,
                                              bool leader_block,
                                              const PrologueLoads *pre = nullptr) {
    SolverState *st = C.st;
    IterStart r;
    r.explicit_phi = true;
    r.gamma = 0.0f;
    const int k = C.k;
    const bool leader = leader_block && (threadIdx.x == 0);
    bool conv = false;
    bool finish = k >= C.max_iters;
    if (k > 0) {
        // Wave 0 alone reduces the finalize partials and forms the scalars;
        // the block's other waves take gamma and the stop decision from two
        // LDS slots behind one barrier.  Lane l adds the entries l, l + 64,
        // l + 128, ... in that order (their loads in flight together), then a
        // butterfly: the same fixed order in every block.  Slots 2W and 2W+1
        // (W = kWavesPerBlock) are ones that block_sum, the only later user of
        // s_red in a pass, never writes.
        if (threadIdx.x < kWave) {
            const float alpha_prev = pre ? pre->alpha_prev : st->alpha_slot[k & 1];  // after k-2
            double rr = 0.0, ss = 0.0, oo = 0.0;
            for (int t0 = (int)threadIdx.x; t0 < C.n_fin; t0 += 4 * kWave) {
                double v[4][3];
                if (pre && t0 == (int)threadIdx.x) {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        v[u][0] = pre->v[u][0];
                        v[u][1] = pre->v[u][1];
                        v[u][2] = pre->v[u][2];
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int t = t0 + u * kWave;
                        const bool ok = t < C.n_fin;
                        v[u][0] = ok ? C.pfin[3 * t] : 0.0;
                        v[u][1] = ok ? C.pfin[3 * t + 1] : 0.0;
                        v[u][2] = ok ? C.pfin[3 * t + 2] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    rr += v[u][0];
                    ss += v[u][1];
                    oo += v[u][2];
                }
            }
            rr = wave_sum(rr);
            ss = wave_sum(ss);
            oo = wave_sum(oo);
            const double S = ss;
            const double obj = oo;
            const double res = sqrt(rr);
            const float gamma =
                (float)((2.0 * (S - 1.0) - (double)alpha_prev) / (double)(C.n_rows + C.n_cols));
            const float alpha_new = (float)(((double)alpha_prev + 1.0) - S);
            const bool prev_even = ((k - 1) & 1) == 0;
            const bool every = C.stop_rule == kStopEveryLeq;
            const bool check = every || prev_even;
            const bool has_obj = !C.obj_even_only || prev_even;
            conv = check && (every ? (res <= (double)C.eps) : (res < (double)C.eps));
            if (leader) {
                st->alpha_slot[(k - 1) & 1] = alpha_new;
                st->alpha = alpha_new;
                st->gamma = gamma;
                st->phi_mode = 0;
                st->xsum = S;
                st->res_last = res;
                if (check) st->res = res;
                if (has_obj) st->obj = obj;
                st->last_k = k - 1;
            }
            if (threadIdx.x == 0) {
                s_red[2 * kWavesPerBlock] = (double)gamma;
                s_red[2 * kWavesPerBlock + 1] = (finish || conv) ? 1.0 : 0.0;
            }
        }
        __syncthreads();
        r.explicit_phi = false;
        r.gamma = (float)s_red[2 * kWavesPerBlock];
        finish = s_red[2 * kWavesPerBlock + 1] != 0.0;
    }
    if (finish && leader) {
        st->converged = conv ? 1 : 0;
        st->n_iter = conv ? k
                          : (C.stop_rule == kStopEveryLeq ? C.max_iters
                                                          : qr_unconverged_n_iter(C.max_iters));
        st->k_run = k;
        st->done = 1;
    }
    r.finish = finish;
    return r;
}

}  // namespace drot
