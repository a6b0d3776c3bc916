// gfx950 kernels of the DR regularized-OT solver.
//
// One DR iteration = one fused pass over X (and C) + one small finalize launch:
//   qr_pass / gl_pass : X-update (prox), and in the same pass the contiguous-axis
//                       sums (wave butterfly over the lanes of one column), the
//                       strided-axis sums (per-lane fp64 accumulators combined
//                       through LDS into one partial per tile) and the objective.
//   finalize          : fixed-order reduction of the tile partials, the dual /
//                       aux update of core/kernel_qr.hpp:405-467, the residual
//                       and the stopping test -- all on device, no host round trip.
// The pass is HBM-bound elementwise work (about 1 flop per byte), so it uses
// plain VALU, 128-bit coalesced accesses and no matrix cores.
#include "drot_finalize.h"
#include "drot_kernels.h"

#include <math.h>

namespace drot {

// ----------------------------------------------------------------- helpers

constexpr int kGlPolicySc1 = 16;  // cache-policy operand of a buffer store: sc1

// one fp64 value through a buffer resource, device-scope write-through (sc1)
__device__ __forceinline__ void store_f64_wt(__amdgpu_buffer_rsrc_t r, unsigned off, double v) {
    typedef unsigned int u2_t __attribute__((ext_vector_type(2)));
    const unsigned long long a = (unsigned long long)__double_as_longlong(v);
    const u2_t d = {(unsigned)a, (unsigned)(a >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(d, r, (int)off, 0, kGlPolicySc1);
}

template <int VEC>
struct VecIO;

template <>
struct VecIO<4> {
    // raw buffer load: an offset past the resource's range reads 0
    static __device__ __forceinline__ void load_buf(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                    float (&v)[4]) {
        typedef unsigned int u4_t __attribute__((ext_vector_type(4)));
        const u4_t d = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
        v[0] = __uint_as_float(d.x); v[1] = __uint_as_float(d.y);
        v[2] = __uint_as_float(d.z); v[3] = __uint_as_float(d.w);
    }
    static __device__ __forceinline__ void load(const float *p, float (&v)[4]) {
        const float4 t = *reinterpret_cast<const float4 *>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    }
    static __device__ __forceinline__ void store(float *p, const float (&v)[4]) {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    }
    // raw buffer store, device-scope write-through (sc1): the stored columns
    // stay out of the L2 dirty set the kernel's end writes back
    static __device__ __forceinline__ void store_wt(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                    const float (&v)[4]) {
        typedef unsigned int u4_t __attribute__((ext_vector_type(4)));
        const u4_t d = {__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]),
                        __float_as_uint(v[3])};
        __builtin_amdgcn_raw_buffer_store_b128(d, r, (int)off, 0, kGlPolicySc1);
    }
};

template <>
struct VecIO<1> {
    static __device__ __forceinline__ void load_buf(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                    float (&v)[1]) {
        v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    }
    static __device__ __forceinline__ void load(const float *p, float (&v)[1]) { v[0] = *p; }
    static __device__ __forceinline__ void store(float *p, const float (&v)[1]) { *p = v[0]; }
    static __device__ __forceinline__ void store_wt(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                    const float (&v)[1]) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[0]), r, (int)off, 0, kGlPolicySc1);
    }
};

// raw buffer load of VEC ints: an offset past the resource's range reads 0
template <int VEC>
__device__ __forceinline__ void load_ints_buf(__amdgpu_buffer_rsrc_t r, unsigned off,
                                              int (&v)[VEC]) {
    if constexpr (VEC == 4) {
        typedef unsigned int u4_t __attribute__((ext_vector_type(4)));
        const u4_t d = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
        v[0] = (int)d.x; v[1] = (int)d.y; v[2] = (int)d.z; v[3] = (int)d.w;
    } else {
#pragma unroll
        for (int e = 0; e < VEC; e++)
            v[e] = (int)__builtin_amdgcn_raw_buffer_load_b32(r, (int)(off + 4u * e), 0, 0);
    }
}

template <int VEC>
__device__ __forceinline__ void load_ints(const int *p, int (&v)[VEC]) {
    if constexpr (VEC == 4) {
        const int4 t = *reinterpret_cast<const int4 *>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
#pragma unroll
        for (int e = 0; e < VEC; e++) v[e] = p[e];
    }
}

// phi of one row / column: explicit array on the first iteration after an
// initialisation, (hat + gamma) / n afterwards.  The two fp32 roundings are
// those of `(ab_val - rs_val + gamma) / T(n)` in core/kernel_qr.hpp:431,448
// (the `ab_val - rs_val` part is stored in hat by the finalize kernel).
__device__ __forceinline__ float phi_of(int mode, const float *phi, const float *hat,
                                        int idx, float gamma, float n) {
    return mode ? phi[idx] : (hat[idx] + gamma) / n;
}

// ----------------------------------------------------------------- init

// Warm start (pytorch/src/qr_drot_cuda_kernel.cu:45-71: X = 0 and constant
// phi/a/b/alpha computed on the host) or cold start (core/drot_qr.hpp:308-320:
// X = p q^T, phi = a = b = 0).  Fill kernels only, no memset nodes.
__global__ __launch_bounds__(kBlock) void init_kernel(InitArgs A) {
    const size_t total = (size_t)A.n_rows * (size_t)A.n_cols;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nthr = (size_t)gridDim.x * blockDim.x;
    if (A.warm) {
        for (size_t e = tid; e < total; e += nthr) A.x[e] = 0.0f;
    } else {
        for (size_t e = tid; e < total; e += nthr) {
            const int j = (int)(e / (size_t)A.n_rows);
            const int i = (int)(e - (size_t)j * (size_t)A.n_rows);
            A.x[e] = A.p[i] * A.q[j];
        }
    }
    for (size_t i = tid; i < (size_t)A.n_rows; i += nthr) {
        A.phi1[i] = A.warm ? A.v_phi1 : 0.0f;
        A.a[i] = A.warm ? A.v_a : 0.0f;
        A.ahat[i] = 0.0f;
    }
    for (size_t j = tid; j < (size_t)A.n_cols; j += nthr) {
        A.phi2[j] = A.warm ? A.v_phi2 : 0.0f;
        A.b[j] = A.warm ? A.v_b : 0.0f;
        A.bhat[j] = 0.0f;
    }
    // no segment starts out flagged: the first pass reads all of X
    if (A.zf)
        for (size_t i = tid; i < A.zf_bytes; i += nthr) A.zf[i] = 0;
    // no headroom known yet: nothing is screened before its C has been read
    if (A.seg_e) {
        for (size_t i = tid; i < A.seg_e_count; i += nthr) A.seg_e[i] = INFINITY;
        for (size_t i = tid; i < A.g1_count; i += nthr) A.g1[i] = 0.0;
    }
    if (tid == 0) {
        SolverState s;
        s.done = 0;
        s.converged = 0;
        s.n_iter = 0;
        s.k_run = 0;
        s.gamma = 0.0f;
        s.alpha = A.warm ? A.v_alpha : 0.0f;
        s.alpha_slot[0] = 0.0f;
        s.alpha_slot[1] = s.alpha;  // "after iteration -1"
        s.phi_mode = 1;
        s.last_k = -1;
        s.obj = 0.0;
        s.res = 1e9;
        s.xsum = 0.0;
        s.res_last = 1e9;
        *A.st = s;
    }

}

// ----------------------------------------------------------------- GL pass

// Soft-threshold scale of one group column, core/kernel_glr.hpp:125,320:
// max(1 - lambda / (||x||_2 + 1e-9), 0), from the fp64 sum of squares.
__device__ __forceinline__ float group_scale(double n2, double lam) {
    const double sc = 1.0 - lam / (sqrt(n2) + 1e-9);
    return (float)(sc > 0.0 ? sc : 0.0);
}

constexpr int kNormBatch = 8;  // LDS reads in flight per lane in the group norms

// Group row bounds are read from LDS when there are at most kGlBoundsLds
// groups (copied at kernel start).
constexpr int kGlBoundsLds = 1024;


// Fused GL-DROT pass for one tile of `tc` whole columns.  Restates
// group_lasso_regularizer_update_x_for_small (core/kernel_glr.hpp:59-161) and
// the large-group chain update_x / update_norm / apply_norm
// (core/kernel_glr.hpp:163-399) in ONE pass: the block holds the whole column
// in registers, so the group norms are formed before X is written and the
// reference's second apply_norm pass over X disappears.
//   X' = max(X + phi1 + phi2 - s C, 0)
//   scale_{g,j} = max(1 - lambda / (||X'_{g,j}||_2 + 1e-9), 0)
//   X_{g,j} := X'_{g,j} * scale_{g,j}
// Thread layout: row chunk r of thread t covers rows (r*kBlock + t)*VEC ...+VEC-1.
template <int VEC, int RCH>
__global__ __launch_bounds__(kBlock) void gl_pass_kernel(GlPassArgs A) {
    extern __shared__ double s_dyn[];
    __shared__ double s_red[3 * kWavesPerBlock];
    __shared__ int s_bnd[kGlBoundsLds + 1];
    __shared__ int s_done;
    const unsigned long long t0 = A.ctl.tspan ? (unsigned long long)wall_clock64() : 0ull;
    // Loads before the prologue, in this order and as raw buffer loads with an
    // out-of-range offset for a lane with nothing to read (see gl_rows_kernel):
    // the prologue's finalize entries, the group bounds, the rows' potentials
    // and groups.
    PrologueLoads pre;
    prologue_issue(A.ctl, pre);
    const int done_v = A.ctl.st->done;  // uniform: a scalar load
    const int G0 = A.n_groups;
    const bool bnd_lds0 = G0 <= kGlBoundsLds;
    const __amdgpu_buffer_rsrc_t br =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(A.bounds), 0, (G0 + 1) * 4, 0x00020000);
    const int bnd_v = (int)__builtin_amdgcn_raw_buffer_load_b32(
        br, (int)(bnd_lds0 && (int)threadIdx.x <= G0 ? threadIdx.x * 4u : 0x80000000u), 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    const bool explicit_phi = A.ctl.k == 0;
    const float *hsrc = explicit_phi ? A.phi1 : A.ahat;
    const __amdgpu_buffer_rsrc_t h1r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(hsrc), 0, A.n_rows * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t gr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int *>(A.row_group), 0, A.n_rows * 4, 0x00020000);
    float h1[RCH][VEC];
    int gid[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
        const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
        const unsigned off = i0 < A.n_rows ? (unsigned)i0 * 4u : 0x80000000u;
        VecIO<VEC>::load_buf(h1r, off, h1[r]);
        load_ints_buf<VEC>(gr, off, gid[r]);
    }
    if (threadIdx.x == 0) s_done = done_v;
    if (bnd_lds0 && (int)threadIdx.x <= G0) s_bnd[threadIdx.x] = bnd_v;
    if (bnd_lds0 && G0 >= kBlock) {  // more bounds than threads (rare): the rest directly
        for (int g = threadIdx.x + kBlock; g <= G0; g += kBlock) s_bnd[g] = A.bounds[g];
    }
    __syncthreads();
    if (s_done) return;  // the solve ended before this launch
    const IterStart is = iteration_prologue(A.ctl, s_red, blockIdx.x == 0, &pre);
    if (is.finish) return;
    const int phi_mode = is.explicit_phi ? 1 : 0;
    const float gamma = is.gamma;
    const int N = A.n_rows, M = A.n_cols, G = A.n_groups;
    const bool bnd_lds = G <= kGlBoundsLds;
    auto bnd = [&](int g) { return bnd_lds ? s_bnd[g] : A.bounds[g]; };
    double *s_sq = s_dyn;                                   // [N]
    float *s_scale = reinterpret_cast<float *>(s_dyn + N);  // [G]
    const float fM = (float)M, fN = (float)N;
    const float s = A.step;
    const double lam = (double)A.lambda;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_uniform(threadIdx.x / kWave);

    float f1[RCH][VEC];
    double racc[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            racc[r][e] = 0.0;
            f1[r][e] = phi_mode ? h1[r][e] : (h1[r][e] + gamma) / fM;
        }
    }
    double oacc = 0.0;

    const int j0 = blockIdx.x * A.tc;
    const int j1 = min(j0 + A.tc, M);
    for (int j = j0; j < j1; j++) {
        const float f2 = phi_of(phi_mode, A.phi2, A.bhat, j, gamma, fN);
        const float *xc = A.x + (size_t)j * N;
        const float *cc = A.c + (size_t)j * N;
        float xv[RCH][VEC], cv[RCH][VEC];
#pragma unroll
        for (int r = 0; r < RCH; r++) {
            const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
            if (i0 < N) {
                VecIO<VEC>::load(xc + i0, xv[r]);
                VecIO<VEC>::load(cc + i0, cv[r]);
            } else {
#pragma unroll
                for (int e = 0; e < VEC; e++) { xv[r][e] = 0.0f; cv[r][e] = 0.0f; }
            }
        }
#pragma unroll
        for (int r = 0; r < RCH; r++) {
            const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float tt = (xv[r][e] + f1[r][e]) + f2;
                const float y = fmaxf(__builtin_fmaf(-s, cv[r][e], tt), 0.0f);
                xv[r][e] = y;
                if (i0 + e < N) s_sq[i0 + e] = (double)y * (double)y;
            }
        }
        __syncthreads();
        // group norms -> scales, fixed order
        // (LDS reads are issued kNormBatch at a time; the additions keep their
        // sequential order)
        if (A.lane_per_group) {
            for (int g = threadIdx.x; g < G; g += kBlock) {
                const int hi = bnd(g + 1);
                int i = bnd(g);
                double n2 = 0.0;
                for (; i + kNormBatch <= hi; i += kNormBatch) {
                    double v[kNormBatch];
#pragma unroll
                    for (int u = 0; u < kNormBatch; u++) v[u] = s_sq[i + u];
#pragma unroll
                    for (int u = 0; u < kNormBatch; u++) n2 += v[u];
                }
                for (; i < hi; i++) n2 += s_sq[i];
                s_scale[g] = group_scale(n2, lam);
            }
        } else {
            for (int g = w; g < G; g += kWavesPerBlock) {
                const int hi = bnd(g + 1);
                int i = bnd(g) + lane;
                double n2 = 0.0;
                for (; i + (kNormBatch - 1) * kWave < hi; i += kNormBatch * kWave) {
                    double v[kNormBatch];
#pragma unroll
                    for (int u = 0; u < kNormBatch; u++) v[u] = s_sq[i + u * kWave];
#pragma unroll
                    for (int u = 0; u < kNormBatch; u++) n2 += v[u];
                }
                for (; i < hi; i += kWave) n2 += s_sq[i];
                n2 = wave_sum(n2);
                if (lane == 0) s_scale[g] = group_scale(n2, lam);
            }
        }
        __syncthreads();
        double csum = 0.0;
#pragma unroll
        for (int r = 0; r < RCH; r++) {
            const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
            if (i0 < N) {
                float xo[VEC];
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    const float v = xv[r][e] * s_scale[gid[r][e]];
                    xo[e] = v;
                    racc[r][e] += (double)v;
                    csum += (double)v;
                    oacc += (double)v * (double)cv[r][e];
                }
                VecIO<VEC>::store(A.x + (size_t)j * N + i0, xo);
            }
        }
        csum = block_sum(csum, s_red);  // also fences s_sq / s_scale reuse
        if (threadIdx.x == 0) A.pcol[j] = csum;
    }
#pragma unroll
    for (int r = 0; r < RCH; r++) {
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            const int i = (r * kBlock + (int)threadIdx.x) * VEC + e;
            if (i < N) A.prow[(size_t)blockIdx.x * N + i] = racc[r][e];
        }
    }
    const double o = block_sum(oacc, s_red);
    if (threadIdx.x == 0) A.pobj[blockIdx.x] = o;
    if (A.ctl.tspan) record_span(A.ctl.tspan, t0);
}

// Group scales of one column from its prox values y[0 .. n_rows) in LDS, by
// one wave, written to sc[0 .. G): one lane per group summing its rows in
// order (groups under 64 rows), or the whole wave per group, lane l adding
// rows lo + l, lo + l + 64, ... before a butterfly.  Fixed order either way.
// (LDS reads are issued kNormBatch at a time; the additions keep their
// sequential order.)
template <typename Bnd>
__device__ __forceinline__ void column_group_scales(const float *y, float *sc, int G,
                                                    bool lane_per_group, const Bnd &bnd,
                                                    double lam, int lane) {
    if (lane_per_group) {
        for (int g = lane; g < G; g += kWave) {
            const int hi = bnd(g + 1);
            const int lo = bnd(g);
            int i = lo;
            double n2 = 0.0;
            for (; i + kNormBatch <= hi; i += kNormBatch) {
                float v[kNormBatch];
#pragma unroll
                for (int u = 0; u < kNormBatch; u++) v[u] = y[i + u];
#pragma unroll
                for (int u = 0; u < kNormBatch; u++) n2 += (double)v[u] * (double)v[u];
            }
            // the last rows (fewer than kNormBatch) read at once, clamped to
            // the group's first row; a row past the group adds +0.0, which
            // leaves the sum's bits unchanged
            if (i < hi) {
                float v[kNormBatch - 1];
#pragma unroll
                for (int u = 0; u < kNormBatch - 1; u++) v[u] = y[i + u < hi ? i + u : lo];
#pragma unroll
                for (int u = 0; u < kNormBatch - 1; u++) {
                    const double d = (double)v[u];
                    n2 += i + u < hi ? d * d : 0.0;
                }
            }
            sc[g] = group_scale(n2, lam);
        }
    } else {
        // lane g % 64 keeps group g's total; the scales of up to 64 groups
        // are formed at once
        double mine = 0.0;
        for (int g = 0; g < G; g++) {
            const int hi = bnd(g + 1);
            const int lo = bnd(g);
            int i = lo + lane;
            double n2 = 0.0;
            for (; i + (kNormBatch - 1) * kWave < hi; i += kNormBatch * kWave) {
                float v[kNormBatch];
#pragma unroll
                for (int u = 0; u < kNormBatch; u++) v[u] = y[i + u * kWave];
#pragma unroll
                for (int u = 0; u < kNormBatch; u++) n2 += (double)v[u] * (double)v[u];
            }
            if (i < hi) {  // the lane's last rows (fewer than kNormBatch) read at once, as above
                float v[kNormBatch - 1];
#pragma unroll
                for (int u = 0; u < kNormBatch - 1; u++) {
                    const int ii = i + u * kWave;
                    v[u] = y[ii < hi ? ii : lo];
                }
#pragma unroll
                for (int u = 0; u < kNormBatch - 1; u++) {
                    const double d = (double)v[u];
                    n2 += i + u * kWave < hi ? d * d : 0.0;
                }
            }
            n2 = wave_sum(n2);
            if ((g & (kWave - 1)) == lane) mine = n2;
            if ((g & (kWave - 1)) == kWave - 1 || g == G - 1) {
                const int gb = g & ~(kWave - 1);
                if (gb + lane <= g) sc[gb + lane] = group_scale(mine, lam);
            }
        }
    }
}

// GL pass with one wave per column (n_rows <= 2048): the 4 waves of a block
// take different columns at once, so a block never waits for a column-wide
// barrier.  A wave keeps its column's prox values in registers and a copy in
// its own LDS slice, forms the group norms from that slice (one wave per group
// when groups have >= 64 rows, one lane per group otherwise, fp64 in a fixed
// order), and reads back the group scales by row.  Column sums are wave sums;
// the row partials of the block's columns are combined across the waves
// through LDS at the end, in wave order.  Same arithmetic as gl_pass_kernel.

template <int VEC, int RCH>
__global__ __launch_bounds__(kBlock) void gl_wave_kernel(GlPassArgs A) {
    constexpr int kChunk = kWave * VEC;  // rows per register chunk
    extern __shared__ float s_wave[];    // per wave: y[n_rows], scale[n_groups]
    __shared__ double s_red[3 * kWavesPerBlock];
    __shared__ double s_comb[2][kWavesPerBlock][kChunk];
    __shared__ int s_bnd[kGlBoundsLds + 1];
    __shared__ int s_done;
    const unsigned long long t0 = A.ctl.tspan ? (unsigned long long)wall_clock64() : 0ull;
    // Loads before the prologue in this order, all raw buffer loads with an
    // out-of-range offset for a lane with nothing to read (see gl_rows_kernel):
    // the prologue's finalize entries, the group bounds, the rows' potentials
    // and groups, the first column's X and C.
    PrologueLoads pre;
    prologue_issue(A.ctl, pre);
    const int done_v = A.ctl.st->done;  // uniform: a scalar load
    const int G0 = A.n_groups;
    const bool bnd_lds0 = G0 <= kGlBoundsLds;
    const __amdgpu_buffer_rsrc_t br =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(A.bounds), 0, (G0 + 1) * 4, 0x00020000);
    const int bnd_v = (int)__builtin_amdgcn_raw_buffer_load_b32(
        br, (int)(bnd_lds0 && (int)threadIdx.x <= G0 ? threadIdx.x * 4u : 0x80000000u), 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    const int lane0 = threadIdx.x & (kWave - 1);
    const bool explicit_phi = A.ctl.k == 0;
    const float *hsrc = explicit_phi ? A.phi1 : A.ahat;
    const __amdgpu_buffer_rsrc_t h1r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(hsrc), 0, A.n_rows * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t gr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int *>(A.row_group), 0, A.n_rows * 4, 0x00020000);
    float h1[RCH][VEC];
    int gid[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
        const int i0 = (r * kWave + lane0) * VEC;
        const unsigned off = i0 < A.n_rows ? (unsigned)i0 * 4u : 0x80000000u;
        VecIO<VEC>::load_buf(h1r, off, h1[r]);
        load_ints_buf<VEC>(gr, off, gid[r]);
    }
    // this block's columns: [j0, j1); wave w takes j0 + w, j0 + w + 4, ...
    // The first column's X and C are loaded before the prologue as well.
    const int w = wave_uniform(threadIdx.x / kWave);
    const int j0 = blockIdx.x * A.tc;
    const int j1 = min(j0 + A.tc, A.n_cols);
    float xv[RCH][VEC], cv[RCH][VEC];
    float h2 = 0.0f;
    const unsigned blk_bytes = (unsigned)(j1 - j0) * (unsigned)A.n_rows * 4u;
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
        A.x + (size_t)j0 * A.n_rows, 0, (int)blk_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t cr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(A.c) + (size_t)j0 * A.n_rows, 0, (int)blk_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t h2r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(explicit_phi ? A.phi2 : A.bhat) + j0, 0, (j1 - j0) * 4, 0x00020000);
    auto load_col = [&](int j) {  // j < j1
        const unsigned jc = (unsigned)(j - j0);
        h2 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(h2r, (int)(jc * 4u), 0, 0));
#pragma unroll
        for (int r = 0; r < RCH; r++) {
            const int i0 = (r * kWave + lane0) * VEC;
            const unsigned off =
                i0 < A.n_rows ? (jc * (unsigned)A.n_rows + (unsigned)i0) * 4u : 0x80000000u;
            VecIO<VEC>::load_buf(xr, off, xv[r]);
            VecIO<VEC>::load_buf(cr, off, cv[r]);
        }
    };
    if (j0 + w < j1) load_col(j0 + w);
    if (threadIdx.x == 0) s_done = done_v;
    if (bnd_lds0 && (int)threadIdx.x <= G0) s_bnd[threadIdx.x] = bnd_v;
    if (bnd_lds0 && G0 >= kBlock) {  // more bounds than threads (rare): the rest directly
        for (int g = threadIdx.x + kBlock; g <= G0; g += kBlock) s_bnd[g] = A.bounds[g];
    }
    __syncthreads();
    if (s_done) return;  // the solve ended before this launch
    const IterStart is = iteration_prologue(A.ctl, s_red, blockIdx.x == 0, &pre);
    if (is.finish) return;
    const int phi_mode = is.explicit_phi ? 1 : 0;
    const float gamma = is.gamma;
    const int N = A.n_rows, M = A.n_cols, G = A.n_groups;
    const bool bnd_lds = G <= kGlBoundsLds;
    auto bnd = [&](int g) { return bnd_lds ? s_bnd[g] : A.bounds[g]; };
    const float fM = (float)M, fN = (float)N;
    const float s = A.step;
    const double lam = (double)A.lambda;
    const int lane = lane0;
    float *s_y = s_wave + (size_t)w * (N + G);
    float *s_sc = s_y + N;

    float f1[RCH][VEC];
    double racc[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            racc[r][e] = 0.0;
            f1[r][e] = phi_mode ? h1[r][e] : (h1[r][e] + gamma) / fM;
        }
    }
    double oacc = 0.0;
    for (int j = j0 + w; j < j1; j += kWavesPerBlock) {
        if (j != j0 + w) load_col(j);
        const float f2 = phi_mode ? h2 : (h2 + gamma) / fN;
#pragma unroll
        for (int r = 0; r < RCH; r++) {
            const int i0 = (r * kWave + lane) * VEC;
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float tt = (xv[r][e] + f1[r][e]) + f2;
                const float y = fmaxf(__builtin_fmaf(-s, cv[r][e], tt), 0.0f);
                xv[r][e] = y;
                if (i0 + e < N) s_y[i0 + e] = y;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        column_group_scales(s_y, s_sc, G, A.lane_per_group != 0, bnd, lam, lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // every scale read goes out at once (a row past n_rows has group 0)
        float scv[RCH][VEC];
#pragma unroll
        for (int r = 0; r < RCH; r++) {
#pragma unroll
            for (int e = 0; e < VEC; e++) scv[r][e] = s_sc[gid[r][e]];
        }
        double csum = 0.0;
#pragma unroll
        for (int r = 0; r < RCH; r++) {
            const int i0 = (r * kWave + lane) * VEC;
            if (i0 < N) {
                float xo[VEC];
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    const float v = xv[r][e] * scv[r][e];
                    xo[e] = v;
                    racc[r][e] += (double)v;
                    csum += (double)v;
                    oacc += (double)v * (double)cv[r][e];
                }
                VecIO<VEC>::store_wt(xr, ((unsigned)(j - j0) * (unsigned)N + (unsigned)i0) * 4u, xo);
            }
        }
        csum = wave_sum(csum);
        if (lane == 0) A.pcol[j] = csum;
        // the next column rewrites s_y only after every lane has read the scales
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // row partials of the block: waves combined in a fixed order, chunk by
    // chunk through two alternating LDS buffers (one barrier per chunk; the
    // barrier of chunk r + 1 orders the reads of chunk r before the writes of
    // chunk r + 2 into the same buffer)
    const __amdgpu_buffer_rsrc_t pr = __builtin_amdgcn_make_buffer_rsrc(
        A.prow + (size_t)blockIdx.x * N, 0, N * 8, 0x00020000);
#pragma unroll
    for (int r = 0; r < RCH; r++) {
        if (r * kChunk >= N) break;
        double (*cb)[kChunk] = s_comb[r & 1];
#pragma unroll
        for (int e = 0; e < VEC; e++) cb[w][lane * VEC + e] = racc[r][e];
        __syncthreads();
        for (int t = threadIdx.x; t < kChunk; t += kBlock) {
            const int i = r * kChunk + t;
            if (i < N) {
                double sum = 0.0;
#pragma unroll
                for (int q = 0; q < kWavesPerBlock; q++) sum += cb[q][t];
                store_f64_wt(pr, (unsigned)i * 8u, sum);  // written through, as X
            }
        }
    }
    const double o = block_sum(oacc, s_red);
    if (threadIdx.x == 0) A.pobj[blockIdx.x] = o;
    if (A.ctl.tspan) record_span(A.ctl.tspan, t0);
}

// GL pass with the block's threads over rows and its columns taken four at a
// time (n_rows <= kBlock*VEC*RCH, at most 2048).  Thread t owns the rows
// (r*kBlock + t)*VEC + e of every column of the block, so its row partials
// add up in registers over the columns in order and are stored as they are:
// no combine across waves.  Per round of four columns:
//   * every thread forms y = max(X + phi1 + phi2 - sC, 0) of its rows of the
//     four columns and writes them to the columns' LDS slices (one barrier);
//   * wave w forms the group scales of column w from its slice in the fixed
//     order of column_group_scales (one barrier);
//   * every thread scales and stores its elements; the column sums and the
//     objective are wave sums added in wave order (one barrier).
// All loads of the first round (X, C, both potentials, the groups of the
// rows, the done flag) are issued before the bounds staging and the
// prologue's reductions.  Same arithmetic as gl_pass_kernel.
template <int VEC, int RCH>
__global__ __launch_bounds__(kBlock) void gl_rows_kernel(GlPassArgs A) {
    constexpr int kCols = kWavesPerBlock;  // columns per round, one per wave for the norms
    // per column of a round: y[n_rows], scale[n_groups] (gl_rows_slice floats)
    extern __shared__ __attribute__((aligned(16))) float s_cols[];
    __shared__ double s_red[3 * kWavesPerBlock];
    __shared__ double s_cs[kWavesPerBlock][kCols + 1];
    __shared__ int s_bnd[kGlBoundsLds + 1];
    __shared__ int s_done;
    const unsigned long long t0 = A.ctl.tspan ? (unsigned long long)wall_clock64() : 0ull;
    const int N = A.n_rows, M = A.n_cols, G = A.n_groups;
    const int slice = gl_rows_slice(N, G);  // floats per column slice (16-byte multiple)
    const bool explicit_phi = A.ctl.k == 0;
    const float *hsrc = explicit_phi ? A.phi1 : A.ahat;
    const float *hat2 = explicit_phi ? A.phi2 : A.bhat;
    const int j0 = blockIdx.x * A.tc;
    const int j1 = min(j0 + A.tc, M);
    // Every load before the prologue is a buffer load whose offset is pushed
    // out of range for a lane with nothing to read (the hardware returns 0),
    // in this order: the prologue's finalize entries, the group bounds, the
    // rows' potentials and groups, the first round of columns.  A load under
    // a lane condition becomes a branch that waits for it and for everything
    // issued before it (a wave's loads complete in order), which serialised
    // the pass's start; in this order each wait covers only what it needs.
    PrologueLoads pre;
    prologue_issue(A.ctl, pre);
    const int done_v = A.ctl.st->done;  // uniform: a scalar load
    const bool bnd_lds = G <= kGlBoundsLds;
    const __amdgpu_buffer_rsrc_t br =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(A.bounds), 0, (G + 1) * 4, 0x00020000);
    const int bnd_v = (int)__builtin_amdgcn_raw_buffer_load_b32(
        br, (int)(bnd_lds && (int)threadIdx.x <= G ? threadIdx.x * 4u : 0x80000000u), 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    const __amdgpu_buffer_rsrc_t h1r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(hsrc), 0, N * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t gr =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(A.row_group), 0, N * 4, 0x00020000);
    float h1[RCH][VEC];
    int gid[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
        const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
        const unsigned off = i0 < N ? (unsigned)i0 * 4u : 0x80000000u;
        VecIO<VEC>::load_buf(h1r, off, h1[r]);
        load_ints_buf<VEC>(gr, off, gid[r]);
    }
    // one register set: a second one, to keep the next round's loads in
    // flight, took the register count from 168 to 246 and made the pass
    // slower at every shape measured
    float xv[kCols][RCH][VEC], cv[kCols][RCH][VEC], h2[kCols];
    // the block's columns [j0, j1) of X and C, and of the column potentials
    const unsigned blk_bytes = (unsigned)(j1 - j0) * (unsigned)N * 4u;
    const __amdgpu_buffer_rsrc_t xr =
        __builtin_amdgcn_make_buffer_rsrc(A.x + (size_t)j0 * N, 0, (int)blk_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t cr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(A.c) + (size_t)j0 * N, 0, (int)blk_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t h2r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(hat2) + j0, 0, (j1 - j0) * 4, 0x00020000);
    auto load_round = [&](int jr) __attribute__((always_inline)) {
#pragma unroll
        for (int jj = 0; jj < kCols; jj++) {
            const int jc = jr + jj - j0;  // column within the block
            const bool col_ok = jr + jj < j1;
            h2[jj] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                h2r, (int)(col_ok ? (unsigned)jc * 4u : 0x80000000u), 0, 0));
#pragma unroll
            for (int r = 0; r < RCH; r++) {
                const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
                const unsigned off =
                    (col_ok && i0 < N) ? ((unsigned)jc * (unsigned)N + (unsigned)i0) * 4u : 0x80000000u;
                VecIO<VEC>::load_buf(xr, off, xv[jj][r]);
                VecIO<VEC>::load_buf(cr, off, cv[jj][r]);
            }
        }
    };
    load_round(j0);
    if (threadIdx.x == 0) s_done = done_v;
    if (bnd_lds && (int)threadIdx.x <= G) s_bnd[threadIdx.x] = bnd_v;
    if (bnd_lds && G >= kBlock) {  // more bounds than threads (rare): the rest directly
        for (int g = threadIdx.x + kBlock; g <= G; g += kBlock) s_bnd[g] = A.bounds[g];
    }
    __syncthreads();
    if (s_done) return;  // the solve ended before this launch
    const IterStart is = iteration_prologue(A.ctl, s_red, blockIdx.x == 0, &pre);
    if (is.finish) return;
    const float gamma = is.gamma;
    auto bnd = [&](int g) { return bnd_lds ? s_bnd[g] : A.bounds[g]; };
    const float fM = (float)M, fN = (float)N;
    const float s = A.step;
    const double lam = (double)A.lambda;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_uniform(threadIdx.x / kWave);

    float f1[RCH][VEC];
    double racc[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            racc[r][e] = 0.0;
            f1[r][e] = is.explicit_phi ? h1[r][e] : (h1[r][e] + gamma) / fM;
        }
    }
    double oacc = 0.0;
    for (int jr = j0; jr < j1; jr += kCols) {
        if (jr != j0) load_round(jr);
        // prox values of the round's columns -> their LDS slices
#pragma unroll
        for (int jj = 0; jj < kCols; jj++) {
            const float f2 = is.explicit_phi ? h2[jj] : (h2[jj] + gamma) / fN;
            float *ys = s_cols + jj * slice;
#pragma unroll
            for (int r = 0; r < RCH; r++) {
                const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
                float yv[VEC];
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    const float tt = (xv[jj][r][e] + f1[r][e]) + f2;
                    yv[e] = fmaxf(__builtin_fmaf(-s, cv[jj][r][e], tt), 0.0f);
                    xv[jj][r][e] = yv[e];
                }
                if (i0 < N) VecIO<VEC>::store(ys + i0, yv);
            }
        }
        __syncthreads();
        if (jr + w < j1) {
            float *ys = s_cols + w * slice;
            column_group_scales(ys, ys + N, G, A.lane_per_group != 0, bnd, lam, lane);
        }
        __syncthreads();
        // every scale read of the round goes out at once (a row past n_rows
        // has group 0; the slices of columns past j1 exist and are not used)
        float scv[kCols][RCH][VEC];
#pragma unroll
        for (int jj = 0; jj < kCols; jj++) {
            const float *sc = s_cols + jj * slice + N;
#pragma unroll
            for (int r = 0; r < RCH; r++) {
#pragma unroll
                for (int e = 0; e < VEC; e++) scv[jj][r][e] = sc[gid[r][e]];
            }
        }
        double cs[kCols];
#pragma unroll
        for (int jj = 0; jj < kCols; jj++) {
            cs[jj] = 0.0;
            if (jr + jj >= j1) continue;  // block-uniform
            const unsigned col_off = (unsigned)(jr + jj - j0) * (unsigned)N;
#pragma unroll
            for (int r = 0; r < RCH; r++) {
                const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
                if (i0 < N) {
                    float xo[VEC];
#pragma unroll
                    for (int e = 0; e < VEC; e++) {
                        const float v = xv[jj][r][e] * scv[jj][r][e];
                        xo[e] = v;
                        racc[r][e] += (double)v;
                        cs[jj] += (double)v;
                        oacc += (double)v * (double)cv[jj][r][e];
                    }
                    VecIO<VEC>::store_wt(xr, (col_off + (unsigned)i0) * 4u, xo);
                }
            }
        }
        // column sums: wave sums, then the waves in order
#pragma unroll
        for (int jj = 0; jj < kCols; jj++) cs[jj] = wave_sum(cs[jj]);
        if (lane == 0) {
#pragma unroll
            for (int jj = 0; jj < kCols; jj++) s_cs[w][jj] = cs[jj];
        }
        __syncthreads();
        if ((int)threadIdx.x < kCols && jr + (int)threadIdx.x < j1) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < kWavesPerBlock; q++) t += s_cs[q][threadIdx.x];
            A.pcol[jr + threadIdx.x] = t;
        }
        // the next round writes s_cs only after two more barriers, and its
        // LDS slices only after every thread has passed this one
    }
    // the block's row partials, written through like X (rows past n_rows
    // take an out-of-range offset and are dropped)
    const __amdgpu_buffer_rsrc_t pr = __builtin_amdgcn_make_buffer_rsrc(
        A.prow + (size_t)blockIdx.x * N, 0, N * 8, 0x00020000);
#pragma unroll
    for (int r = 0; r < RCH; r++) {
        const int i0 = (r * kBlock + (int)threadIdx.x) * VEC;
        const unsigned off = i0 < N ? (unsigned)i0 * 8u : 0x80000000u;
        if constexpr (VEC == 4) {
            typedef unsigned int u4_t __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const unsigned long long a = (unsigned long long)__double_as_longlong(racc[r][2 * h]);
                const unsigned long long b = (unsigned long long)__double_as_longlong(racc[r][2 * h + 1]);
                const u4_t d = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
                __builtin_amdgcn_raw_buffer_store_b128(d, pr, (int)(off + 16u * h), 0, kGlPolicySc1);
            }
        } else {
            typedef unsigned int u2_t __attribute__((ext_vector_type(2)));
            const unsigned long long a = (unsigned long long)__double_as_longlong(racc[r][0]);
            const u2_t d = {(unsigned)a, (unsigned)(a >> 32)};
            __builtin_amdgcn_raw_buffer_store_b64(d, pr, (int)off, 0, kGlPolicySc1);
        }
    }
    const double o = block_sum(oacc, s_red);
    if (threadIdx.x == 0) A.pobj[blockIdx.x] = o;
    if (A.ctl.tspan) record_span(A.ctl.tspan, t0);
}

// ------------------------------------------------- GL pass for long columns
//
// Columns too long for one block (the fused kernels above keep a whole column
// in registers and LDS) are cut into row chunks of R = kBlock*VEC*RCH rows,
// block (rc, ct) taking chunk rc of the columns of tile ct, and the group
// norms are completed across chunks between two launches, as the reference's
// large-group chain does between update_x and apply_norm
// (core/kernel_glr.hpp:163-399, core/drot_glr.hpp:145-240):
//   gl_split_norm_kernel   iteration prologue; y = max(X + phi1 + phi2 - sC, 0)
//                          for the chunk, and per (group segment, column) the
//                          fp64 sum of y^2 over the group's rows in the chunk;
//                          X is not written
//   gl_split_apply_kernel  the same y again (X is unchanged, so the bits are
//                          the same), the group norm as the sum of its
//                          segments in chunk order, X := y * scale, and the
//                          row / column / objective partials of the fused pass
// 20 bytes per element (read X, C twice; write X) against the fused
// kernels' 12, the same as the reference's update_x + apply_norm.

template <int VEC, int RCH>
struct SplitChunk {
    static constexpr int R = kBlock * VEC * RCH;
};

// Loads of one chunk's column j: y = max((x + f1) + f2 - s c, 0) per element
// (rows past n_rows give 0).  Rows of chunk element (r, e): r0 + (r*kBlock + t)*VEC + e.
template <int VEC, int RCH>
__device__ __forceinline__ void split_prox(const GlPassArgs &A, int j, int r0, float f2,
                                           const float (&f1)[RCH][VEC], float (&y)[RCH][VEC],
                                           float (&cv)[RCH][VEC]) {
    const int N = A.n_rows;
    const float s = A.step;
    const float *xc = A.x + (size_t)j * N;
    const float *cc = A.c + (size_t)j * N;
    float xv[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
        const int i0 = r0 + (r * kBlock + (int)threadIdx.x) * VEC;
        if (i0 < N) {
            VecIO<VEC>::load(xc + i0, xv[r]);
            VecIO<VEC>::load(cc + i0, cv[r]);
        } else {
#pragma unroll
            for (int e = 0; e < VEC; e++) { xv[r][e] = 0.0f; cv[r][e] = 0.0f; }
        }
    }
#pragma unroll
    for (int r = 0; r < RCH; r++) {
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            const float tt = (xv[r][e] + f1[r][e]) + f2;
            y[r][e] = fmaxf(__builtin_fmaf(-s, cv[r][e], tt), 0.0f);
        }
    }
}

// Row potentials and groups of the chunk's rows (issued before the prologue's
// reductions in the norm kernel).
template <int VEC, int RCH>
__device__ __forceinline__ void split_rows(const GlPassArgs &A, bool explicit_phi, int r0,
                                           float (&h1)[RCH][VEC], int (&gid)[RCH][VEC]) {
    const float *hsrc = explicit_phi ? A.phi1 : A.ahat;
#pragma unroll
    for (int r = 0; r < RCH; r++) {
        const int i0 = r0 + (r * kBlock + (int)threadIdx.x) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; e++) { h1[r][e] = 0.0f; gid[r][e] = 0; }
        if (i0 < A.n_rows) {
            VecIO<VEC>::load(hsrc + i0, h1[r]);
            load_ints<VEC>(A.row_group + i0, gid[r]);
        }
    }
}

template <int VEC, int RCH>
__global__ __launch_bounds__(kBlock) void gl_split_norm_kernel(GlPassArgs A) {
    constexpr int R = SplitChunk<VEC, RCH>::R;
    __shared__ float s_y[R];
    __shared__ double s_red[3 * kWavesPerBlock];
    if (A.ctl.st->done) return;
    PrologueLoads pre;  // the prologue's first entries go out first (see gl_rows_kernel)
    prologue_issue(A.ctl, pre);
    __builtin_amdgcn_sched_barrier(0);
    const int rc = blockIdx.x % A.n_rt;
    const int ct = blockIdx.x / A.n_rt;
    const int N = A.n_rows, M = A.n_cols;
    const int r0 = rc * R;
    const int r1 = min(r0 + R, N);
    const bool explicit_phi = A.ctl.k == 0;
    float h1[RCH][VEC];
    int gid[RCH][VEC];
    split_rows<VEC, RCH>(A, explicit_phi, r0, h1, gid);
    const int g_first = A.row_group[r0];
    const int g_last = A.row_group[r1 - 1];
    const IterStart is = iteration_prologue(A.ctl, s_red, blockIdx.x == 0, &pre);
    if (is.finish) return;
    const float gamma = is.gamma;
    const float fM = (float)M, fN = (float)N;
    float f1[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++)
#pragma unroll
        for (int e = 0; e < VEC; e++) f1[r][e] = explicit_phi ? h1[r][e] : (h1[r][e] + gamma) / fM;
    const int ng = g_last - g_first + 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_uniform(threadIdx.x / kWave);
    const int j0 = ct * A.tc;
    const int j1 = min(j0 + A.tc, M);
    for (int j = j0; j < j1; j++) {
        const float h2 = explicit_phi ? A.phi2[j] : A.bhat[j];
        const float f2 = explicit_phi ? h2 : (h2 + gamma) / fN;
        float y[RCH][VEC], cv[RCH][VEC];
        split_prox<VEC, RCH>(A, j, r0, f2, f1, y, cv);
#pragma unroll
        for (int r = 0; r < RCH; r++)
#pragma unroll
            for (int e = 0; e < VEC; e++) s_y[(r * kBlock + (int)threadIdx.x) * VEC + e] = y[r][e];
        __syncthreads();
        // per group touching the chunk: fp64 sum of y^2 over its rows here
        if (A.lane_per_group) {
            for (int gl = threadIdx.x; gl < ng; gl += kBlock) {
                const int g = g_first + gl;
                const int lo = max(A.bounds[g], r0) - r0, hi = min(A.bounds[g + 1], r1) - r0;
                double n2 = 0.0;
                for (int i = lo; i < hi; i++) n2 += (double)s_y[i] * (double)s_y[i];
                A.pnorm[(size_t)(A.gseg[g] + rc - A.gclo[g]) * M + j] = n2;
            }
        } else {
            for (int gl = w; gl < ng; gl += kWavesPerBlock) {
                const int g = g_first + gl;
                const int lo = max(A.bounds[g], r0) - r0, hi = min(A.bounds[g + 1], r1) - r0;
                double n2 = 0.0;
                for (int i = lo + lane; i < hi; i += kWave) n2 += (double)s_y[i] * (double)s_y[i];
                n2 = wave_sum(n2);
                if (lane == 0) A.pnorm[(size_t)(A.gseg[g] + rc - A.gclo[g]) * M + j] = n2;
            }
        }
        __syncthreads();  // s_y is rewritten by the next column
    }
}

template <int VEC, int RCH>
__global__ __launch_bounds__(kBlock) void gl_split_apply_kernel(GlPassArgs A) {
    constexpr int R = SplitChunk<VEC, RCH>::R;
    __shared__ float s_scale[R];
    __shared__ double s_red[kWavesPerBlock];
    // the norm kernel of this iteration ran the prologue: done is set when the
    // solve ended with iteration k-1, and gamma is that of iteration k-1
    const SolverState *st = A.ctl.st;
    if (st->done) return;
    const int rc = blockIdx.x % A.n_rt;
    const int ct = blockIdx.x / A.n_rt;
    const int N = A.n_rows, M = A.n_cols;
    const int r0 = rc * R;
    const int r1 = min(r0 + R, N);
    const bool explicit_phi = A.ctl.k == 0;
    float h1[RCH][VEC];
    int gid[RCH][VEC];
    split_rows<VEC, RCH>(A, explicit_phi, r0, h1, gid);
    const int g_first = A.row_group[r0];
    const int g_last = A.row_group[r1 - 1];
    const float gamma = explicit_phi ? 0.0f : st->gamma;
    const float fM = (float)M, fN = (float)N;
    const double lam = (double)A.lambda;
    float f1[RCH][VEC];
    double racc[RCH][VEC];
#pragma unroll
    for (int r = 0; r < RCH; r++) {
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            f1[r][e] = explicit_phi ? h1[r][e] : (h1[r][e] + gamma) / fM;
            gid[r][e] -= g_first;  // local group index within the chunk
            racc[r][e] = 0.0;
        }
    }
    const int ng = g_last - g_first + 1;
    double oacc = 0.0;
    const int j0 = ct * A.tc;
    const int j1 = min(j0 + A.tc, M);
    for (int j = j0; j < j1; j++) {
        const float h2 = explicit_phi ? A.phi2[j] : A.bhat[j];
        const float f2 = explicit_phi ? h2 : (h2 + gamma) / fN;
        float y[RCH][VEC], cv[RCH][VEC];
        split_prox<VEC, RCH>(A, j, r0, f2, f1, y, cv);
        // group scales from the segments' partial norms, in chunk order
        for (int gl = threadIdx.x; gl < ng; gl += kBlock) {
            const int g = g_first + gl;
            double n2 = 0.0;
            for (int sg = A.gseg[g]; sg < A.gseg[g + 1]; sg++) n2 += A.pnorm[(size_t)sg * M + j];
            s_scale[gl] = group_scale(n2, lam);
        }
        __syncthreads();
        double csum = 0.0;
#pragma unroll
        for (int r = 0; r < RCH; r++) {
            const int i0 = r0 + (r * kBlock + (int)threadIdx.x) * VEC;
            if (i0 < N) {
                float xo[VEC];
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    const float v = y[r][e] * s_scale[gid[r][e]];
                    xo[e] = v;
                    racc[r][e] += (double)v;
                    csum += (double)v;
                    oacc += (double)v * (double)cv[r][e];
                }
                VecIO<VEC>::store(A.x + (size_t)j * N + i0, xo);
            }
        }
        csum = block_sum(csum, s_red);  // its barriers also order s_scale's reuse
        if (threadIdx.x == 0) A.pcol[(size_t)rc * M + j] = csum;
    }
#pragma unroll
    for (int r = 0; r < RCH; r++) {
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            const int i = r0 + (r * kBlock + (int)threadIdx.x) * VEC + e;
            if (i < N) A.prow[(size_t)ct * N + i] = racc[r][e];
        }
    }
    const double o = block_sum(oacc, s_red);
    if (threadIdx.x == 0) A.pobj[blockIdx.x] = o;
}

// ----------------------------------------------------------------- finalize

template <int BATCH>
__global__ __launch_bounds__(kBlock) void finalize_kernel(FinalizeArgs A) {
    __shared__ double s_part[kBlock / kFinRows][kFinRows];
    __shared__ double s_red[3 * kWavesPerBlock];
    FinJob F;
    F.done = &A.st->done;
    F.prow = A.prow;
    F.pcol = A.pcol;
    F.pobj = A.pobj;
    F.p = A.p;
    F.q = A.q;
    F.a = A.a;
    F.b = A.b;
    F.ahat = A.ahat;
    F.bhat = A.bhat;
    F.pfin = A.pfin + 3 * blockIdx.x;
    F.n_rows = A.n_rows;
    F.n_cols = A.n_cols;
    F.n_rt = A.n_rt;
    F.n_ct = A.n_ct;
    F.n_obj = A.n_obj;
    F.rb = A.rb;
    F.n_fin = A.rb + A.cb;
    F.blk = blockIdx.x;
    F.with_obj = A.with_obj;
    finalize_block<BATCH>(F, s_part, s_red);
}

// ----------------------------------------------------------------- fix_x

// core/kernel_qr.hpp:469-482: undo the -sC shift once a QR solve has converged
// (convergence is only tested on even iterations, whose stored X is X' - sC).
__global__ __launch_bounds__(kBlock) void fix_x_kernel(float *x, const float *c, float step,
                                                      size_t total, const SolverState *st) {
    // fix_x runs once after the loop; st->converged is final by then
    if (!st->converged) return;
    const size_t nthr = (size_t)gridDim.x * blockDim.x;
    const size_t n4 = total / 4;
    const bool aligned = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(c)) & 15u) == 0;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (aligned) {
        for (; e < n4; e += nthr) {
            float4 xv = reinterpret_cast<float4 *>(x)[e];
            const float4 cv = reinterpret_cast<const float4 *>(c)[e];
            xv.x = fmaxf(__builtin_fmaf(step, cv.x, xv.x), 0.0f);
            xv.y = fmaxf(__builtin_fmaf(step, cv.y, xv.y), 0.0f);
            xv.z = fmaxf(__builtin_fmaf(step, cv.z, xv.z), 0.0f);
            xv.w = fmaxf(__builtin_fmaf(step, cv.w, xv.w), 0.0f);
            reinterpret_cast<float4 *>(x)[e] = xv;
        }
        e = n4 * 4 + ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    }
    for (; e < total; e += nthr) x[e] = fmaxf(__builtin_fmaf(step, c[e], x[e]), 0.0f);
}

// After the loop: the segments the last executed pass (k_run - 1) flagged as
// all-zero were not stored, so write what that pass would have stored there:
// X' - sC with X' = +0 (fmaf(-s, c, 0), the pass's own expression) after an
// even pass, X' = +0 after an odd one.  Segment (rt, j) covers rows
// [rt*tr, rt*tr + tr) of column j; its flag is zf[(k & 1) * zf_bytes + rt*n_cols + j].
__global__ __launch_bounds__(kBlock) void qr_unflag_kernel(float *x, const float *c, float step,
                                                          const unsigned char *zf,
                                                          size_t zf_bytes, int n_rows,
                                                          int n_cols, int tr,
                                                          const SolverState *st) {
    const int k_last = st->k_run - 1;
    if (k_last < 0) return;
    const unsigned char *z = zf + (size_t)(k_last & 1) * zf_bytes;
    const bool even = (k_last & 1) == 0;
    const size_t total = (size_t)n_rows * (size_t)n_cols;
    const size_t nthr = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += nthr) {
        const int j = (int)(e / (size_t)n_rows);
        const int i = (int)(e - (size_t)j * (size_t)n_rows);
        if (z[(size_t)(i / tr) * (size_t)n_cols + j])
            x[e] = even ? __builtin_fmaf(-step, c[e], 0.0f) : 0.0f;
    }
}

// Materialise phi1/phi2 of the current state (for state export / resume).
__global__ __launch_bounds__(kBlock) void export_phi_kernel(const float *ahat, const float *bhat,
                                                           const float *phi1, const float *phi2,
                                                           const SolverState *st, int n_rows,
                                                           int n_cols, float *o1, float *o2) {
    const int mode = st->phi_mode;
    const float gamma = st->gamma;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nthr = gridDim.x * blockDim.x;
    for (int i = tid; i < n_rows; i += nthr) o1[i] = phi_of(mode, phi1, ahat, i, gamma, (float)n_cols);
    for (int j = tid; j < n_cols; j += nthr) o2[j] = phi_of(mode, phi2, bhat, j, gamma, (float)n_rows);
}

// ----------------------------------------------------------------- launchers

static int grid_for(size_t work, int per_block_items, int cap) {
    size_t g = (work + (size_t)per_block_items - 1) / (size_t)per_block_items;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = (size_t)cap;
    return (int)g;
}

hipError_t launch_init(const InitArgs &a, hipStream_t s) {
    const size_t total = (size_t)a.n_rows * (size_t)a.n_cols;
    const int grid = grid_for(total, kBlock * 8, 8192);
    hipLaunchKernelGGL(init_kernel, dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

bool gl_tile_supported(const GlTile &t) {
    if (t.vec == 4) return t.rch == 1 || t.rch == 2 || t.rch == 4 || t.rch == 8;
    if (t.vec == 1) return t.rch == 1 || t.rch == 2 || t.rch == 4 || t.rch == 8 ||
                           t.rch == 16 || t.rch == 32;
    return false;
}

hipError_t launch_gl_pass(const GlTile &t, int grid, size_t lds, const GlPassArgs &a,
                          hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (t.split) {
        // two launches; timing events bracket both (they include the dispatches)
        if (e0 && e1) {
            hipError_t er = hipEventRecord(e0, s);
            if (er != hipSuccess) return er;
        }
#define DROT_GLS_CASE(V, R)                                                                \
        if (t.vec == V && t.rch == R) {                                                    \
            hipLaunchKernelGGL((gl_split_norm_kernel<V, R>), dim3(grid), dim3(kBlock), 0, s, a); \
            hipLaunchKernelGGL((gl_split_apply_kernel<V, R>), dim3(grid), dim3(kBlock), 0, s, a); \
        } else
        DROT_GLS_CASE(4, 1) DROT_GLS_CASE(4, 4) DROT_GLS_CASE(1, 4) DROT_GLS_CASE(1, 16)
        return hipErrorInvalidValue;
#undef DROT_GLS_CASE
        hipError_t er = hipGetLastError();
        if (er == hipSuccess && e0 && e1) er = hipEventRecord(e1, s);
        return er;
    }
    if (t.wave == 2) {
#define DROT_GLR_CASE(V, R)                                                               \
        if (t.vec == V && t.rch == R) {                                                   \
            if (e0 && e1)                                                                 \
                hipExtLaunchKernelGGL((gl_rows_kernel<V, R>), dim3(grid), dim3(kBlock), lds, s, \
                                      e0, e1, 0, a);                                      \
            else                                                                          \
                hipLaunchKernelGGL((gl_rows_kernel<V, R>), dim3(grid), dim3(kBlock), lds, s, a); \
            return hipGetLastError();                                                     \
        }
        DROT_GLR_CASE(4, 1) DROT_GLR_CASE(4, 2)
        DROT_GLR_CASE(1, 1) DROT_GLR_CASE(1, 2) DROT_GLR_CASE(1, 4) DROT_GLR_CASE(1, 8)
#undef DROT_GLR_CASE
        return hipErrorInvalidValue;
    }
    if (t.wave) {
#define DROT_GLW_CASE(V, R)                                                               \
        if (t.vec == V && t.rch == R) {                                                   \
            if (e0 && e1)                                                                 \
                hipExtLaunchKernelGGL((gl_wave_kernel<V, R>), dim3(grid), dim3(kBlock), lds, s, \
                                      e0, e1, 0, a);                                      \
            else                                                                          \
                hipLaunchKernelGGL((gl_wave_kernel<V, R>), dim3(grid), dim3(kBlock), lds, s, a); \
            return hipGetLastError();                                                     \
        }
        DROT_GLW_CASE(4, 1) DROT_GLW_CASE(4, 2) DROT_GLW_CASE(4, 4) DROT_GLW_CASE(4, 8)
        DROT_GLW_CASE(1, 1) DROT_GLW_CASE(1, 2) DROT_GLW_CASE(1, 4) DROT_GLW_CASE(1, 8)
#undef DROT_GLW_CASE
        return hipErrorInvalidValue;
    }
#define DROT_GL_CASE(V, R)                                                                \
    if (t.vec == V && t.rch == R) {                                                       \
        if (e0 && e1)                                                                     \
            hipExtLaunchKernelGGL((gl_pass_kernel<V, R>), dim3(grid), dim3(kBlock), lds, s, \
                                  e0, e1, 0, a);                                          \
        else                                                                              \
            hipLaunchKernelGGL((gl_pass_kernel<V, R>), dim3(grid), dim3(kBlock), lds, s, a); \
        return hipGetLastError();                                                         \
    }
    DROT_GL_CASE(4, 1) DROT_GL_CASE(4, 2) DROT_GL_CASE(4, 4) DROT_GL_CASE(4, 8)
    DROT_GL_CASE(1, 1) DROT_GL_CASE(1, 2) DROT_GL_CASE(1, 4) DROT_GL_CASE(1, 8)
    DROT_GL_CASE(1, 16) DROT_GL_CASE(1, 32)
#undef DROT_GL_CASE
    return hipErrorInvalidValue;
}

// Rows with many partials (GL's column tiles: 250 at 1500x1000) get more loads
// in flight per thread; few partials keep the short unrolled batch, whose
// extra predicated loads would cost more than they save.
hipError_t launch_finalize(int grid, const FinalizeArgs &a, hipStream_t s) {
    constexpr int kParts = kBlock / kFinRows;
    const int cnt = a.n_ct > a.n_rt ? a.n_ct : a.n_rt;
    if (cnt <= 8 * kParts)
        hipLaunchKernelGGL(finalize_kernel<8>, dim3(grid), dim3(kBlock), 0, s, a);
    else if (cnt <= 16 * kParts)
        hipLaunchKernelGGL(finalize_kernel<16>, dim3(grid), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(finalize_kernel<32>, dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fix_x(float *x, const float *c, float step, size_t total,
                        const SolverState *st, hipStream_t s) {
    const int grid = grid_for(total / 4 + 1, kBlock * 4, 8192);
    hipLaunchKernelGGL(fix_x_kernel, dim3(grid), dim3(kBlock), 0, s, x, c, step, total, st);
    return hipGetLastError();
}

hipError_t launch_qr_unflag(float *x, const float *c, float step, const unsigned char *zf,
                            size_t zf_bytes, int n_rows, int n_cols, int tr,
                            const SolverState *st, hipStream_t s) {
    const int grid = grid_for((size_t)n_rows * (size_t)n_cols, kBlock, 8192);
    hipLaunchKernelGGL(qr_unflag_kernel, dim3(grid), dim3(kBlock), 0, s, x, c, step, zf,
                       zf_bytes, n_rows, n_cols, tr, st);
    return hipGetLastError();
}

hipError_t launch_export_phi(const float *ahat, const float *bhat, const float *phi1,
                             const float *phi2, const SolverState *st, int n_rows,
                             int n_cols, float *o1, float *o2, hipStream_t s) {
    const int grid = grid_for((size_t)(n_rows > n_cols ? n_rows : n_cols), kBlock, 1024);
    hipLaunchKernelGGL(export_phi_kernel, dim3(grid), dim3(kBlock), 0, s, ahat, bhat, phi1,
                       phi2, st, n_rows, n_cols, o1, o2);
    return hipGetLastError();
}

}  // namespace drot
