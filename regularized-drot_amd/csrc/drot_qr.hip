// QR-DROT fused pass over X (gfx950): one launch streams the whole matrix once
// and produces every per-tile partial the dual update needs
// (core/kernel_qr.hpp:171-376 restated):
//   X' = max(X + phi1 + phi2 - sC, 0)/kappa,  stored X = X' - sC (even k) | X'
//   (odd k, whose kernel never reads C: the -sC is already folded into X),
//   phi = (hat + gamma)/n formed on the fly from the finalize kernel's output,
//   column partials by a butterfly over the lanes of a column, row partials in
//   fp64 registers combined through LDS, objective partials on even k.
// The pass contains no cross-block synchronisation: an agent-scope release
// (buffer_wbl2) per block costs more than the whole pass on this chip, so the
// grid-wide steps of an iteration are the kernel boundaries pass -> finalize
// -> next pass, whose prologue finishes the previous iteration
// (iteration_prologue in drot_device.h).
#include "drot_kernels.h"

#include <math.h>

namespace drot {

namespace {

// 16-byte / 4-byte global accesses.  Store policy ST: 0 = default (L2
// write-back global store), 2 = device-scope write-through (sc1) buffer store
// through a per-tile buffer resource whose range is the tile's columns, so a
// store outside the tile is dropped by the hardware range check.
// Write-through keeps the streamed X out of the L2 dirty set, which shortens
// the end-of-kernel cache write-back.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBufDword3 = 0x00020000;  // raw buffer, 32-bit data format
constexpr int kPolicySc1 = 16;          // cache-policy operand: sc1

template <int VEC>
struct Io;

template <>
struct Io<4> {
    static __device__ __forceinline__ void load(const float *p, float (&v)[4]) {
        const f32x4 t = *reinterpret_cast<const f32x4 *>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    }
    static __device__ __forceinline__ void store(float *p, const float (&v)[4]) {
        f32x4 t;
        t.x = v[0]; t.y = v[1]; t.z = v[2]; t.w = v[3];
        *reinterpret_cast<f32x4 *>(p) = t;
    }
    static __device__ __forceinline__ void store_wt(__amdgpu_buffer_rsrc_t r, int byte_off,
                                                    const float (&v)[4]) {
        u32x4 d;
        d.x = __float_as_uint(v[0]);
        d.y = __float_as_uint(v[1]);
        d.z = __float_as_uint(v[2]);
        d.w = __float_as_uint(v[3]);
        __builtin_amdgcn_raw_buffer_store_b128(d, r, byte_off, 0, kPolicySc1);
    }
};

template <>
struct Io<1> {
    static __device__ __forceinline__ void load(const float *p, float (&v)[1]) { v[0] = *p; }
    static __device__ __forceinline__ void store(float *p, const float (&v)[1]) { *p = v[0]; }
    static __device__ __forceinline__ void store_wt(__amdgpu_buffer_rsrc_t r, int byte_off,
                                                    const float (&v)[1]) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[0]), r, byte_off, 0, kPolicySc1);
    }
};

}  // namespace

// ST: store policy of X (see Io); ST == 2 needs tc * n_rows * 4 < 2^31.
//
// Software pipelined: the first batch of X/C loads (and the row potentials)
// is issued before the prologue's reductions, and every batch's loads are in
// flight while the previous batch is computed and stored.
template <int VEC, int LPC, bool EVEN, int ST, int U>
__global__ __launch_bounds__(kBlock) void qr_pass_kernel(QrPassArgs A) {
    constexpr int CPW = kWave / LPC;  // columns per wave step
    constexpr int TR = LPC * VEC;     // tile rows
    // U: column steps per batch (two batches in flight)
    __shared__ double s_row[kWavesPerBlock * CPW][TR];
    __shared__ double s_red[3 * kWavesPerBlock];

    if (A.ctl.st->done) return;
    const int N = A.n_rows, M = A.n_cols;
    const int n_rt = A.n_rt;
    const int nstep = A.nstep;
    const float fM = (float)M, fN = (float)N;
    // XCD-aware tile order: blocks are dispatched round-robin over the 8 XCDs
    // (block b on XCD b % 8), so give each XCD a contiguous run of tiles in
    // row-tile-major order.  Vertically adjacent tiles then share an L2 and the
    // cache lines straddling their boundary are fetched once.
    const int n_tiles = n_rt * A.n_ct;
    const int per_xcd = (n_tiles + kXcds - 1) / kXcds;
    const int tile = (int)(blockIdx.x % kXcds) * per_xcd + (int)(blockIdx.x / kXcds);
    if (tile >= n_tiles) return;
    const int rt = tile % n_rt;
    const int ct = tile / n_rt;
    const int tc = nstep * kWavesPerBlock * CPW;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_uniform(threadIdx.x / kWave);
    const int cs = lane / LPC;
    const int rl = lane % LPC;
    const int row = rt * TR + rl * VEC;
    const bool row_ok = row < N;  // VEC == 4 only when N % 4 == 0
    const int col0 = ct * tc + w * CPW + cs;  // column of step t: col0 + t * 4 * CPW
    constexpr int kColStride = kWavesPerBlock * CPW;
    const bool explicit_phi = (A.ctl.k == 0);
    const float *hat2 = explicit_phi ? A.phi2 : A.bhat;

    float xa[U][VEC], ca[U][VEC], h2a[U];
    auto load_batch = [&](int t0, float (&xv)[U][VEC], float (&cv)[U][VEC], float (&h2)[U]) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u;
            const int j = col0 + t * kColStride;
            const bool col_ok = (t < nstep) && (j < M);
            h2[u] = col_ok ? hat2[j] : 0.0f;
            if (col_ok && row_ok) {
                const size_t off = (size_t)j * (size_t)N + (size_t)row;
                Io<VEC>::load(A.x + off, xv[u]);
                if (EVEN) Io<VEC>::load(A.c + off, cv[u]);
                else {
#pragma unroll
                    for (int e = 0; e < VEC; e++) cv[u][e] = 0.0f;
                }
            } else {
#pragma unroll
                for (int e = 0; e < VEC; e++) { xv[u][e] = 0.0f; cv[u][e] = 0.0f; }
            }
        }
    };
    load_batch(0, xa, ca, h2a);
    float h1[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++)
        h1[e] = row_ok ? (explicit_phi ? A.phi1[row + e] : A.ahat[row + e]) : 0.0f;

    // write-through stores address the tile's column range [tile_col0, +tc)
    const int tile_col0 = ct * tc;
    const int tile_cols = min(tc, M - tile_col0);
    __amdgpu_buffer_rsrc_t xrsrc = __builtin_amdgcn_make_buffer_rsrc(
        A.x + (size_t)tile_col0 * (size_t)N, 0, tile_cols * N * 4, kBufDword3);

    const IterStart is = iteration_prologue(A.ctl, s_red);
    if (is.finish) return;
    const float gamma = is.gamma;
    const float s = A.step, kappa = A.kappa;

    float f1[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) f1[e] = explicit_phi ? h1[e] : (h1[e] + gamma) / fM;
    double racc[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) racc[e] = 0.0;
    double oacc = 0.0;

    for (int t0 = 0; t0 < nstep; t0 += U) {
        float xb[U][VEC], cb[U][VEC], h2b[U];
        if (t0 + U < nstep) load_batch(t0 + U, xb, cb, h2b);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u;
            if (t >= nstep) break;  // wave-uniform
            const int j = col0 + t * kColStride;
            const bool col_ok = j < M;
            const bool ok = col_ok && row_ok;
            const float f2 = explicit_phi ? h2a[u] : (h2a[u] + gamma) / fN;
            float xo[VEC];
            double csum = 0.0;
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float tt = (xa[u][e] + f1[e]) + f2;
                float y = EVEN ? __builtin_fmaf(-s, ca[u][e], tt) : tt;
                y = fmaxf(y, 0.0f);
                const float xp = y / kappa;
                xo[e] = EVEN ? __builtin_fmaf(-s, ca[u][e], xp) : xp;
                if (ok) {
                    racc[e] += (double)xp;
                    csum += (double)xp;
                    if (EVEN) oacc += (double)xp * (double)ca[u][e];
                }
            }
            if (ok) {
                if (ST == 2) {
                    Io<VEC>::store_wt(xrsrc, ((j - tile_col0) * N + row) * 4, xo);
                } else {
                    const size_t off = (size_t)j * (size_t)N + (size_t)row;
                    Io<VEC>::store(A.x + off, xo);
                }
            }
            csum = seg_sum<LPC>(csum);
            if (rl == 0 && col_ok) A.pcol[(size_t)rt * M + j] = csum;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            h2a[u] = h2b[u];
#pragma unroll
            for (int e = 0; e < VEC; e++) { xa[u][e] = xb[u][e]; ca[u][e] = cb[u][e]; }
        }
    }

    // strided-axis partials: fixed-order combine of the 4*CPW column slots
    const int slot = w * CPW + cs;
#pragma unroll
    for (int e = 0; e < VEC; e++) s_row[slot][rl * VEC + e] = racc[e];
    __syncthreads();
    for (int r = threadIdx.x; r < TR; r += kBlock) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < kWavesPerBlock * CPW; q++) sum += s_row[q][r];
        const int i = rt * TR + r;
        if (i < N) A.prow[(size_t)ct * N + i] = sum;
    }
    if (EVEN) {
        const double o = block_sum(oacc, s_red);
        if (threadIdx.x == 0) A.pobj[tile] = o;
    }
}

// ----------------------------------------------------------------- launch

bool qr_tile_supported(const QrTile &t) {
    return (t.vec == 1 || t.vec == 4) && (t.lpc == 16 || t.lpc == 32 || t.lpc == 64) &&
           (t.nt == 0 || t.nt == 2) && (t.u == 2 || t.u == 4);
}

template <int VEC, int LPC, int ST, int U>
static void launch_pass_t(bool even, int grid, const QrPassArgs &a, hipStream_t s,
                          hipEvent_t e0, hipEvent_t e1) {
    auto k = even ? qr_pass_kernel<VEC, LPC, true, ST, U> : qr_pass_kernel<VEC, LPC, false, ST, U>;
    if (e0 && e1)  // timing: events bracket the kernel's own execution
        hipExtLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, a);
    else
        hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, a);
}

template <int VEC, int LPC>
static void launch_pass_v(const QrTile &t, bool even, int grid, const QrPassArgs &a,
                          hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (t.nt == 2) {
        if (t.u == 2) launch_pass_t<VEC, LPC, 2, 2>(even, grid, a, s, e0, e1);
        else launch_pass_t<VEC, LPC, 2, 4>(even, grid, a, s, e0, e1);
    } else {
        if (t.u == 2) launch_pass_t<VEC, LPC, 0, 2>(even, grid, a, s, e0, e1);
        else launch_pass_t<VEC, LPC, 0, 4>(even, grid, a, s, e0, e1);
    }
}

int qr_pass_grid(int n_rt, int n_ct) {
    const int per = (n_rt * n_ct + kXcds - 1) / kXcds;
    return per * kXcds;
}

hipError_t launch_qr_pass(const QrTile &t, bool even, int grid, const QrPassArgs &a,
                          hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (t.vec == 4) {
        if (t.lpc == 16) launch_pass_v<4, 16>(t, even, grid, a, s, e0, e1);
        else if (t.lpc == 32) launch_pass_v<4, 32>(t, even, grid, a, s, e0, e1);
        else launch_pass_v<4, 64>(t, even, grid, a, s, e0, e1);
    } else {
        if (t.lpc == 16) launch_pass_v<1, 16>(t, even, grid, a, s, e0, e1);
        else if (t.lpc == 32) launch_pass_v<1, 32>(t, even, grid, a, s, e0, e1);
        else launch_pass_v<1, 64>(t, even, grid, a, s, e0, e1);
    }
    return hipGetLastError();
}

}  // namespace drot
