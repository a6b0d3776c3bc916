// QR-DROT pass over one tile (gfx950), shared by the single-problem and the
// batched pass kernels so both run exactly the same per-element arithmetic.
//
// Restates quadratic_regularizer_update_x_{even,odd} (core/kernel_qr.hpp:171-376):
//   X' = max(X + phi1 + phi2 - sC, 0)/kappa,  stored X = X' - sC (even k) | X'
//   (odd k, whose kernel never reads C: the -sC is already folded into X),
//   phi = (hat + gamma)/n formed on the fly from the finalize kernel's output.
// A tile is TR = LPC*VEC rows by tc columns of one problem; LPC lanes cover
// the tile rows of one column with 16-byte accesses, so a wave handles 64/LPC
// columns per step.  Column partials: butterfly over those lanes.  Row
// partials: fp64 registers across the tile's columns, combined through LDS in
// a fixed order.  Objective partials on even k.
//
// Software pipelined: the first batch of X/C loads (and the row potentials)
// is issued before the prologue's reductions, and every batch's loads are in
// flight while the previous batch is computed and stored.
//
// Zero segments.  The plan is sparse long before convergence (at 3000^2 about
// 95% of the 64-row column segments of X' are all zero after a tenth of the
// iterations).  A segment is the TR rows of one column inside a tile.  When
// every X' of a segment is +0 (bit pattern 0), the pass stores nothing there
// and sets the segment's flag in zout; what it would have stored is known
// without the store: +0 after an odd pass, fmaf(-s, c, +0) after an even one.
// The next pass reads its flags (zin) first: on a flagged segment the even
// pass skips the X load (X = +0) and the odd pass loads C in place of X and
// forms fmaf(-s, c, +0).  Every element sees the same operands as without the
// flags, so results keep their bits, and a settled segment moves 4 bytes per
// element (C) instead of 12 (even) or 8 (odd).
//
// Screening (SC, with the flags).  A flagged segment also skips its C load,
// and with it every load, when a bound proves that all its X' stay +0.  For
// element i of segment (tile rows, column j), y = f1_i + f2_j - s c_ij up to
// the roundings of the pass; the segment keeps a headroom E >= max_i (f1_i -
// s c_ij) - G1 from the last pass that read its C, where G1 is the tile's
// running sum of its rows' largest rise of f1 from one pass to the next, so
// that E + G1(now) bounds max_i (f1_i - s c_ij) at any later pass.  A flagged
// segment is skipped when E + G1 + f2_j lies below -2^-21 (max|f1| + |f2_j|),
// which covers the roundings of y (at most three, each of 2^-24 of the
// largest operand): every y is then negative, X' = +0 as the full pass would
// give, and the flag stays set.  The first batch of a tile is issued before
// the prologue and is not screened.
#pragma once

#include <type_traits>

#include "drot_kernels.h"

namespace drot {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int kBufDword3 = 0x00020000;  // raw buffer, 32-bit data format
constexpr int kPolicySc1 = 16;          // cache-policy operand: sc1

// 16-byte / 4-byte global accesses.  Store policy ST: 0 = default (L2
// write-back global store), 2 = device-scope write-through (sc1) buffer store
// through a per-tile buffer resource whose range is the tile's columns, so a
// store outside the tile is dropped by the hardware range check.
// Write-through keeps the streamed X out of the L2 dirty set, which shortens
// the end-of-kernel cache write-back.
template <int VEC>
struct Io;

// The 16-byte accesses need only 4-byte alignment (gfx950's multi-dword
// global and buffer accesses are dword-aligned), so a column may start
// anywhere: p is read and written through 4-byte-aligned copies.
template <>
struct Io<4> {
    static __device__ __forceinline__ void load(const float *p, float (&v)[4]) {
        f32x4 t;
        __builtin_memcpy(&t, p, sizeof(t));
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    }
    static __device__ __forceinline__ void store(float *p, const float (&v)[4]) {
        f32x4 t;
        t.x = v[0]; t.y = v[1]; t.z = v[2]; t.w = v[3];
        __builtin_memcpy(p, &t, sizeof(t));
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

// Everything one tile needs.  x/c point at the problem's matrix (leading
// dimension ld >= n_rows), the vectors at the problem's slices.  When ld is
// not a multiple of VEC, the last chunk of a column's rows is partial: it is
// read and written element by element (a 16-byte access would reach into the
// next column, or past the end of the matrix).
struct QrTileJob {
    float *x;
    const float *c;
    const float *phi1, *phi2, *ahat, *bhat;
    double *prow;   // problem's [n_ct][n_rows] row partials
    double *pcol;   // problem's [n_rt][n_cols] column partials
    double *pobj;   // this tile's objective partial
    int n_rows, n_cols, ld;
    int rt, ct, tc, nstep;
    int prio;       // 1: waves raise their issue priority while many column steps remain
    float step, kappa;
    // zero-segment flags, [n_rt][n_cols] bytes, both null or both set
    // (needs nstep <= LPC, 2 LPC in the branch-free 16-byte tiles: lane rl of
    // a column group loads the flag of step rl, and of step rl + LPC)
    const unsigned char *zin;
    unsigned char *zout;
    // screening (SC): the problem's [n_rt][n_cols] headrooms E, this tile's G1,
    // and the problem's row potentials of the previous pass (read) and of this
    // one (written by the tiles of column tile 0)
    float *seg_e;
    double *g1;
    const float *f1_prev;
    float *f1_out;
};

// y / kappa, correctly rounded, from the correctly rounded reciprocal
// rkappa = RN(1/kappa): q0 = RN(y * rkappa) is within one ulp of y/kappa, the
// remainder y - kappa*q0 is then exact in one fma, and RN(q0 + rem*rkappa) is
// the correctly rounded quotient (Markstein's theorem, round-to-nearest).
// The remainder is exact only while it does not underflow, so quotients below
// 2^-90 (and above 2^120, or not finite) go through the true division.
__device__ __forceinline__ float div_kappa(float y, float kappa, float rkappa) {
    const float q0 = y * rkappa;
    const float rem = __builtin_fmaf(-q0, kappa, y);
    return __builtin_fmaf(rem, rkappa, q0);
}

__device__ __forceinline__ bool div_kappa_slow(float y, float q) {
    return y > 0.0f && !(q >= 0x1p-90f && q <= 0x1p+120f);
}

// Issue priority from the share of column steps still ahead (3 at the start,
// 0 for the last batch).  When every tile of a launch is resident at once,
// the SIMD arbiter favours the oldest waves, so the blocks dispatched last on
// a CU fall behind and finish alone; ranking waves by remaining work keeps the
// blocks of the launch in step and shortens that tail.
__device__ __forceinline__ void set_issue_priority(int remaining, int nstep) {
    const int pr = min(3, max(0, (4 * remaining - 1) / nstep));
    if (pr == 3) __builtin_amdgcn_s_setprio(3);
    else if (pr == 2) __builtin_amdgcn_s_setprio(2);
    else if (pr == 1) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// The LPC bits of ballot b that belong to column group cs (lanes cs*LPC ...).
template <int LPC>
__device__ __forceinline__ unsigned long long group_bits(unsigned long long b, int cs) {
    if constexpr (LPC == 64) {
        return b;
    } else {
        return (b >> (cs * LPC)) & ((1ull << LPC) - 1ull);
    }
}

// End of a tile: the row partials of the tile's columns (fixed-order combine of
// the kWavesPerBlock*CPW column slots) and, on even passes, the objective
// partial.  The objective's wave sums meet behind the same barrier (in their
// own LDS slots: s_red may still be read by the prologue of a slower wave) and
// are added in wave order as block_sum does.
// Rows of a lane's slice.  A lane holds VEC rows of a column as NCH chunks of
// CH = min(VEC, 4) consecutive rows; chunk h of lane rl starts at tile row
// h*CH*LPC + rl*CH, so one 16-byte access per chunk covers CH*LPC contiguous
// rows of the column across the column group (VEC <= 4: one chunk, rows
// rl*VEC ... rl*VEC + VEC-1).
template <int VEC>
struct LaneSlice {
    static constexpr int CH = VEC < 4 ? VEC : 4;
    static constexpr int NCH = VEC / CH;
};

template <int VEC, int LPC>
__device__ __forceinline__ int tile_row(int rl, int e) {
    constexpr int CH = LaneSlice<VEC>::CH;
    return (e / CH) * (CH * LPC) + rl * CH + (e % CH);
}

template <int VEC, int LPC, bool EVEN>
__device__ __forceinline__ void tile_partials_out(const QrTileJob &J, const double (&racc)[VEC],
                                                  double oacc, double (*s_row)[LPC * VEC]) {
    constexpr int CPW = kWave / LPC;
    constexpr int TR = LPC * VEC;
    __shared__ double s_obj[kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_uniform(threadIdx.x / kWave);
    const int slot = w * CPW + lane / LPC;
    const int rl = lane % LPC;
#pragma unroll
    for (int e = 0; e < VEC; e++) s_row[slot][tile_row<VEC, LPC>(rl, e)] = racc[e];
    if (EVEN) {
        const double ow = wave_sum(oacc);
        if (lane == 0) s_obj[w] = ow;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < TR; r += kBlock) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < kWavesPerBlock * CPW; q++) sum += s_row[q][r];
        const int i = J.rt * TR + r;
        if (i < J.n_rows) J.prow[(size_t)J.ct * J.n_rows + i] = sum;
    }
    if (EVEN && threadIdx.x == 0) {
        double o = 0.0;
#pragma unroll
        for (int q = 0; q < kWavesPerBlock; q++) o += s_obj[q];
        *J.pobj = o;
    }
}

// ------------------------------------------------------------ lean tile loop
//
// When every VEC-row chunk lies inside one column (ld % VEC == 0) and a tile's
// columns span less than 2^31 bytes, a tile runs without a branch per element
// or per lane.  X, C, the column potentials, the column partials and the zero
// flags are accessed through buffer resources over the tile's own columns; a
// chunk outside the problem (a row past n_rows, a column past n_cols) or a
// zero-flagged X gets an offset past the resource's range, so the hardware
// returns zeros for its loads and drops its stores.  Rows past n_rows and
// columns past n_cols get the potential -inf, which makes their X' +0, so they
// add +0 to every partial sum: the sums, the flags and the stored X keep the
// bits of the masked formulation of qr_tile_generic.  Padding rows of a batch
// column (n_rows <= row < ld) read C as 0 on even passes, so whatever the
// padding holds, the objective partial gets +0 from them and they store +0.
// The batches of loads alternate between two register sets, so no copy
// rotates the software pipeline.

// Byte offset past every tile's range (tiles span less than 2^31 bytes).
constexpr unsigned kOob = 0x80000000u;

// The largest number of columns a batch of loads reaches past the tile
// (U * kWavesPerBlock * CPW <= 4 * 4 * 4).
constexpr int kLeanColSlack = 64;

template <int VEC>
__device__ __forceinline__ bool qr_lean_ok(const QrTileJob &J) {
    return (J.ld % LaneSlice<VEC>::CH) == 0 &&
           (unsigned long long)(J.tc + kLeanColSlack) * (unsigned long long)J.ld * 4ull <
               (1ull << 31);
}

template <int VEC>
struct BufIo;

template <>
struct BufIo<4> {
    static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                float (&v)[4]) {
        const u32x4 d = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
        v[0] = __uint_as_float(d.x);
        v[1] = __uint_as_float(d.y);
        v[2] = __uint_as_float(d.z);
        v[3] = __uint_as_float(d.w);
    }
    template <int POL>
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                 const float (&v)[4]) {
        u32x4 d;
        d.x = __float_as_uint(v[0]);
        d.y = __float_as_uint(v[1]);
        d.z = __float_as_uint(v[2]);
        d.w = __float_as_uint(v[3]);
        __builtin_amdgcn_raw_buffer_store_b128(d, r, (int)off, 0, POL);
    }
};

template <>
struct BufIo<1> {
    static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                float (&v)[1]) {
        v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    }
    template <int POL>
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                 const float (&v)[1]) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[0]), r, (int)off, 0, POL);
    }
};

// A lane's VEC rows of one column: one access per chunk, at its own offset.
template <int VEC>
struct LaneIo {
    static constexpr int CH = LaneSlice<VEC>::CH;
    static constexpr int NCH = LaneSlice<VEC>::NCH;
    static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t r,
                                                const unsigned (&off)[NCH], float (&v)[VEC]) {
#pragma unroll
        for (int h = 0; h < NCH; h++) {
            float t[CH];
            BufIo<CH>::load(r, off[h], t);
#pragma unroll
            for (int i = 0; i < CH; i++) v[h * CH + i] = t[i];
        }
    }
    template <int POL>
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t r,
                                                 const unsigned (&off)[NCH],
                                                 const float (&v)[VEC]) {
#pragma unroll
        for (int h = 0; h < NCH; h++) {
            float t[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) t[i] = v[h * CH + i];
            BufIo<CH>::template store<POL>(r, off[h], t);
        }
    }
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, kBufDword3);
}

// div_kappa's quotient a/b (any sign of a) leaves the range where its
// correction step is exact
__device__ __forceinline__ bool div_rn_slow(float a, float q) {
    return a != 0.0f && !(fabsf(q) >= 0x1p-90f && fabsf(q) <= 0x1p+120f);
}

template <int VEC, int LPC, bool EVEN, int ST, int U, bool ZS, bool SC>
__device__ __forceinline__ void qr_tile_lean(const QrTileJob &J, const IterCtl &ctl,
                                             bool leader_block, double (*s_row)[LPC * VEC],
                                             double *s_red) {
    constexpr int CPW = kWave / LPC;
    constexpr int TR = LPC * VEC;
    constexpr int kColStride = kWavesPerBlock * CPW;
    constexpr int kPol = ST == 2 ? kPolicySc1 : 0;
    static_assert(U * kColStride <= kLeanColSlack, "batch reaches past the column slack");
    const int N = J.n_rows, M = J.n_cols, ld = J.ld;
    const int nstep = J.nstep;
    const float fM = (float)M, fN = (float)N;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_uniform(threadIdx.x / kWave);
    const int cs = lane / LPC;
    const int rl = lane % LPC;
    using IoV = LaneIo<VEC>;
    constexpr int NCH = IoV::NCH;
    int rows[VEC];  // problem row of each of the lane's elements
#pragma unroll
    for (int e = 0; e < VEC; e++) rows[e] = J.rt * TR + tile_row<VEC, LPC>(rl, e);
    const int tile_col0 = J.ct * J.tc;
    const int tile_cols = min(J.tc, M - tile_col0);
    const int jt0 = w * CPW + cs;  // tile column of step t: jt0 + t * kColStride
    const bool explicit_phi = (ctl.k == 0);
    const float *hat2 = explicit_phi ? J.phi2 : J.bhat;
    // the tile holds rows past n_rows (padding of a batch column, or none at
    // all); a scalar int, so that its branch is a scalar one
    const int pad = __builtin_amdgcn_readfirstlane((J.rt + 1) * TR > N ? 1 : 0);

    const size_t col_base = (size_t)tile_col0 * (size_t)ld;
    const unsigned tile_bytes = (unsigned)tile_cols * (unsigned)ld * 4u;
    const __amdgpu_buffer_rsrc_t xr = buffer_over(J.x + col_base, tile_bytes);
    const __amdgpu_buffer_rsrc_t cr = buffer_over(J.c + col_base, tile_bytes);
    const __amdgpu_buffer_rsrc_t hr = buffer_over(hat2 + tile_col0, (unsigned)tile_cols * 4u);
    const size_t part_base = (size_t)J.rt * M + tile_col0;
    const __amdgpu_buffer_rsrc_t pr = buffer_over(J.pcol + part_base, (unsigned)tile_cols * 8u);
    const __amdgpu_buffer_rsrc_t zr =
        ZS ? buffer_over(J.zout + part_base, (unsigned)tile_cols) : pr;
    unsigned x_lane[NCH];  // a chunk starting past n_rows reads zeros, stores nothing
#pragma unroll
    for (int h = 0; h < NCH; h++) {
        const int r0 = rows[h * IoV::CH];
        x_lane[h] = r0 < N ? ((unsigned)jt0 * (unsigned)ld + (unsigned)r0) * 4u : kOob;
    }
    const unsigned x_step = (unsigned)kColStride * (unsigned)ld * 4u;
    const unsigned h_lane = (unsigned)jt0 * 4u;
    const unsigned p_lane = rl == 0 ? (unsigned)jt0 * 8u : kOob;
    const unsigned z_lane = rl == 0 ? (unsigned)jt0 : kOob;
    const int col_lim = tile_cols - jt0;  // step t is a column of the problem iff t * kColStride < col_lim

    // the prologue's first loads go out first: waiting for the flags (needed
    // to form the first batch) then brings them too, and the prologue's sums
    // run while the first batch is in flight
    PrologueLoads pre;
    prologue_issue(ctl, pre);
    __builtin_amdgcn_sched_barrier(0);

    // zero flags of this column group's segments from the previous pass, bit t
    // for column step t (see qr_tile_generic).  Lane rl of the column group
    // loads the flag of step rl and, in 16-byte tiles (up to 2 LPC steps), of
    // step rl + LPC as well (slot q: step rl + q LPC).
    constexpr int kZSlots = VEC >= 4 ? 2 : 1;
    using ZMask = std::conditional_t<(kZSlots * LPC > 32), unsigned long long, unsigned>;
    constexpr int kZBits = 8 * (int)sizeof(ZMask);
    ZMask zmask = 0;
    bool zb[kZSlots];
    // screening data: slot q of lane rl holds step rl + q LPC's headroom and
    // column potential
    float e_rel[kZSlots], h_scr[kZSlots];
#pragma unroll
    for (int q = 0; q < kZSlots; q++) {
        zb[q] = false;
        e_rel[q] = INFINITY;
        h_scr[q] = 0.0f;
    }
    // The loads at the start of a tile are buffer loads whose offset is
    // pushed out of range for a lane with nothing to read (the hardware
    // returns 0): a load under a lane condition becomes a branch that waits
    // for it, and for every load issued before it, on the spot, which
    // serialised the tile's start behind its first batch.
    if (ZS) {
        const __amdgpu_buffer_rsrc_t zir = buffer_over(J.zin + part_base, (unsigned)tile_cols);
        const __amdgpu_buffer_rsrc_t eir =
            SC ? buffer_over(J.seg_e + part_base, (unsigned)tile_cols * 4u) : zir;
        bool zok[kZSlots];
        unsigned char zv[kZSlots];
#pragma unroll
        for (int q = 0; q < kZSlots; q++) {
            const int ts = rl + q * LPC;
            const int jc = jt0 + ts * kColStride;  // column within the tile
            zok[q] = ts < nstep && jc < tile_cols;
            const unsigned zo = zok[q] ? (unsigned)jc : kOob;
            const unsigned zo4 = zok[q] ? (unsigned)jc * 4u : kOob;
            zv[q] = __builtin_amdgcn_raw_buffer_load_b8(zir, (int)zo, 0, 0);
            if constexpr (SC) {
                e_rel[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(eir, (int)zo4, 0, 0));
                h_scr[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(hr, (int)zo4, 0, 0));
            }
        }
#pragma unroll
        for (int q = 0; q < kZSlots; q++) {
            zb[q] = zok[q] && zv[q] != 0;
            if constexpr (SC) e_rel[q] = zok[q] ? e_rel[q] : INFINITY;
            zmask |= (ZMask)group_bits<LPC>(__builtin_amdgcn_ballot_w64(zb[q]), cs) << (q * LPC);
        }
    }
    static_assert(!SC || ZS, "screening needs the zero-segment flags");
    double g1_old = 0.0;
    float f1p[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) f1p[e] = 0.0f;
    if constexpr (SC) {
        g1_old = *J.g1;
        if (ctl.k > 0) {
            const __amdgpu_buffer_rsrc_t fpr = buffer_over(J.f1_prev, (unsigned)N * 4u);
#pragma unroll
            for (int e = 0; e < VEC; e++)  // rows past n_rows read 0
                f1p[e] = __uint_as_float(
                    __builtin_amdgcn_raw_buffer_load_b32(fpr, (int)((unsigned)rows[e] * 4u), 0, 0));
        }
    }
    ZMask smask = 0;  // screened steps of this column group (after the prologue)

    float xa[U][VEC], ca[U][VEC], ha[U];
    float xb[U][VEC], cb[U][VEC], hb[U];
    // always inlined: the batches are register arrays (a call would put them in scratch)
    auto load_batch = [&](int t0, float (&xv)[U][VEC], float (&cv)[U][VEC],
                          float (&hv)[U]) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u;  // steps past nstep read columns past the tile: zeros
            unsigned off[NCH];
#pragma unroll
            for (int h = 0; h < NCH; h++) off[h] = x_lane[h] + (unsigned)t * x_step;
            hv[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                hr, (int)(h_lane + (unsigned)(t * kColStride) * 4u), 0, 0));
            if constexpr (ZS) {
                const bool z = (zmask >> (t & (kZBits - 1))) & 1;
                const bool sk = SC && ((smask >> (t & (kZBits - 1))) & 1);
                // a flagged segment: the even pass reads no X (+0), the odd
                // pass reads C (and X as +0) to form what the even pass left;
                // a screened one reads nothing
                unsigned xoff[NCH], coff[NCH], eoff[NCH];
#pragma unroll
                for (int h = 0; h < NCH; h++) {
                    xoff[h] = z ? kOob : off[h];
                    coff[h] = (z && !sk) ? off[h] : kOob;
                    eoff[h] = sk ? kOob : off[h];
                }
                if (EVEN) IoV::load(cr, eoff, cv[u]);
                IoV::load(xr, xoff, xv[u]);
                if (!EVEN) IoV::load(cr, coff, cv[u]);
            } else {
                IoV::load(xr, off, xv[u]);
                if (EVEN) IoV::load(cr, off, cv[u]);
            }
        }
    };
    load_batch(0, xa, ca, ha);
    float h1[VEC];
    const __amdgpu_buffer_rsrc_t h1r =
        buffer_over(explicit_phi ? J.phi1 : J.ahat, (unsigned)N * 4u);
#pragma unroll
    for (int e = 0; e < VEC; e++)  // rows past n_rows read 0
        h1[e] = __uint_as_float(
            __builtin_amdgcn_raw_buffer_load_b32(h1r, (int)((unsigned)rows[e] * 4u), 0, 0));

    const IterStart is = iteration_prologue(ctl, s_red, leader_block, &pre);
    if (is.finish) return;
    const float gamma = is.gamma;
    const float s = J.step, kappa = J.kappa;
    const float rkappa = 1.0f / kappa;
    const float g2 = explicit_phi ? 0.0f : gamma;
    const float d2 = explicit_phi ? 1.0f : fN;
    const float rd2 = 1.0f / d2;
    constexpr unsigned kQLo = 0x12800000u;  // bits of 2^-90
    constexpr unsigned kQHi = 0x7B800000u;  // bits of 2^120
    // conservative screen for div_kappa_slow: y in [2^-88 kappa, 2^118 kappa]
    // gives a quotient in [2^-90, 2^120]; both bounds are exact for kappa in
    // [2^-30, 2^8], and outside that every positive y goes to the exact test
    const bool screen = kappa >= 0x1p-30f && kappa <= 0x1p+8f;
    const unsigned ylo_m1 = screen ? __float_as_uint(0x1p-88f * kappa) - 1u : 0xFFFFFFFEu;
    const unsigned yhi = screen ? __float_as_uint(0x1p+118f * kappa) : 0u;

    float f1[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++)
        f1[e] = rows[e] < N ? (explicit_phi ? h1[e] : (h1[e] + gamma) / fM) : -INFINITY;
    double racc[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) racc[e] = 0.0;
    double oacc = 0.0;

    // screening: G1 += the largest rise of the tile's f1 since the previous
    // pass (exact differences in fp64; every column group holds the tile's
    // rows, so every lane ends with the same sum), then the screen of step rl
    // of this lane's column group
    double g1_new = 0.0;
    if constexpr (SC) {
        // (the rise of a row is exact in fp64 and rounded up to fp32)
        float rise = -INFINITY, f1abs = 0.0f;
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            if (rows[e] < N) {
                rise = fmaxf(rise, __double2float_ru((double)f1[e] - (double)f1p[e]));
                f1abs = fmaxf(f1abs, fabsf(f1[e]));
            }
        }
        rise = seg_max<LPC>(rise);
        f1abs = seg_max<LPC>(f1abs);
        g1_new = ctl.k > 0 ? g1_old + (double)rise : 0.0;
        if (threadIdx.x == 0) *J.g1 = g1_new;
        if (J.ct == 0 && w == 0 && cs == 0) {
#pragma unroll
            for (int e = 0; e < VEC; e++)
                if (rows[e] < N) J.f1_out[rows[e]] = f1[e];
        }
#pragma unroll
        for (int q = 0; q < kZSlots; q++) {
            // f2 of the slot's step within an ulp or two of the pass's own
            // (the bound's slack covers it); an unset headroom (+inf) never
            // screens
            const float f2s = explicit_phi ? h_scr[q] : (h_scr[q] + gamma) / fN;
            const double bnd = (double)e_rel[q] + g1_new + (double)f2s;
            const double slack = 0x1p-21 * ((double)f1abs + fabs((double)f2s)) +
                                 0x1p-40 * (fabs((double)e_rel[q]) + fabs(g1_new)) + 0x1p-120;
            // a step past the problem's columns counts as screened as well:
            // its loads read nothing and its stores are dropped either way
            const int ts = rl + q * LPC;
            const bool past = ts < nstep && tile_col0 + jt0 + ts * kColStride >= M;
            const bool sk = (zb[q] && ctl.k > 0 && bnd + slack < 0.0) || past;
            smask |= (ZMask)group_bits<LPC>(__builtin_amdgcn_ballot_w64(sk), cs) << (q * LPC);
        }
    }
    // headroom of a segment whose C this pass reads, stored by lane 0 of the
    // column group: max_i RN(f1_i - s c_ij) = R bounds the exact maximum by
    // R + 2^-23 |R|, kept relative to G1 and rounded up to fp32.  Refreshed on
    // every eighth pass (an even one, which reads the C of every segment it
    // does not screen); a headroom keeps its validity as G1 grows, and with
    // this interval the simulated screen still certifies 99.8% of the flagged
    // segments at 1000^2, as it does when refreshing every pass.
    const __amdgpu_buffer_rsrc_t er =
        SC ? buffer_over(J.seg_e + part_base, (unsigned)tile_cols * 4u) : pr;
    const unsigned e_lane = rl == 0 ? (unsigned)jt0 * 4u : kOob;
    const bool e_refresh = (ctl.k & 7) == 0;

    auto run_batch = [&](int t0, float (&xv)[U][VEC], float (&cv)[U][VEC],
                         float (&hv)[U]) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u;
            if (t >= nstep) break;  // wave-uniform
            if constexpr (SC && U > 1) {
                // every column group of the wave screened at this step: the
                // step's results are known without arithmetic (X' = +0: the
                // flag stays set, the column partial is +0, X is not stored,
                // the row and objective sums gain +0) and it loaded nothing
                if (__builtin_amdgcn_ballot_w64(((smask >> (t & (kZBits - 1))) & 1) == 0) == 0ull) {
                    __builtin_amdgcn_raw_buffer_store_b8(
                        (unsigned char)1, zr, (int)(z_lane + (unsigned)(t * kColStride)), 0, 0);
                    __builtin_amdgcn_raw_buffer_store_b64(
                        (u32x2)(0u), pr, (int)(p_lane + (unsigned)(t * kColStride) * 8u), 0, 0);
                    continue;
                }
            }
            // phi2 = (hat + gamma) / n_rows; at k == 0 the array value itself,
            // as (value + 0) / 1 (a zero's sign aside, which no X' sees)
            const float num = hv[u] + g2;
            float f2 = div_kappa(num, d2, rd2);
            const unsigned qb = __float_as_uint(f2) & 0x7FFFFFFFu;
            // (two ballots of plain compares: their lane masks meet in one
            // scalar and, with no per-lane boolean to materialise)
            if (__builtin_amdgcn_ballot_w64(num != 0.0f) &
                __builtin_amdgcn_ballot_w64(qb - kQLo > kQHi - kQLo))
                if (div_rn_slow(num, f2)) f2 = num / d2;
            // a screened step computes as a column past the problem: every
            // y is -inf, X' = +0, nothing is stored and the flag stays set
            const bool sk = SC && ((smask >> (t & (kZBits - 1))) & 1);
            f2 = (t * kColStride < col_lim && !sk) ? f2 : -INFINITY;
            if (EVEN && pad) {
#pragma unroll
                for (int e = 0; e < VEC; e++)
                    if (rows[e] >= N) cv[u][e] = 0.0f;
            }
            // max(X + phi1 + phi2 - sC, 0): odd pass, flagged segment: X = +0
            // and C read, giving fmaf(-s, c, +0); elsewhere C reads 0 and
            // fmaf(-s, 0, x) == x
            auto y_of = [&](int e) __attribute__((always_inline)) {
                const float xe =
                    (!EVEN && ZS) ? __builtin_fmaf(-s, cv[u][e], xv[u][e]) : xv[u][e];
                const float tt = (xe + f1[e]) + f2;
                return fmaxf(EVEN ? __builtin_fmaf(-s, cv[u][e], tt) : tt, 0.0f);
            };
            float xq[VEC];
            unsigned lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float yv = y_of(e);
                xq[e] = div_kappa(yv, kappa, rkappa);
                const unsigned b = __float_as_uint(yv);
                lo = min(lo, b - 1u);  // +0 maps to the largest value
                hi = max(hi, b);
            }
            if (__builtin_amdgcn_ballot_w64(lo < ylo_m1 || hi > yhi)) {
                // y again from the operands (still live), so that it need not
                // stay in registers for this rare branch
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    const float yv = y_of(e);
                    if (div_kappa_slow(yv, xq[e])) xq[e] = yv / kappa;
                }
            }
            double csum = 0.0;
            unsigned nzb = 0u;
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float xp = xq[e];  // +0 or positive: never -0
                const double d = (double)xp;
                racc[e] += d;
                csum = e == 0 ? d : csum + d;
                // the fp64 product of two floats is exact, so the fma rounds
                // once, as the separate product and sum do
                if (EVEN) oacc = __builtin_fma(d, (double)cv[u][e], oacc);
                nzb |= __float_as_uint(xp);
            }
            unsigned st_off[NCH];
#pragma unroll
            for (int h = 0; h < NCH; h++) st_off[h] = x_lane[h] + (unsigned)t * x_step;
            if constexpr (ZS) {
                const bool zseg =
                    group_bits<LPC>(__builtin_amdgcn_ballot_w64(nzb != 0u), cs) == 0ull;
                __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(zseg ? 1 : 0), zr,
                                                     (int)(z_lane + (unsigned)(t * kColStride)),
                                                     0, 0);
#pragma unroll
                for (int h = 0; h < NCH; h++) st_off[h] = zseg ? kOob : st_off[h];
            }
            float xo[VEC];
#pragma unroll
            for (int e = 0; e < VEC; e++) xo[e] = EVEN ? __builtin_fmaf(-s, cv[u][e], xq[e]) : xq[e];
            IoV::template store<kPol>(xr, st_off, xo);
            if (SC && EVEN && e_refresh) {  // k % 8 == 0: an even pass, C read unless screened
                const bool upd = !sk;
                float r = -INFINITY;
#pragma unroll
                for (int e = 0; e < VEC; e++) r = fmaxf(r, __builtin_fmaf(-s, cv[u][e], f1[e]));
                r = seg_max<LPC>(r);
                const double ev = (double)r + 0x1p-23 * fabs((double)r) - g1_new;
                __builtin_amdgcn_raw_buffer_store_b32(
                    __float_as_uint(__double2float_ru(ev)), er,
                    (int)(upd ? e_lane + (unsigned)(t * kColStride) * 4u : kOob), 0, 0);
            }
            csum = seg_sum<LPC>(csum);
            const unsigned long long cb64 = (unsigned long long)__double_as_longlong(csum);
            u32x2 cw;
            cw.x = (unsigned)cb64;
            cw.y = (unsigned)(cb64 >> 32);
            __builtin_amdgcn_raw_buffer_store_b64(
                cw, pr, (int)(p_lane + (unsigned)(t * kColStride) * 8u), 0, 0);
        }
    };

    if constexpr (SC && U == 1) {
        // steps at which every column group of the wave is screened
        // (wave-uniform) have results known without arithmetic (X' = +0: the
        // flag stays set, the column partial is +0, X is not stored, the row
        // and objective sums gain +0) and load nothing; the loop runs over
        // the other steps only.  (A skip branch inside the step instead kept
        // 26 more registers live and cost a wave per SIMD.)
        unsigned long long all_sk = ~0ull;
#pragma unroll
        for (int g = 0; g < CPW; g++) {
            const unsigned long long m = (unsigned long long)smask;
            const unsigned lo = __builtin_amdgcn_readlane((unsigned)m, g * LPC);
            const unsigned hi = __builtin_amdgcn_readlane((unsigned)(m >> 32), g * LPC);
            all_sk &= ((unsigned long long)hi << 32) | lo;
        }
        const unsigned long long steps = nstep >= 64 ? ~0ull : ((1ull << nstep) - 1ull);
        unsigned long long live = steps & ~all_sk;
        for (unsigned long long m = steps & all_sk; m; m &= m - 1ull) {
            const int t = __builtin_ctzll(m);
            __builtin_amdgcn_raw_buffer_store_b8(
                (unsigned char)1, zr, (int)(z_lane + (unsigned)(t * kColStride)), 0, 0);
            __builtin_amdgcn_raw_buffer_store_b64(
                (u32x2)(0u), pr, (int)(p_lane + (unsigned)(t * kColStride) * 8u), 0, 0);
        }
        if (live) {
            int ta = __builtin_ctzll(live);
            live &= live - 1ull;
            if (ta != 0) load_batch(ta, xa, ca, ha);  // step 0's batch went out before the prologue
            while (true) {
                if (J.prio) set_issue_priority(__builtin_popcountll(live) + 1, nstep);
                const int tb = live ? __builtin_ctzll(live) : nstep;
                live &= live - 1ull;
                if (tb < nstep) load_batch(tb, xb, cb, hb);
                run_batch(ta, xa, ca, ha);
                if (tb >= nstep) break;
                if (J.prio) set_issue_priority(__builtin_popcountll(live) + 1, nstep);
                ta = live ? __builtin_ctzll(live) : nstep;
                live &= live - 1ull;
                if (ta < nstep) load_batch(ta, xa, ca, ha);
                run_batch(tb, xb, cb, hb);
                if (ta >= nstep) break;
            }
        }
    } else {
        for (int t0 = 0; t0 < nstep; t0 += 2 * U) {
            if (J.prio) set_issue_priority(nstep - t0, nstep);
            if (t0 + U < nstep) load_batch(t0 + U, xb, cb, hb);
            run_batch(t0, xa, ca, ha);
            if (t0 + U >= nstep) break;
            if (J.prio) set_issue_priority(nstep - t0 - U, nstep);
            if (t0 + 2 * U < nstep) load_batch(t0 + 2 * U, xa, ca, ha);
            run_batch(t0 + U, xb, cb, hb);
        }
    }
    tile_partials_out<VEC, LPC, EVEN>(J, racc, oacc, s_row);
}

// The masked formulation, for any leading dimension: chunks cut by the column
// end are read and written element by element.  ZS: zero-segment flags in use
// (J.zin and J.zout set); without them the loads are the plain ones below.
template <int VEC, int LPC, bool EVEN, int ST, int U, bool ZS>
__device__ __forceinline__ void qr_tile_generic(const QrTileJob &J, const IterCtl &ctl,
                                                bool leader_block, double (*s_row)[LPC * VEC],
                                                double *s_red) {
    constexpr int CPW = kWave / LPC;  // columns per wave step
    constexpr int TR = LPC * VEC;     // tile rows
    constexpr int kColStride = kWavesPerBlock * CPW;
    const int N = J.n_rows, M = J.n_cols, ld = J.ld;
    const int nstep = J.nstep;
    const float fM = (float)M, fN = (float)N;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_uniform(threadIdx.x / kWave);
    const int cs = lane / LPC;
    const int rl = lane % LPC;
    const int row = J.rt * TR + rl * VEC;
    const bool row_ok = row < N;
    bool ev_ok[VEC];  // element rows inside the problem (padding rows excluded)
#pragma unroll
    for (int e = 0; e < VEC; e++) ev_ok[e] = row + e < N;
    // the whole chunk lies in the column (always when ld % VEC == 0: rows
    // N..ld-1 are then padding of the same column)
    const bool row_full = row + VEC <= ld;
    const int tile_col0 = J.ct * J.tc;
    const int col0 = tile_col0 + w * CPW + cs;  // column of step t: col0 + t * kColStride
    const bool explicit_phi = (ctl.k == 0);
    const float *hat2 = explicit_phi ? J.phi2 : J.bhat;

    // the prologue's first entries go out before everything else (see qr_tile_lean)
    PrologueLoads pre;
    prologue_issue(ctl, pre);
    __builtin_amdgcn_sched_barrier(0);

    // flags of this column group's segments from the previous pass: bit t is
    // column step t (loaded one per lane, gathered by a ballot).  A 64-lane
    // column group can hold up to 64 steps, so its mask is 64 bits wide.
    constexpr bool zon = ZS;
    using ZMask = std::conditional_t<(LPC > 32), unsigned long long, unsigned>;
    ZMask zmask = 0;
    if (zon) {
        const int jz = col0 + rl * kColStride;
        const bool zb = rl < nstep && jz < M && J.zin[(size_t)J.rt * M + jz] != 0;
        const unsigned long long bz = __builtin_amdgcn_ballot_w64(zb);
        zmask = (ZMask)group_bits<LPC>(bz, cs);
    }

    // the matrix pointers as locals: a select between two members of J
    // would be a load through a selected address, which keeps J in scratch
    const float *const xptr = J.x;
    const float *const cptr = J.c;
    float xa[U][VEC], ca[U][VEC], h2a[U];
    // always inlined: the batches are register arrays (a call would put them in scratch)
    auto load_batch = [&](int t0, float (&xv)[U][VEC], float (&cv)[U][VEC],
                          float (&h2)[U]) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u;
            const int j = col0 + t * kColStride;
            const bool col_ok = (t < nstep) && (j < M);
            h2[u] = col_ok ? hat2[j] : 0.0f;
            if constexpr (!ZS) {
                if (col_ok && row_ok) {
                    const size_t off = (size_t)j * (size_t)ld + (size_t)row;
                    if (row_full) {
                        Io<VEC>::load(J.x + off, xv[u]);
                        if (EVEN) Io<VEC>::load(J.c + off, cv[u]);
                    } else {
#pragma unroll
                        for (int e = 0; e < VEC; e++) {
                            xv[u][e] = ev_ok[e] ? J.x[off + e] : 0.0f;
                            cv[u][e] = (EVEN && ev_ok[e]) ? J.c[off + e] : 0.0f;
                        }
                    }
                    if (!EVEN) {
#pragma unroll
                        for (int e = 0; e < VEC; e++) cv[u][e] = 0.0f;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; e++) { xv[u][e] = 0.0f; cv[u][e] = 0.0f; }
                }
                continue;
            }
            // flagged segment: even reads no X (+0), odd reads C in X's place
            const bool z = (zmask >> t) & 1;
            const bool live = col_ok && row_ok;
            const size_t off = (size_t)j * (size_t)ld + (size_t)row;
            const float *xs = (!EVEN && z) ? cptr : xptr;
            float xt[VEC], ct[VEC];
#pragma unroll
            for (int e = 0; e < VEC; e++) { xt[e] = 0.0f; ct[e] = 0.0f; }
            if (live && row_full) {
                // C first: its load does not wait for the flags
                if (EVEN) Io<VEC>::load(J.c + off, ct);
                if (!(EVEN && z)) Io<VEC>::load(xs + off, xt);
            } else if (live) {
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    if (EVEN && ev_ok[e]) ct[e] = J.c[off + e];
                    if (ev_ok[e] && !(EVEN && z)) xt[e] = xs[off + e];
                }
            }
#pragma unroll
            for (int e = 0; e < VEC; e++) { xv[u][e] = xt[e]; cv[u][e] = ct[e]; }
        }
    };
    load_batch(0, xa, ca, h2a);
    float h1[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++)
        h1[e] = ev_ok[e] ? (explicit_phi ? J.phi1[row + e] : J.ahat[row + e]) : 0.0f;
    // write-through stores address the tile's column range [tile_col0, +tc)
    const int tile_cols = min(J.tc, M - tile_col0);
    __amdgpu_buffer_rsrc_t xrsrc = __builtin_amdgcn_make_buffer_rsrc(
        J.x + (size_t)tile_col0 * (size_t)ld, 0, tile_cols * ld * 4, kBufDword3);

    const IterStart is = iteration_prologue(ctl, s_red, leader_block, &pre);
    if (is.finish) return;
    const float gamma = is.gamma;
    const float s = J.step, kappa = J.kappa;
    const float rkappa = 1.0f / kappa;

    float f1[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) f1[e] = explicit_phi ? h1[e] : (h1[e] + gamma) / fM;
    double racc[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) racc[e] = 0.0;
    double oacc = 0.0;

    for (int t0 = 0; t0 < nstep; t0 += U) {
        if (J.prio) set_issue_priority(nstep - t0, nstep);
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
            if (!EVEN && ((zmask >> t) & 1)) {
                // what the even pass would have stored: X' - sC with X' = +0
#pragma unroll
                for (int e = 0; e < VEC; e++) xa[u][e] = __builtin_fmaf(-s, xa[u][e], 0.0f);
            }
            float xo[VEC], yv[VEC], xq[VEC];
            double csum = 0.0;
            bool slow = false;
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float tt = (xa[u][e] + f1[e]) + f2;
                float y = EVEN ? __builtin_fmaf(-s, ca[u][e], tt) : tt;
                yv[e] = fmaxf(y, 0.0f);
                xq[e] = div_kappa(yv[e], kappa, rkappa);
                slow |= div_kappa_slow(yv[e], xq[e]);
            }
            // true division for quotients outside the range where the
            // correction step is exact (wave-uniform branch, normally skipped)
            if (__builtin_amdgcn_ballot_w64(slow)) {
#pragma unroll
                for (int e = 0; e < VEC; e++)
                    if (div_kappa_slow(yv[e], xq[e])) xq[e] = yv[e] / kappa;
            }
            bool nz = false;
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float xp = xq[e];
                xo[e] = EVEN ? __builtin_fmaf(-s, ca[u][e], xp) : xp;
                if (ok && ev_ok[e]) {
                    racc[e] += (double)xp;
                    csum += (double)xp;
                    if (EVEN) oacc += (double)xp * (double)ca[u][e];
                    nz |= __float_as_uint(xp) != 0u;
                }
            }
            // all of the segment's X' are +0: flag it and store nothing
            bool zseg = false;
            if (zon) {
                const unsigned long long bn = __builtin_amdgcn_ballot_w64(nz);
                zseg = group_bits<LPC>(bn, cs) == 0ull;
                if (rl == 0 && col_ok) J.zout[(size_t)J.rt * M + j] = zseg ? 1 : 0;
            }
            if (zseg) {
                // nothing to store
            } else if (ok && row_full) {
                if (ST == 2) {
                    Io<VEC>::store_wt(xrsrc, ((j - tile_col0) * ld + row) * 4, xo);
                } else {
                    const size_t off = (size_t)j * (size_t)ld + (size_t)row;
                    Io<VEC>::store(J.x + off, xo);
                }
            } else if (ok) {
                const size_t off = (size_t)j * (size_t)ld + (size_t)row;
#pragma unroll
                for (int e = 0; e < VEC; e++)
                    if (ev_ok[e]) J.x[off + e] = xo[e];
            }
            csum = seg_sum<LPC>(csum);
            if (rl == 0 && col_ok) J.pcol[(size_t)J.rt * M + j] = csum;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            h2a[u] = h2b[u];
#pragma unroll
            for (int e = 0; e < VEC; e++) { xa[u][e] = xb[u][e]; ca[u][e] = cb[u][e]; }
        }
    }

    tile_partials_out<VEC, LPC, EVEN>(J, racc, oacc, s_row);
}

// SC (screening) runs in the branch-free form only; the host enables it for
// shapes that satisfy qr_lean_ok.
template <int VEC, int LPC, bool EVEN, int ST, int U, bool ZS, bool SC>
__device__ __forceinline__ void qr_tile(const QrTileJob &J, const IterCtl &ctl, bool leader_block,
                                        double (*s_row)[LPC * VEC], double *s_red) {
    if constexpr (VEC == 8 || SC) {
        // two-chunk slices exist in the lean form only; the host gives such
        // tiles to shapes that satisfy qr_lean_ok (rows a multiple of 4)
        if (qr_lean_ok<VEC>(J))
            qr_tile_lean<VEC, LPC, EVEN, ST, U, ZS, SC>(J, ctl, leader_block, s_row, s_red);
    } else {
        if (qr_lean_ok<VEC>(J)) {  // wave-uniform
            qr_tile_lean<VEC, LPC, EVEN, ST, U, ZS, false>(J, ctl, leader_block, s_row, s_red);
            return;
        }
        qr_tile_generic<VEC, LPC, EVEN, ST, U, ZS>(J, ctl, leader_block, s_row, s_red);
    }
}

}  // namespace drot
