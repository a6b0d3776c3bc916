// Host side of the solver: tile planning, workspace layout, the iteration
// driver and the C ABI declared in include/drot.h.
//
// The iteration driver replaces the reference's host loops
// (core/drot_qr.hpp:190-232, core/drot_glr.hpp:278-353).  Those issue 9-16
// kernels and 3-5 memsets per iteration and block on a D2H copy of aux every
// one or two iterations.  Here one iteration is two launches (fused pass +
// finalize) and the stopping test runs on device: the host enqueues iterations
// in chunks, and every finished chunk is answered by an asynchronous 4-byte
// snapshot of the device `done` flag that the host reads one chunk later, so
// the GPU queue never drains while the host polls.  Launches enqueued after
// convergence return at their first instruction.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/drot.h"
#include "drot_kernels.h"

namespace drot {
namespace {

thread_local std::string g_last_error;

drot_status fail(drot_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

#define DROT_HIP_CHECK(expr)                                                          \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(DROT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
int ceil_div(int a, int b) { return (a + b - 1) / b; }

enum Kind { kQR = 0, kGL = 1 };

// Everything a solve of one shape needs to know about tiles and workspace.
struct Plan {
    int kind = kQR;
    int n_rows = 0, n_cols = 0, n_groups = 1;
    QrTile qt{4, 32};
    GlTile gt{4, 1, 0};
    int n_rt = 1, n_ct = 1, nstep = 1, tc = 1;
    int grid_pass = 1, n_obj = 1;
    int rb = 1, cb = 1;
    int lane_per_group = 0;
    size_t lds = 0;
    size_t off_state = 0, off_phi1 = 0, off_phi2 = 0, off_a = 0, off_b = 0;
    size_t off_ahat = 0, off_bhat = 0, off_prow = 0, off_pcol = 0, off_pobj = 0;
    size_t off_pfin = 0, off_bounds = 0, off_rowgroup = 0, total = 0;
    // GL long-column path: partial group norms [n_seg][n_cols] (fp64) and the
    // per-group segment tables
    size_t off_pnorm = 0, off_gseg = 0, off_gclo = 0;
    int n_seg = 0;
    // QR zero-segment flags: two buffers of zf_bytes ([n_rt][n_cols]) at
    // off_zf, or none (zf_bytes == 0)
    size_t off_zf = 0, zf_bytes = 0;
    // QR screening of flagged segments (drot_qr_tile.h): headrooms
    // [n_rt][n_cols] (fp32), one G1 per tile (fp64), two row-potential buffers
    int scr = 0;
    size_t off_seg_e = 0, off_g1 = 0, off_f1p = 0;
};

// Whether a supported tile may run a shape.  16-byte tiles need 16-byte
// aligned buffers (v4); two-chunk (vec 8) tiles exist only in the branch-free
// form, which needs every 4-row chunk inside one column (n_rows % 4 == 0) and
// a column tile, with the load batches' slack, under 2^31 bytes.
bool qr_tile_fits(const QrTile &t, int n_rows, int nstep, bool v4) {
    if (t.vec == 1) return true;
    if (!v4) return false;
    if (t.vec == 4) return true;
    const long long tc = (long long)nstep * kWavesPerBlock * (kWave / t.lpc);
    return n_rows % 4 == 0 && (tc + 64) * (long long)n_rows * 4 < (1LL << 31);
}

// Optional override of the QR tile, "vec,lpc,nstep[,nt[,u]]" (tuning sweeps only).
bool qr_tile_from_env(int vec_allowed, int n_rows, QrTile *t, int *nstep) {
    const char *s = getenv("DROT_QR_TILE");
    if (!s || !*s) return false;
    int v = 0, l = 0, n = 0, nt = 0, u = 4;
    const int got = sscanf(s, "%d,%d,%d,%d,%d", &v, &l, &n, &nt, &u);
    if (got < 3) return false;
    QrTile cand{v, l, got >= 4 ? nt : 0, got >= 5 ? u : 4};
    if (!qr_tile_supported(cand) || n < 1) return false;
    if (!qr_tile_fits(cand, n_rows, n, vec_allowed != 0)) return false;
    *t = cand;
    *nstep = n;
    return true;
}

void layout(Plan &P) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t N = (size_t)P.n_rows, M = (size_t)P.n_cols;
    P.off_state = take(sizeof(SolverState));
    P.off_phi1 = take(N * 4);
    P.off_phi2 = take(M * 4);
    P.off_a = take(N * 4);
    P.off_b = take(M * 4);
    P.off_ahat = take(N * 4);
    P.off_bhat = take(M * 4);
    P.off_prow = take((size_t)P.n_ct * N * 8);
    P.off_pcol = take((size_t)P.n_rt * M * 8);
    P.off_pobj = take((size_t)P.n_obj * 8);
    P.off_pfin = take((size_t)(P.rb + P.cb) * 24);
    P.off_bounds = take((size_t)(P.n_groups + 1) * 4);
    P.off_rowgroup = take(P.kind == kGL ? N * 4 : 0);
    P.off_zf = take(2 * P.zf_bytes);
    P.off_seg_e = take(P.scr ? (size_t)P.n_rt * M * 4 : 0);
    P.off_g1 = take(P.scr ? (size_t)P.n_rt * P.n_ct * 8 : 0);
    P.off_f1p = take(P.scr ? 2 * N * 4 : 0);
    const bool split = P.kind == kGL && P.gt.split;
    P.off_pnorm = take(split ? (size_t)P.n_seg * M * 8 : 0);
    P.off_gseg = take(split ? (size_t)(P.n_groups + 1) * 4 : 0);
    P.off_gclo = take(split ? (size_t)P.n_groups * 4 : 0);
    P.total = off;
}

// Group row bounds, core/kernel_glr.hpp:79-80 (groups <= 256 rows) and
// :185-186 (larger groups), evaluated in fp32 as there.
std::vector<int> gl_bounds(int n_rows, int n_groups) {
    std::vector<int> hb((size_t)n_groups + 1);
    const bool small = ceil_div(n_rows, n_groups) <= 256;
    const double off = small ? 0.49 : 0.49999;
    const float per = float(n_rows) / float(n_groups);
    for (int g = 0; g <= n_groups; g++) hb[(size_t)g] = (int)((double)(per * float(g)) + off);
    return hb;
}

// Segment tables of the long-column GL path for row chunks of `chunk` rows:
// group g touches chunks gclo[g] .. and owns the partial-norm rows
// gseg[g] .. gseg[g+1]-1, one per chunk (at least one, so an empty group's
// slot stays inside its own range).
// Returns the largest number of groups one chunk touches (the apply kernel
// keeps one scale per such group in LDS, at most `chunk` of them).
int gl_segments(const std::vector<int> &hb, int chunk, std::vector<int> *gseg,
                std::vector<int> *gclo) {
    const int G = (int)hb.size() - 1;
    gseg->assign((size_t)G + 1, 0);
    gclo->assign((size_t)G, 0);
    for (int g = 0; g < G; g++) {
        const int lo = hb[(size_t)g], hi = hb[(size_t)g + 1];
        const int clo = lo / chunk;
        const int chi = hi > lo ? (hi - 1) / chunk : clo;
        (*gclo)[(size_t)g] = clo;
        (*gseg)[(size_t)g + 1] = (*gseg)[(size_t)g] + (chi - clo + 1);
    }
    // groups of chunk c: from the group of its first row to that of its last
    const int n_rows = hb.back();
    int most = 0, g = 0;
    for (int r0 = 0; r0 < n_rows; r0 += chunk) {
        const int r1 = std::min(r0 + chunk, n_rows);
        while (g + 1 < G && hb[(size_t)g + 1] <= r0) g++;
        int g_last = g;
        while (g_last + 1 < G && hb[(size_t)g_last + 1] <= r1 - 1) g_last++;
        most = std::max(most, g_last - g + 1);
    }
    return most;
}

// Matrix size (elements) from which QR solves use zero-segment flags.  Time
// to 1e-4 with / without them on MI355X: 1000^2 to 2048^2 1.01-1.03x,
// 2500^2 0.99x, 3000^2 0.94x, 10000^2 0.69x per iteration.
constexpr long long kZeroSkipMinElems = 6000000;
// Screening of flagged segments (with the flags) of single problems from this
// many elements up, in tiles of one column step per load batch (their loop
// visits only the steps a wave does not screen whole).  MI355X, time to 1e-4
// with and without, two-chunk tiles: 10000^2 198 vs 378 ms, 8000^2 136 vs
// 191, 6000^2 103 vs 133, 5000^2 76 vs 88, 4000^2 66.6 vs 65.8, 3000^2 50.5
// vs 44.8 (C and X in the 256 MB last-level cache: the bound's bookkeeping
// costs more than the C reads it saves).
constexpr long long kScreenMinElems = 20000000;
// The same for a batch's total elements: mixed 256-problem batch forward 734
// vs 814 ms, 64 x 512^2 forward 45 vs 42 ms.
constexpr long long kBatchScreenMinElems = 50000000;

// Plan of a QR solve with a given pass tile; nstep <= 0 picks the column steps.
Plan plan_qr_tile(int n_rows, int n_cols, QrTile t, int nstep) {
    Plan P;
    P.kind = kQR;
    P.n_rows = n_rows;
    P.n_cols = n_cols;
    P.qt = t;
    const int tr = P.qt.lpc * P.qt.vec;
    const int cpw = kWave / P.qt.lpc;
    const int tc_min = kWavesPerBlock * cpw;
    P.n_rt = ceil_div(n_rows, tr);
    if (nstep <= 0) {
        nstep = std::max(1, 4096 / (tr * tc_min));
        while (nstep > 1 && (long long)P.n_rt * ceil_div(n_cols, nstep * tc_min) < 1024) nstep--;
    }
    P.nstep = nstep;
    P.tc = nstep * tc_min;
    P.n_ct = ceil_div(n_cols, P.tc);
    // the write-through store path addresses one column tile with 32-bit offsets
    if ((long long)P.tc * n_rows * 4 >= (1LL << 31)) P.qt.nt = 0;
    P.grid_pass = qr_pass_grid(P.n_rt, P.n_ct);
    P.n_obj = P.n_rt * P.n_ct;
    P.rb = ceil_div(n_rows, kFinRows);
    P.cb = ceil_div(n_cols, kFinRows);
    // zero-segment flags (drot_qr_tile.h) from kZeroSkipMinElems elements up:
    // they save HBM traffic, while smaller matrices are bound by latency,
    // which the flag load at the start of a tile adds to.  A column group
    // gathers its steps' flags one per lane (two in two-chunk tiles), so
    // nstep <= lpc (2 lpc).  DROT_QR_ZSKIP=1 / 0 turns them on / off for
    // every size.
    const char *zs = getenv("DROT_QR_ZSKIP");
    const bool big = (long long)n_rows * n_cols >= kZeroSkipMinElems;
    const bool want = (zs && zs[0] == '1') || (big && !(zs && zs[0] == '0'));
    const bool zskip = nstep <= (P.qt.vec == 8 ? 2 : 1) * P.qt.lpc && want;
    P.zf_bytes = zskip ? (size_t)P.n_rt * (size_t)n_cols : 0;
    // screening of flagged segments, with the flags, in the branch-free tile
    // (rows in whole 16-byte chunks for vec 4 and 8), by default in tiles of
    // one step per batch.  DROT_QR_SCREEN=1 / 0 turns it on / off wherever
    // the flags are on.
    const char *sc = getenv("DROT_QR_SCREEN");
    const bool lean = (P.qt.vec == 1 || n_rows % 4 == 0) &&
                      (long long)(P.tc + 64) * n_rows * 4 < (1LL << 31);
    const bool want_sc = (sc && sc[0] == '1') ||
                         ((long long)n_rows * n_cols >= kScreenMinElems && P.qt.u == 1 &&
                          !(sc && sc[0] == '0'));
    P.scr = (zskip && lean && want_sc) ? 1 : 0;
    layout(P);
    return P;
}

// Tiles chosen by the autotuner, per shape, for the whole process.
struct TunedTile {
    QrTile t;
    int nstep;
    double us;
};
std::mutex g_tune_mu;
std::map<std::pair<int, int>, TunedTile> g_tuned;

bool tuned_tile(int n_rows, int n_cols, bool v4, QrTile *t, int *nstep) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({n_rows, n_cols});
    if (it == g_tuned.end()) return false;
    // e.g. unaligned buffers: the heuristic scalar tile
    if (!qr_tile_fits(it->second.t, n_rows, it->second.nstep, v4)) return false;
    *t = it->second.t;
    *nstep = it->second.nstep;
    return true;
}

// QR tile choice (measured on MI355X at 3000^2 and 10000^2): 16-byte accesses,
// 32 lanes per column for tall matrices, write-through X stores, two batches
// of two column steps in flight, about 4096 elements per block, and at least
// ~1024 blocks when the matrix allows (256 CUs, several blocks resident each).
// The strided-axis partials cost 8 bytes per row per column tile.  A tile set
// by the autotuner for this shape (or DROT_QR_TILE) takes precedence.
// Built-in 16-byte tile for shapes without a tuned entry, the rule the tuned
// MI355X table follows (512^2 .. 10000^2): 16 lanes per column (64-row
// tiles) and the fewest column steps (at most 8) whose grid fits one resident
// round of the even pass (kResidentPass blocks: 5 per CU on 256 CUs), grown
// while tiles hold under kSmallTile elements and the grid keeps kMinGrid
// blocks.  A grid beyond one round even at 8 steps takes two 16-byte chunks
// per lane when the rows allow it: 8 rows per lane, 16 lanes per column
// (128-row tiles) and the fewest of 8, 16, 24 or 32 column steps whose grid
// fits one round of these tiles (kResidentTwoChunk blocks: four per CU), 32
// when none does; write-through stores while C and X fit the 256 MB
// last-level cache and plain ones beyond.  MI355X, time to 1e-4: 4000^2 65.7
// vs 76.2 ms (against 16 lanes, 8 steps, one chunk), 5000^2 75.9 (16 steps;
// 84.0 at 32), 6000^2 86.4 (24 steps; 106.5 at 16, 95.8 at 32), 8000^2 115.9
// (32 steps; 137.1 at 16), 10000^2 166.6 (32 steps; 200.2 at 16).  Other row
// counts beyond two rounds: 32 lanes per column, 8 steps and plain stores.
constexpr long long kResidentPass = 1280;
constexpr long long kResidentTwoChunk = 1024;
constexpr long long kMinGrid = 512;
constexpr int kSmallTile = 4096;

void builtin_tile16(int n_rows, int n_cols, QrTile *t, int *nstep) {
    constexpr int tr = 64, tc_min = kWavesPerBlock * (kWave / 16);
    const long long n_rt = ceil_div(n_rows, tr);
    auto grid = [&](int ns) { return n_rt * ceil_div(n_cols, ns * tc_min); };
    int ns = 1;
    while (ns < 8 && grid(ns) > kResidentPass) ns++;
    const long long cx_bytes = 8LL * n_rows * n_cols;  // C and X
    if (grid(ns) > kResidentPass && n_rows % 4 == 0 && (512LL + 64) * n_rows * 4 < (1LL << 31)) {
        int steps = 8;
        while (steps < 32 &&
               ceil_div(n_rows, 128) * (long long)ceil_div(n_cols, 16 * steps) > kResidentTwoChunk)
            steps += 8;
        *t = QrTile{8, 16, cx_bytes < (200LL << 20) ? 2 : 0, 1};
        *nstep = steps;
        return;
    }
    if (grid(ns) > 2 * kResidentPass) {
        *t = QrTile{4, 32, 0, 2};
        *nstep = 8;
        return;
    }
    while (ns < 8 && tr * tc_min * ns < kSmallTile && grid(ns + 1) >= kMinGrid) ns++;
    *t = QrTile{4, 16, 2, 2};
    *nstep = ns;
}

Plan plan_qr(int n_rows, int n_cols, bool vec4_ok) {
    const bool v4 = vec4_ok;
    QrTile t;
    int nstep = 0;
    if (!qr_tile_from_env(v4, n_rows, &t, &nstep) &&
        !tuned_tile(n_rows, n_cols, v4, &t, &nstep)) {
        if (v4) {
            builtin_tile16(n_rows, n_cols, &t, &nstep);
        } else {
            t = QrTile{1, n_rows >= 256 ? 64 : (n_rows >= 32 ? 32 : 16), 2, 2};
            nstep = 0;
        }
    }
    return plan_qr_tile(n_rows, n_cols, t, nstep);
}

// Candidate tiles of the autotuner for one shape.
std::vector<std::pair<QrTile, int>> tune_candidates(int n_rows, int n_cols, bool v4) {
    std::vector<std::pair<QrTile, int>> out;
    // two-chunk tiles: 8 or 16 lanes per column, one step per batch
    for (int lpc : {8, 16})
        for (int nt : {2, 0})
            for (int nstep : {1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 24, 32}) {
                const QrTile t{8, lpc, nt, 1};
                const int tc = nstep * kWavesPerBlock * (kWave / lpc);
                if (lpc * 8 > 4 * std::max(n_rows, 16) || tc > 4 * std::max(n_cols, 4)) continue;
                if (qr_tile_fits(t, n_rows, nstep, v4)) out.push_back({t, nstep});
            }
    for (int vec : {4, 1}) {
        if (vec == 4 && !v4) continue;
        for (int lpc : {16, 32, 64}) {
            if (lpc * vec > 4 * std::max(n_rows, 16)) continue;  // mostly padding
            for (int nt : {2, 0})
                for (int u : {2, 4})
                    for (int nstep : {1, 2, 3, 4, 5, 6, 7, 8, 12, 16}) {
                        // a 4-step batch only pays with enough steps to fill it
                        if (u == 4 && nstep < 4) continue;
                        const int tc = nstep * kWavesPerBlock * (kWave / lpc);
                        if (tc > 4 * std::max(n_cols, 4)) continue;
                        out.push_back({QrTile{vec, lpc, nt, u}, nstep});
                    }
        }
    }
    return out;
}

constexpr size_t kGlLdsLimit = 64 * 1024 - 256;  // leaves room for static LDS

// GL tile choice: a block owns whole columns (all row groups of a column) so
// the group norms are complete before X is written.
bool plan_gl(int n_rows, int n_cols, int n_groups, bool vec4_ok, Plan *out,
             std::string *why) {
    Plan P;
    P.kind = kGL;
    P.n_rows = n_rows;
    P.n_cols = n_cols;
    P.n_groups = n_groups;
    const bool v4 = vec4_ok && (n_rows % 4 == 0);
    P.gt.vec = v4 ? 4 : 1;
    // one wave per column when a column fits a wave's registers and the
    // per-wave LDS slices fit (DROT_GL_WAVE=0 selects the block-per-column
    // kernel, for tests that compare the two)
    const char *wave_env = getenv("DROT_GL_WAVE");
    const bool wave_ok = !(wave_env && wave_env[0] == '0');
    // about 256 blocks of four waves: row partials cost 8 B per row per block
    const int cpw = std::max(1, ceil_div(n_cols, kWavesPerBlock * 256));
    const int wchunk = kWave * P.gt.vec;
    int wrch = 1;
    while (wrch < 8 && wchunk * wrch < n_rows) wrch *= 2;
    const size_t wave_lds = (size_t)kWavesPerBlock * ((size_t)n_rows + n_groups) * 4;
    const bool wave_fits = wave_ok && wchunk * wrch >= n_rows && wave_lds <= 48 * 1024;
    // columns of up to 2048 rows: threads over rows, four columns per round.
    // With more than two rounds per block the wave-per-column kernel, whose
    // waves drift apart instead of meeting at three barriers per round, is
    // faster where it applies (MI355X, GL iteration: 1500x1000 15.4 -> 14.1
    // us, 1501x999 17.4 -> 14.6, 2048x2048 23.4 -> 21.8, 2048x4000 equal,
    // 1024x8000 37.1 -> 39.6).  DROT_GL_ROWS=0 turns it off, for tests that
    // compare the kernels.
    const char *rows_env = getenv("DROT_GL_ROWS");
    const bool rows_ok = wave_ok && !(rows_env && rows_env[0] == '0') && (cpw <= 2 || !wave_fits);
    const int rchunk = kBlock * P.gt.vec;
    int rrch = 1;
    while (rrch < (v4 ? 2 : 8) && rchunk * rrch < n_rows) rrch *= 2;
    const size_t rows_lds = (size_t)kWavesPerBlock * gl_rows_slice(n_rows, n_groups) * 4;
    if (rows_ok && rchunk * rrch >= n_rows && rows_lds <= 48 * 1024) {
        P.gt.wave = 2;
        P.gt.rch = rrch;
        P.lds = rows_lds;
        P.tc = kWavesPerBlock * cpw;
        P.n_ct = ceil_div(n_cols, P.tc);
        P.n_rt = 1;
        P.grid_pass = P.n_ct;
        P.n_obj = P.n_ct;
        P.lane_per_group = (n_rows / std::max(1, n_groups)) < 64 ? 1 : 0;
        P.rb = ceil_div(n_rows, kFinRows);
        P.cb = ceil_div(n_cols, kFinRows);
        layout(P);
        *out = P;
        return true;
    }
    if (wave_fits) {
        P.gt.wave = 1;
        P.gt.rch = wrch;
        P.lds = wave_lds;
        P.tc = kWavesPerBlock * cpw;
        P.n_ct = ceil_div(n_cols, P.tc);
        P.n_rt = 1;
        P.grid_pass = P.n_ct;
        P.n_obj = P.n_ct;
        P.lane_per_group = (n_rows / std::max(1, n_groups)) < 64 ? 1 : 0;
        P.rb = ceil_div(n_rows, kFinRows);
        P.cb = ceil_div(n_cols, kFinRows);
        layout(P);
        *out = P;
        return true;
    }
    const int maxr = v4 ? 8 : 32;
    int rch = 1;
    while (rch < maxr && kBlock * P.gt.vec * rch < n_rows) rch *= 2;
    const size_t block_lds = (size_t)n_rows * 8 + (size_t)n_groups * 4 + 16;
    // DROT_GL_SPLIT=1 forces the long-column path with 1024-row chunks (tests
    // compare it with the fused kernels on columns of a few chunks)
    const char *split_env = getenv("DROT_GL_SPLIT");
    const bool force_split = split_env && split_env[0] == '1';
    if (force_split || kBlock * P.gt.vec * rch < n_rows || block_lds > kGlLdsLimit) {
        // columns too long for one block: row chunks over blocks, group norms
        // completed between two launches (gl_split_*_kernel)
        P.gt.split = 1;
        P.gt.wave = 0;
        P.gt.rch = force_split ? (v4 ? 1 : 4) : (v4 ? 4 : 16);
        const int chunk = kBlock * P.gt.vec * P.gt.rch;
        P.n_rt = ceil_div(n_rows, chunk);
        // about 1024 blocks over (chunk, column tile) pairs
        P.tc = std::max(1, ceil_div(P.n_rt * n_cols, 1024));
        P.n_ct = ceil_div(n_cols, P.tc);
        P.grid_pass = P.n_rt * P.n_ct;
        P.n_obj = P.grid_pass;
        P.lane_per_group = (n_rows / std::max(1, n_groups)) < 64 ? 1 : 0;
        P.rb = ceil_div(n_rows, kFinRows);
        P.cb = ceil_div(n_cols, kFinRows);
        std::vector<int> gseg, gclo;
        if (gl_segments(gl_bounds(n_rows, n_groups), chunk, &gseg, &gclo) > chunk) {
            *why = "group bounds put more groups than rows into one row chunk";
            return false;
        }
        P.n_seg = gseg.back();
        P.lds = 0;
        layout(P);
        *out = P;
        return true;
    }
    P.gt.rch = rch;
    P.lds = block_lds;
    // long columns: about 512 blocks (two per CU) hide more of each column's
    // load-and-reduce latency than 256, at the cost of twice the row partials
    // for the finalize (3000x3000: 44.4 -> 34.8 us/iteration; for columns
    // short enough for the wave kernel, 256 blocks stay faster)
    P.tc = std::max(1, ceil_div(n_cols, n_rows > 2048 ? 512 : 256));
    P.n_ct = ceil_div(n_cols, P.tc);
    P.n_rt = 1;
    P.grid_pass = P.n_ct;
    P.n_obj = P.n_ct;
    P.lane_per_group = (n_rows / std::max(1, n_groups)) < 64 ? 1 : 0;
    P.rb = ceil_div(n_rows, kFinRows);
    P.cb = ceil_div(n_cols, kFinRows);
    layout(P);
    *out = P;
    return true;
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Per-thread pinned snapshot slots and events for the asynchronous done-flag poll.
struct PollCtx {
    int *pinned = nullptr;   // 2 done flags + a full SolverState copy
    hipEvent_t ev[2] = {nullptr, nullptr};
    int device = -1;
};
thread_local PollCtx g_poll;

drot_status poll_ctx(PollCtx **out) {
    int dev = 0;
    DROT_HIP_CHECK(hipGetDevice(&dev));
    PollCtx &c = g_poll;
    if (c.pinned == nullptr || c.device != dev) {
        // Contexts of a previous device are kept alive (process lifetime).
        c = PollCtx();
        DROT_HIP_CHECK(hipHostMalloc((void **)&c.pinned, 4096, hipHostMallocDefault));
        DROT_HIP_CHECK(hipEventCreateWithFlags(&c.ev[0], hipEventDisableTiming));
        DROT_HIP_CHECK(hipEventCreateWithFlags(&c.ev[1], hipEventDisableTiming));
        c.device = dev;
    }
    *out = &c;
    return DROT_OK;
}

thread_local bool g_timing = false;

// Per-thread pool of timing events (timing mode only).
struct TimingPool {
    std::vector<hipEvent_t> ev;
    int device = -1;
};
thread_local TimingPool g_tpool;

drot_status timing_pool(int pairs) {
    int dev = 0;
    DROT_HIP_CHECK(hipGetDevice(&dev));
    if (g_tpool.device != dev) {
        g_tpool.ev.clear();  // events of another device are kept alive
        g_tpool.device = dev;
    }
    while ((int)g_tpool.ev.size() < 2 * pairs) {
        hipEvent_t e = nullptr;
        DROT_HIP_CHECK(hipEventCreate(&e));
        g_tpool.ev.push_back(e);
    }
    return DROT_OK;
}

struct DevPtrs {
    SolverState *st;
    float *phi1, *phi2, *a, *b, *ahat, *bhat;
    double *prow, *pcol, *pobj, *pfin;
    int *bounds;
    int *row_group;
    unsigned char *zf;  // null when the plan has no zero-segment flags
    float *seg_e;       // screening (null: off)
    double *g1;
    float *f1p;
};

DevPtrs carve(const Plan &P, void *ws) {
    char *base = static_cast<char *>(ws);
    DevPtrs d;
    d.st = reinterpret_cast<SolverState *>(base + P.off_state);
    d.phi1 = reinterpret_cast<float *>(base + P.off_phi1);
    d.phi2 = reinterpret_cast<float *>(base + P.off_phi2);
    d.a = reinterpret_cast<float *>(base + P.off_a);
    d.b = reinterpret_cast<float *>(base + P.off_b);
    d.ahat = reinterpret_cast<float *>(base + P.off_ahat);
    d.bhat = reinterpret_cast<float *>(base + P.off_bhat);
    d.prow = reinterpret_cast<double *>(base + P.off_prow);
    d.pcol = reinterpret_cast<double *>(base + P.off_pcol);
    d.pobj = reinterpret_cast<double *>(base + P.off_pobj);
    d.pfin = reinterpret_cast<double *>(base + P.off_pfin);
    d.bounds = reinterpret_cast<int *>(base + P.off_bounds);
    d.row_group = reinterpret_cast<int *>(base + P.off_rowgroup);
    d.zf = P.zf_bytes ? reinterpret_cast<unsigned char *>(base + P.off_zf) : nullptr;
    d.seg_e = P.scr ? reinterpret_cast<float *>(base + P.off_seg_e) : nullptr;
    d.g1 = P.scr ? reinterpret_cast<double *>(base + P.off_g1) : nullptr;
    d.f1p = P.scr ? reinterpret_cast<float *>(base + P.off_f1p) : nullptr;
    return d;
}

// Warm-start constants, evaluated in fp32 exactly as the reference's host code
// (pytorch/src/qr_drot_cuda_kernel.cu:53-61, core/drot_qr.hpp:322-330).
void warm_constants(int n_rows, int n_cols, float step, InitArgs *ia) {
    const float n = float(n_rows);
    const float m = float(n_cols);
    const float k = float(1.0) * step * n * m - float(2.);
    ia->v_phi1 = (k + float(2)) / n / (m + n);
    ia->v_phi2 = (k + float(2)) / m / (m + n);
    ia->v_a = (k + float(1)) / n;
    ia->v_b = (k + float(1)) / m;
    ia->v_alpha = k + float(1);
}

// Iteration chunk: keep each chunk around a millisecond of GPU work so the host
// stays ahead, and at most 64 iterations so little work is wasted after
// convergence.  Always even, so the QR even/odd kernels keep their phase.
int chunk_iters(const Plan &P) {
    const double elems = (double)P.n_rows * (double)P.n_cols;
    const double us_per_iter = std::max(4.0, elems * 12.0 / 5.0e6);  // ~5 TB/s
    int ch = (int)(1000.0 / us_per_iter);
    ch = std::max(2, std::min(64, ch));
    return ch & ~1;
}

struct SolveIn {
    const float *c, *p, *q;
    float *x;
    float step, kappa, lambda;
    int max_iters;
    float eps;
    int warm;
};

// noalloc: the solve allocates nothing (no pinned snapshot slots, no events,
// no timing buffers): after every chunk the done flag is copied into pageable
// host memory, which returns once the chunk has finished, and timing mode is
// ignored.  The GPU then idles for the host's turnaround once per chunk.
drot_status run(const Plan &P, const SolveIn &in, void *ws, hipStream_t stream,
                drot_result *res, bool noalloc = false) {
    DevPtrs d = carve(P, ws);
    PollCtx *pc = nullptr;
    if (!noalloc) {
        drot_status ps = poll_ctx(&pc);
        if (ps != DROT_OK) return ps;
    }

    InitArgs ia{};
    ia.x = in.x;
    ia.p = in.p;
    ia.q = in.q;
    ia.phi1 = d.phi1;
    ia.phi2 = d.phi2;
    ia.a = d.a;
    ia.b = d.b;
    ia.ahat = d.ahat;
    ia.bhat = d.bhat;
    ia.st = d.st;
    ia.n_rows = P.n_rows;
    ia.n_cols = P.n_cols;
    ia.warm = in.warm;
    ia.zf = d.zf;
    ia.zf_bytes = 2 * P.zf_bytes;
    ia.seg_e = d.seg_e;
    ia.seg_e_count = P.scr ? (size_t)P.n_rt * (size_t)P.n_cols : 0;
    ia.g1 = d.g1;
    ia.g1_count = P.scr ? (size_t)P.n_rt * (size_t)P.n_ct : 0;
    warm_constants(P.n_rows, P.n_cols, in.step, &ia);
    DROT_HIP_CHECK(launch_init(ia, stream));

    IterCtl ctl{};
    ctl.st = d.st;
    ctl.pfin = d.pfin;
    ctl.n_fin = P.rb + P.cb;
    ctl.k = 0;
    ctl.max_iters = in.max_iters;
    ctl.eps = in.eps;
    ctl.stop_rule = (P.kind == kQR) ? kStopEvenStrict : kStopEveryLeq;
    ctl.obj_even_only = (P.kind == kQR) ? 1 : 0;
    ctl.n_rows = P.n_rows;
    ctl.n_cols = P.n_cols;

    QrPassArgs qa{};
    GlPassArgs ga{};
    FinalizeArgs fa{};
    const int grid_fin = P.rb + P.cb;
    if (P.kind == kQR) {
        qa.x = in.x;
        qa.c = in.c;
        qa.phi1 = d.phi1;
        qa.phi2 = d.phi2;
        qa.ahat = d.ahat;
        qa.bhat = d.bhat;
        qa.prow = d.prow;
        qa.pcol = d.pcol;
        qa.pobj = d.pobj;
        qa.n_rows = P.n_rows;
        qa.n_cols = P.n_cols;
        qa.n_rt = P.n_rt;
        qa.n_ct = P.n_ct;
        qa.nstep = P.nstep;
        qa.step = in.step;
        qa.kappa = in.kappa;
        qa.ctl = ctl;
    } else {
        const int G = P.n_groups;
        const std::vector<int> hb = gl_bounds(P.n_rows, G);
        if (hb[0] != 0 || hb[(size_t)G] != P.n_rows)
            return fail(DROT_ERR_UNSUPPORTED, "group bounds do not cover all rows");
        std::vector<int> rg((size_t)P.n_rows);  // group of each row
        for (int g = 0; g < G; g++)
            for (int i = hb[g]; i < hb[g + 1]; i++) rg[(size_t)i] = g;
        DROT_HIP_CHECK(hipMemcpyAsync(d.bounds, hb.data(), sizeof(int) * (G + 1),
                                      hipMemcpyHostToDevice, stream));
        DROT_HIP_CHECK(hipMemcpyAsync(d.row_group, rg.data(), sizeof(int) * rg.size(),
                                      hipMemcpyHostToDevice, stream));
        std::vector<int> gseg, gclo;
        if (P.gt.split) {
            gl_segments(hb, kBlock * P.gt.vec * P.gt.rch, &gseg, &gclo);
            if (gseg.back() != P.n_seg)
                return fail(DROT_ERR_UNSUPPORTED, "group segment count changed");
            char *base = static_cast<char *>(ws);
            ga.pnorm = reinterpret_cast<double *>(base + P.off_pnorm);
            ga.gseg = reinterpret_cast<const int *>(base + P.off_gseg);
            ga.gclo = reinterpret_cast<const int *>(base + P.off_gclo);
            DROT_HIP_CHECK(hipMemcpyAsync(base + P.off_gseg, gseg.data(), sizeof(int) * gseg.size(),
                                          hipMemcpyHostToDevice, stream));
            DROT_HIP_CHECK(hipMemcpyAsync(base + P.off_gclo, gclo.data(), sizeof(int) * gclo.size(),
                                          hipMemcpyHostToDevice, stream));
        }
        DROT_HIP_CHECK(hipStreamSynchronize(stream));
        ga.x = in.x;
        ga.c = in.c;
        ga.phi1 = d.phi1;
        ga.phi2 = d.phi2;
        ga.ahat = d.ahat;
        ga.bhat = d.bhat;
        ga.prow = d.prow;
        ga.pcol = d.pcol;
        ga.pobj = d.pobj;
        ga.bounds = d.bounds;
        ga.row_group = d.row_group;
        ga.n_rows = P.n_rows;
        ga.n_cols = P.n_cols;
        ga.n_groups = P.n_groups;
        ga.n_ct = P.n_ct;
        ga.n_rt = P.n_rt;
        ga.tc = P.tc;
        ga.lane_per_group = P.lane_per_group;
        ga.step = in.step;
        ga.lambda = in.lambda;
        ga.ctl = ctl;
    }
    fa.prow = d.prow;
    fa.pcol = d.pcol;
    fa.pobj = d.pobj;
    fa.p = in.p;
    fa.q = in.q;
    fa.a = d.a;
    fa.b = d.b;
    fa.ahat = d.ahat;
    fa.bhat = d.bhat;
    fa.pfin = d.pfin;
    fa.st = d.st;
    fa.n_rows = P.n_rows;
    fa.n_cols = P.n_cols;
    fa.n_rt = P.n_rt;
    fa.n_ct = P.n_ct;
    fa.n_obj = P.n_obj;
    fa.rb = P.rb;
    fa.cb = P.cb;

    // launch k (< max_iters): pass + finalize of iteration k; the last launch
    // (k == max_iters) is a pass whose prologue only finishes iteration k-1.
    // Counted in 64 bits: max_iters may be INT_MAX.
    const long long n_launch = (long long)in.max_iters + 1;
    const int chunk = chunk_iters(P);
    // Timing mode: events around every pass; pool halves alternate between the
    // (at most two) chunks in flight, and a half is read back before reuse.
    const bool timing = g_timing && !noalloc;
    std::vector<float> pass_ms;  // per executed launch, grown chunk by chunk
    // Device-clock spans (timing mode): launches with k % 8 < 2 (one even, one
    // odd of every eight) are sampled, launch k into slot block
    // 2 (k / 8) + k % 8 of [2 x grid] each, at most kMaxSpanBytes in total;
    // the other launches run uninstrumented.
    unsigned long long *tspan = nullptr;
    struct SpanFree {
        unsigned long long *p;
        ~SpanFree() {
            if (p) (void)hipFree(p);
        }
    } span_guard{nullptr};
    long long n_samples = 0;
    auto span_slot = [&](long long k) -> long long {
        if (k % 8 >= 2) return -1;
        const long long sl = 2 * (k / 8) + k % 8;
        return sl < n_samples ? sl : -1;
    };
    if (timing) {
        drot_status ts = timing_pool(2 * chunk);
        if (ts != DROT_OK) return ts;
        constexpr size_t kMaxSpanBytes = 64u << 20;
        const size_t per = sizeof(unsigned long long) * 2 * (size_t)P.grid_pass;
        const long long cap = (long long)std::max<size_t>(2, kMaxSpanBytes / per);
        const long long want = 2 * (n_launch / 8) + std::min<long long>(2, n_launch % 8);
        n_samples = std::min(cap, want);
        if (n_samples > 0) {
            DROT_HIP_CHECK(hipMalloc((void **)&tspan, per * (size_t)n_samples));
            span_guard.p = tspan;
            DROT_HIP_CHECK(hipMemsetAsync(tspan, 0, per * (size_t)n_samples, stream));
        }
    }
    auto span_ptr = [&](int k) -> unsigned long long * {
        const long long sl = span_slot(k);
        if (!tspan || sl < 0) return nullptr;
        return tspan + (size_t)sl * 2 * (size_t)P.grid_pass;
    };
    int chunk_start[2] = {0, 0}, chunk_n[2] = {0, 0};
    auto harvest = [&](int half) -> drot_status {
        for (int t = 0; t < chunk_n[half]; t++) {
            hipEvent_t e0 = g_tpool.ev[(size_t)(half * chunk + t) * 2];
            hipEvent_t e1 = g_tpool.ev[(size_t)(half * chunk + t) * 2 + 1];
            float ms = 0.0f;
            DROT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            pass_ms[(size_t)(chunk_start[half] + t)] = ms;
        }
        chunk_n[half] = 0;
        return DROT_OK;
    };
    long long launched = 0;
    int slot = 0;
    bool pending[2] = {false, false};
    volatile int *flags = noalloc ? nullptr : pc->pinned;
    while (launched < n_launch) {
        const int n = (int)std::min<long long>(chunk, n_launch - launched);
        chunk_start[slot] = (int)launched;
        chunk_n[slot] = timing ? n : 0;
        if (timing) pass_ms.resize((size_t)(launched + n), 0.0f);
        for (int t = 0; t < n; t++) {
            const int k = (int)(launched + t);
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timing) {
                e0 = g_tpool.ev[(size_t)(slot * chunk + t) * 2];
                e1 = g_tpool.ev[(size_t)(slot * chunk + t) * 2 + 1];
            }
            if (P.kind == kQR) {
                qa.ctl.k = k;
                qa.ctl.tspan = span_ptr(k);
                // pass k reads the flags pass k-1 wrote and writes the others
                qa.zin = d.zf ? d.zf + (size_t)((k + 1) & 1) * P.zf_bytes : nullptr;
                qa.zout = d.zf ? d.zf + (size_t)(k & 1) * P.zf_bytes : nullptr;
                if (d.seg_e) {
                    qa.seg_e = d.seg_e;
                    qa.g1 = d.g1;
                    qa.f1_prev = d.f1p + (size_t)((k + 1) & 1) * P.n_rows;
                    qa.f1_out = d.f1p + (size_t)(k & 1) * P.n_rows;
                }
                DROT_HIP_CHECK(launch_qr_pass(P.qt, (k % 2) == 0, P.grid_pass, qa, stream, e0, e1));
            } else {
                ga.ctl.k = k;
                ga.ctl.tspan = P.gt.split ? nullptr : span_ptr(k);
                DROT_HIP_CHECK(launch_gl_pass(P.gt, P.grid_pass, P.lds, ga, stream, e0, e1));
            }
            if (k < in.max_iters) {
                fa.with_obj = (P.kind == kGL) || ((k % 2) == 0);
                DROT_HIP_CHECK(launch_finalize(grid_fin, fa, stream));
            }
        }
        launched += n;
        if (noalloc) {
            int done = 0;
            DROT_HIP_CHECK(hipMemcpyAsync(&done, &d.st->done, sizeof(int),
                                          hipMemcpyDeviceToHost, stream));
            DROT_HIP_CHECK(hipStreamSynchronize(stream));
            if (done) break;
            continue;
        }
        DROT_HIP_CHECK(hipMemcpyAsync((void *)&flags[slot], &d.st->done, sizeof(int),
                                      hipMemcpyDeviceToHost, stream));
        DROT_HIP_CHECK(hipEventRecord(pc->ev[slot], stream));
        pending[slot] = true;
        const int other = slot ^ 1;
        if (pending[other]) {
            DROT_HIP_CHECK(hipEventSynchronize(pc->ev[other]));
            pending[other] = false;
            if (timing) {
                drot_status hs = harvest(other);
                if (hs != DROT_OK) return hs;
            }
            if (flags[other]) break;
        }
        slot = other;
    }
    if (P.kind == kQR) {
        if (d.zf)
            DROT_HIP_CHECK(launch_qr_unflag(in.x, in.c, in.step, d.zf, P.zf_bytes, P.n_rows,
                                            P.n_cols, P.qt.lpc * P.qt.vec, d.st, stream));
        DROT_HIP_CHECK(launch_fix_x(in.x, in.c, in.step,
                                    (size_t)P.n_rows * (size_t)P.n_cols, d.st, stream));
    }
    SolverState hs_local;
    SolverState *hs = noalloc ? &hs_local : reinterpret_cast<SolverState *>(pc->pinned + 16);
    DROT_HIP_CHECK(hipMemcpyAsync(hs, d.st, sizeof(SolverState), hipMemcpyDeviceToHost,
                                  stream));
    DROT_HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<double> span_us;  // per launch, -1 when not sampled
    if (timing) {
        for (int h = 0; h < 2; h++) {
            drot_status st = harvest(h);
            if (st != DROT_OK) return st;
        }
        span_us.assign(pass_ms.size(), -1.0);
        if (n_samples > 0) {
            const size_t per = 2 * (size_t)P.grid_pass;
            std::vector<unsigned long long> hsp(per * (size_t)n_samples);
            DROT_HIP_CHECK(hipMemcpy(hsp.data(), tspan, sizeof(unsigned long long) * hsp.size(),
                                     hipMemcpyDeviceToHost));
            int dev = 0, khz = 0;
            DROT_HIP_CHECK(hipGetDevice(&dev));
            DROT_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            for (size_t k = 0; k < span_us.size(); k++) {
                const long long sl = span_slot((long long)k);
                if (sl < 0) continue;
                const unsigned long long *b = hsp.data() + per * (size_t)sl;
                unsigned long long t0 = ~0ull, t1 = 0;
                for (size_t e = 0; e < per; e += 2) {
                    if (b[e] == 0) continue;  // block did no work
                    t0 = std::min(t0, b[e]);
                    t1 = std::max(t1, b[e + 1]);
                }
                if (t1 > t0 && khz > 0) span_us[k] = (double)(t1 - t0) * 1000.0 / (double)khz;
            }
        }
    }
    if (res) {
        res->n_iter = hs->done ? hs->n_iter
                               : (P.kind == kQR ? qr_unconverged_n_iter(in.max_iters) : in.max_iters);
        res->converged = hs->converged;
        res->iterations_run = hs->k_run;
        res->objective = (float)hs->obj;
        res->objective_f64 = hs->obj;
        res->residual = hs->res;
        res->run_ms = 0.0f;
        res->prep_ms = 0.0f;
        res->pass_ms[0] = res->pass_ms[1] = 0.0;
        res->pass_ms_events[0] = res->pass_ms_events[1] = 0.0;
        res->pass_count[0] = res->pass_count[1] = 0;
        res->pass_count_events[0] = res->pass_count_events[1] = 0;
        if (timing) {
            for (int k = 0; k < hs->k_run && k < (int)pass_ms.size(); k++) {
                const int par = (P.kind == kQR && (k % 2) == 1) ? 1 : 0;
                if (span_us[(size_t)k] >= 0.0) {
                    res->pass_ms[par] += span_us[(size_t)k] * 1e-3;
                    res->pass_count[par] += 1;
                }
                res->pass_ms_events[par] += pass_ms[(size_t)k];
                res->pass_count_events[par] += 1;
            }
        }
    }
    return DROT_OK;
}

drot_status check_common(const float *c, const float *p, const float *q, int n_rows,
                         int n_cols, const float *x, int max_iters) {
    if (!c || !p || !q || !x) return fail(DROT_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "n_rows and n_cols must be positive");
    if ((long long)n_rows * (long long)n_cols > (1LL << 40))
        return fail(DROT_ERR_UNSUPPORTED, "matrix too large");
    if (max_iters < 0) return fail(DROT_ERR_INVALID_ARGUMENT, "max_iters must be >= 0");
    return DROT_OK;
}

}  // namespace
}  // namespace drot

using namespace drot;

// ====================================================================== C ABI

extern "C" {

int drot_abi_version(void) { return DROT_ABI_VERSION; }

void drot_set_timing(int enable) { g_timing = enable != 0; }

const char *drot_last_error(void) { return g_last_error.c_str(); }

const char *drot_status_string(drot_status s) {
    switch (s) {
        case DROT_OK: return "ok";
        case DROT_ERR_INVALID_ARGUMENT: return "invalid argument";
        case DROT_ERR_WORKSPACE_TOO_SMALL: return "workspace too small";
        case DROT_ERR_UNSUPPORTED: return "unsupported shape";
        case DROT_ERR_HIP: return "HIP runtime error";
        case DROT_ERR_ALLOC: return "allocation failed";
    }
    return "unknown status";
}

drot_status drot_qr_params_from_rho(float rho, float r_weight, int n_rows, int n_cols,
                                    int max_iters, float eps, drot_qr_params *out) {
    if (!out || n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_qr_params_from_rho");
    // pytorch/src/qr_drot_cuda_kernel.cu:38-40
    const float step = rho / (float(n_rows) + float(n_cols));
    out->step_size = step;
    out->scale = 1 + step * r_weight * float(n_rows + n_cols);
    out->r_weight = r_weight;
    out->max_iters = max_iters;
    out->eps = eps;
    out->warm_start = 1;
    return DROT_OK;
}

drot_status drot_qr_params_from_step(float step_size, float r_weight, int n_rows,
                                     int n_cols, int max_iters, float eps,
                                     int warm_start, drot_qr_params *out) {
    if (!out || n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_qr_params_from_step");
    // core/drot_qr.hpp:262
    out->step_size = step_size;
    out->scale = (1 + step_size * float(n_rows + n_cols) * r_weight);
    out->r_weight = r_weight;
    out->max_iters = max_iters;
    out->eps = eps;
    out->warm_start = warm_start ? 1 : 0;
    return DROT_OK;
}

drot_status drot_glr_params_from_rho(float rho, float r_weight, int n_rows, int n_cols,
                                     int n_groups, int max_iters, float eps,
                                     drot_glr_params *out) {
    if (!out || n_rows <= 0 || n_cols <= 0 || n_groups <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_glr_params_from_rho");
    // pytorch/src/glr_drot_cuda_kernel.cu:45-47
    out->step_size = rho / (float(n_rows) + float(n_cols));
    out->lambda = (float)((double)r_weight * sqrt((double)(n_rows / n_groups)));
    out->n_groups = n_groups;
    out->max_iters = max_iters;
    out->eps = eps;
    out->warm_start = 1;
    return DROT_OK;
}

drot_status drot_glr_params_from_step(float step_size, float r_weight, int n_rows,
                                      int n_cols, int n_groups, int max_iters,
                                      float eps, int warm_start, drot_glr_params *out) {
    if (!out || n_rows <= 0 || n_cols <= 0 || n_groups <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_glr_params_from_step");
    // core/drot_glr.hpp:385
    out->step_size = step_size;
    out->lambda = (float)((double)r_weight / sqrt((double)(n_rows / n_groups)));
    out->n_groups = n_groups;
    out->max_iters = max_iters;
    out->eps = eps;
    out->warm_start = warm_start ? 1 : 0;
    return DROT_OK;
}

drot_status drot_qr_workspace_size(int n_rows, int n_cols, size_t *bytes) {
    if (!bytes || n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_qr_workspace_size");
    const Plan a = plan_qr(n_rows, n_cols, true);
    const Plan b = plan_qr(n_rows, n_cols, false);
    *bytes = std::max(a.total, b.total);
    return DROT_OK;
}

drot_status drot_glr_workspace_size(int n_rows, int n_cols, int n_groups, size_t *bytes) {
    if (!bytes || n_rows <= 0 || n_cols <= 0 || n_groups <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_glr_workspace_size");
    Plan a, b;
    std::string why;
    if (!plan_gl(n_rows, n_cols, n_groups, true, &a, &why)) return fail(DROT_ERR_UNSUPPORTED, why);
    if (!plan_gl(n_rows, n_cols, n_groups, false, &b, &why)) return fail(DROT_ERR_UNSUPPORTED, why);
    *bytes = std::max(a.total, b.total);
    return DROT_OK;
}

drot_status drot_qr_solve_f32(const float *c, const float *p, const float *q, int n_rows,
                              int n_cols, const drot_qr_params *prm, float *x,
                              void *workspace, size_t workspace_bytes, void *stream,
                              drot_result *result) {
    drot_status s = check_common(c, p, q, n_rows, n_cols, x, prm ? prm->max_iters : 0);
    if (s != DROT_OK) return s;
    if (!prm) return fail(DROT_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!workspace) return fail(DROT_ERR_INVALID_ARGUMENT, "workspace is NULL");
    if (!(prm->scale > 0.0f)) return fail(DROT_ERR_INVALID_ARGUMENT, "scale must be > 0");
    const Plan P = plan_qr(n_rows, n_cols, aligned16(c) && aligned16(x));
    if (workspace_bytes < P.total)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL,
                    "workspace needs " + std::to_string(P.total) + " bytes");
    SolveIn in{c, p, q, x, prm->step_size, prm->scale, 0.0f, prm->max_iters, prm->eps,
               prm->warm_start};
    return run(P, in, workspace, static_cast<hipStream_t>(stream), result);
}

}  // extern "C"

// ---------------------------------------------------- allocation-free QR solve

// Device scratch of drot_qr_solve_noalloc_f32, in place of the caller's
// workspace: DROT_NOALLOC_SCRATCH_BYTES per device, part of this library's
// code object, so each device holds its own copy from the library's first use
// there until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code it names a host-side stand-in): it passes the device
// address that hipGetSymbolAddress returned on that device.
static __device__ __align__(256) unsigned char scratch[DROT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNoAllocMaxDevices = 64;

// Per device: the address of `scratch` there (looked up on the first call on
// the device, with the device current) and a lock that makes the calls on one
// device take the shared scratch one at a time.
struct NoAllocDevice {
    std::mutex mu;
    void *ptr = nullptr;
};
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

}  // namespace

extern "C" {

drot_status drot_qr_solve_noalloc_f32(const float *c, const float *p, const float *q,
                                      int n_rows, int n_cols, const drot_qr_params *prm,
                                      float *x, void *stream, drot_result *result) {
    drot_status s = check_common(c, p, q, n_rows, n_cols, x, prm ? prm->max_iters : 0);
    if (s != DROT_OK) return s;
    if (!prm) return fail(DROT_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!(prm->scale > 0.0f)) return fail(DROT_ERR_INVALID_ARGUMENT, "scale must be > 0");
    const Plan P = plan_qr(n_rows, n_cols, aligned16(c) && aligned16(x));
    if (P.total > (size_t)DROT_NOALLOC_SCRATCH_BYTES)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL,
                    "a " + std::to_string(n_rows) + "x" + std::to_string(n_cols) +
                        " solve needs " + std::to_string(P.total) +
                        " bytes of scratch; the allocation-free solve holds " +
                        std::to_string((size_t)DROT_NOALLOC_SCRATCH_BYTES) +
                        " (use drot_qr_solve_f32 with a workspace)");
    int dev = 0;
    DROT_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(DROT_ERR_UNSUPPORTED, "device index " + std::to_string(dev) +
                                              " beyond the allocation-free solve's table");
    NoAllocDevice &D = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(D.mu);
    if (!D.ptr) {
        void *ptr = nullptr;
        DROT_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(DROT_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        D.ptr = ptr;
    }
    SolveIn in{c, p, q, x, prm->step_size, prm->scale, 0.0f, prm->max_iters, prm->eps,
               prm->warm_start};
    return run(P, in, D.ptr, static_cast<hipStream_t>(stream), result, true);
}

drot_status drot_glr_solve_f32(const float *c, const float *p, const float *q, int n_rows,
                               int n_cols, const drot_glr_params *prm, float *x,
                               void *workspace, size_t workspace_bytes, void *stream,
                               drot_result *result) {
    drot_status s = check_common(c, p, q, n_rows, n_cols, x, prm ? prm->max_iters : 0);
    if (s != DROT_OK) return s;
    if (!prm) return fail(DROT_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!workspace) return fail(DROT_ERR_INVALID_ARGUMENT, "workspace is NULL");
    if (prm->n_groups <= 0 || prm->n_groups > n_rows)
        return fail(DROT_ERR_INVALID_ARGUMENT, "n_groups must be in [1, n_rows]");
    Plan P;
    std::string why;
    if (!plan_gl(n_rows, n_cols, prm->n_groups, aligned16(c) && aligned16(x), &P, &why))
        return fail(DROT_ERR_UNSUPPORTED, why);
    if (workspace_bytes < P.total)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL,
                    "workspace needs " + std::to_string(P.total) + " bytes");
    SolveIn in{c, p, q, x, prm->step_size, 1.0f, prm->lambda, prm->max_iters, prm->eps,
               prm->warm_start};
    return run(P, in, workspace, static_cast<hipStream_t>(stream), result);
}

drot_status drot_export_duals_f32(int kind, int n_rows, int n_cols, int n_groups,
                                  const void *workspace, size_t workspace_bytes,
                                  float *phi1, float *phi2, float *alpha, void *stream) {
    if (!workspace || n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_export_duals_f32");
    // Offsets of the dual vectors and state do not depend on the tile choice
    // (they precede the partial buffers), so either plan's layout serves.
    Plan P;
    std::string why;
    if (kind == kQR) {
        P = plan_qr(n_rows, n_cols, true);
    } else if (kind == kGL) {
        if (!plan_gl(n_rows, n_cols, n_groups, true, &P, &why)) return fail(DROT_ERR_UNSUPPORTED, why);
    } else {
        return fail(DROT_ERR_INVALID_ARGUMENT, "kind must be 0 (QR) or 1 (GL)");
    }
    if (workspace_bytes < P.off_prow)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL, "workspace too small for this shape");
    DevPtrs d = carve(P, const_cast<void *>(workspace));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (phi1 || phi2) {
        float *o1 = phi1, *o2 = phi2;
        if (!o1 || !o2)
            return fail(DROT_ERR_INVALID_ARGUMENT, "phi1 and phi2 must both be given");
        DROT_HIP_CHECK(launch_export_phi(d.ahat, d.bhat, d.phi1, d.phi2, d.st, n_rows,
                                         n_cols, o1, o2, s));
    }
    if (alpha) {
        SolverState hs;
        DROT_HIP_CHECK(hipMemcpyAsync(&hs, d.st, sizeof(hs), hipMemcpyDeviceToHost, s));
        DROT_HIP_CHECK(hipStreamSynchronize(s));
        *alpha = hs.alpha;
    } else {
        DROT_HIP_CHECK(hipStreamSynchronize(s));
    }
    return DROT_OK;
}

void drot_free(void *ptr) { free(ptr); }

}  // extern "C"

// ----------------------------------------------------------------- wrappers

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Shared body of the two host-pointer wrappers: allocate, copy, solve, copy
// back, free -- with the prep / run split of core/drot_qr.hpp:271-366.
template <typename SolveFn>
float *host_wrapper(const float *hc, const float *hp, const float *hq, int n_rows,
                    int n_cols, size_t ws_bytes, float *dur_in_ms, float *prep_dur_in_ms,
                    int *n_iter, float *objective, int return_x, SolveFn solve) {
    if (!hc || !hp || !hq || n_rows <= 0 || n_cols <= 0) {
        fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to wrapper");
        return nullptr;
    }
    const size_t mat = (size_t)n_rows * (size_t)n_cols * sizeof(float);
    DevBuf c, p, q, x, ws;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    float *out = nullptr;
    auto ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess) {
            fail(DROT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
            return false;
        }
        return true;
    };
    if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate") ||
        !ok(hipEventCreate(&e2), "hipEventCreate"))
        goto done;
    if (!ok(hipEventRecord(e0, nullptr), "hipEventRecord")) goto done;
    if (!ok(hipMalloc(&c.p, mat), "hipMalloc") || !ok(hipMalloc(&x.p, mat), "hipMalloc") ||
        !ok(hipMalloc(&p.p, (size_t)n_rows * 4), "hipMalloc") ||
        !ok(hipMalloc(&q.p, (size_t)n_cols * 4), "hipMalloc") ||
        !ok(hipMalloc(&ws.p, ws_bytes), "hipMalloc"))
        goto done;
    if (!ok(hipMemcpy(c.p, hc, mat, hipMemcpyHostToDevice), "hipMemcpy") ||
        !ok(hipMemcpy(p.p, hp, (size_t)n_rows * 4, hipMemcpyHostToDevice), "hipMemcpy") ||
        !ok(hipMemcpy(q.p, hq, (size_t)n_cols * 4, hipMemcpyHostToDevice), "hipMemcpy"))
        goto done;
    if (!ok(hipEventRecord(e1, nullptr), "hipEventRecord")) goto done;
    {
        drot_result r{};
        const drot_status s = solve((const float *)c.p, (const float *)p.p, (const float *)q.p,
                                    (float *)x.p, ws.p, ws_bytes, &r);
        if (s != DROT_OK) goto done;
        if (!ok(hipEventRecord(e2, nullptr), "hipEventRecord") ||
            !ok(hipEventSynchronize(e2), "hipEventSynchronize"))
            goto done;
        float prep = 0.0f, run_ms = 0.0f;
        (void)hipEventElapsedTime(&prep, e0, e1);
        (void)hipEventElapsedTime(&run_ms, e1, e2);
        if (dur_in_ms) *dur_in_ms = run_ms;
        if (prep_dur_in_ms) *prep_dur_in_ms = prep;
        if (n_iter) *n_iter = r.n_iter;
        if (objective) *objective = r.objective;
        if (return_x) {
            out = (float *)malloc(mat);
            if (!out) {
                fail(DROT_ERR_ALLOC, "malloc of the returned plan failed");
                goto done;
            }
            if (!ok(hipMemcpy(out, x.p, mat, hipMemcpyDeviceToHost), "hipMemcpy")) {
                free(out);
                out = nullptr;
                goto done;
            }
        }
        g_last_error.clear();
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e2) (void)hipEventDestroy(e2);
    return out;
}

}  // namespace

extern "C" {

float *drot_qr_wrapper_f32(const float *c, const float *p, const float *q, int n_rows,
                           int n_cols, float step_size, float r_weight, int max_iters,
                           float eps, float *dur_in_ms, float *prep_dur_in_ms, int *n_iter,
                           float *objective, int use_warmup_init, int return_x) {
    drot_qr_params prm;
    if (drot_qr_params_from_step(step_size, r_weight, n_rows, n_cols, max_iters, eps,
                                 use_warmup_init, &prm) != DROT_OK)
        return nullptr;
    size_t ws = 0;
    if (drot_qr_workspace_size(n_rows, n_cols, &ws) != DROT_OK) return nullptr;
    return host_wrapper(c, p, q, n_rows, n_cols, ws, dur_in_ms, prep_dur_in_ms, n_iter,
                        objective, return_x,
                        [&](const float *dc, const float *dp, const float *dq, float *dx,
                            void *dws, size_t wsb, drot_result *r) {
                            return drot_qr_solve_f32(dc, dp, dq, n_rows, n_cols, &prm, dx,
                                                     dws, wsb, nullptr, r);
                        });
}

float *drot_glr_wrapper_f32(const float *c, const float *p, const float *q, int n_rows,
                            int n_cols, int n_groups, float step_size, float r_weight,
                            int max_iters, float eps, float *dur_in_ms,
                            float *prep_dur_in_ms, int *n_iter, float *objective,
                            int use_warmup_init, int return_x) {
    drot_glr_params prm;
    if (drot_glr_params_from_step(step_size, r_weight, n_rows, n_cols, n_groups, max_iters,
                                  eps, use_warmup_init, &prm) != DROT_OK)
        return nullptr;
    size_t ws = 0;
    if (drot_glr_workspace_size(n_rows, n_cols, n_groups, &ws) != DROT_OK) return nullptr;
    return host_wrapper(c, p, q, n_rows, n_cols, ws, dur_in_ms, prep_dur_in_ms, n_iter,
                        objective, return_x,
                        [&](const float *dc, const float *dp, const float *dq, float *dx,
                            void *dws, size_t wsb, drot_result *r) {
                            return drot_glr_solve_f32(dc, dp, dq, n_rows, n_cols, &prm, dx,
                                                      dws, wsb, nullptr, r);
                        });
}

}  // extern "C"

// ====================================================================== batches

namespace drot {
namespace {

struct BatchPlan {
    std::vector<BatchProblem> probs;
    std::vector<BatchTile> tiles;  // all tiles, problem by problem
    std::vector<BatchFin> fins;    // all finalize blocks, problem by problem
    std::vector<int> tile_begin, fin_begin;  // per problem, into tiles / fins
    long long row_len = 0, col_len = 0, prow_len = 0, pcol_len = 0;
    int max_iters = 0;
    size_t off_state = 0, off_vec = 0, off_prow = 0, off_pcol = 0, off_pobj = 0, off_pfin = 0;
    size_t off_probs = 0, off_tiles = 0, off_fins = 0, total = 0;
    long long zf_len = 0;  // bytes per zero-segment flag buffer (0: no flags)
    size_t off_zf = 0;
    // screening of flagged segments (with the flags): zf_len fp32 headrooms,
    // one fp64 G1 per tile, two row-potential buffers of row_len floats
    int scr = 0;
    size_t off_seg_e = 0, off_g1 = 0, off_f1p = 0;
};

drot_status plan_batch(int n, const drot_batch_problem *in, BatchPlan *out) {
    if (n <= 0 || !in) return fail(DROT_ERR_INVALID_ARGUMENT, "empty batch");
    BatchPlan B;
    constexpr int cpw = kWave / kBatchLpc;
    constexpr int tc_min = kWavesPerBlock * cpw;
    int tile_off = 0, fin_off = 0;
    long long prow_off = 0, pcol_off = 0, zf_off = 0;
    // zero-segment flags for every problem (the batch keeps the HBM busy, so
    // skipping settled segments pays at all sizes); DROT_QR_ZSKIP=0 turns them
    // off.  Screening with the flags for batches whose matrices add up to
    // kBatchScreenMinElems elements or more; DROT_QR_SCREEN=1 / 0 turns it on
    // / off.
    const char *zs = getenv("DROT_QR_ZSKIP");
    const bool flags = !(zs && zs[0] == '0');
    const char *sc = getenv("DROT_QR_SCREEN");
    long long elems = 0;
    for (int i = 0; i < n; i++)
        elems += (long long)std::max(in[i].n_rows, 0) * (long long)std::max(in[i].n_cols, 0);
    const bool want_sc =
        (sc && sc[0] == '1') || (elems >= kBatchScreenMinElems && !(sc && sc[0] == '0'));
    B.scr = (flags && want_sc) ? 1 : 0;
    // 256-column tiles: with zero-segment flags a block's cost is mostly its
    // latency chain (flag load, first data, prologue), which longer tiles
    // spread over more columns (mixed 256-problem forward with 4-row lane
    // slices: 1.21 s at 64 columns, 1.04 s at 128, 0.99 s at 256).  With
    // screening most steps of a settled tile are skipped whole, and 512
    // columns (two flags per lane) spread the chain further: mixed forward
    // 685 vs 717 ms; the unscreened 64 x 512^2 forward is faster at 256
    // (41.9 vs 48.4 ms: more blocks).  The width follows the batch's size
    // rather than the screening switch, so that a batch keeps its tiles, and
    // its bits, with screening on or off; DROT_BATCH_TILE_COLS=256 / 512
    // sets it.
    const char *bw = getenv("DROT_BATCH_TILE_COLS");
    const int tc = bw && atoi(bw) == kBatchTileColsScreened ? kBatchTileColsScreened
                   : bw && atoi(bw) == kBatchTileCols       ? kBatchTileCols
                   : elems >= kBatchScreenMinElems          ? kBatchTileColsScreened
                                                            : kBatchTileCols;
    for (int i = 0; i < n; i++) {
        const drot_batch_problem &d = in[i];
        if (d.n_rows <= 0 || d.n_cols <= 0)
            return fail(DROT_ERR_INVALID_ARGUMENT, "problem " + std::to_string(i) + ": bad shape");
        if (d.ld < d.n_rows || d.ld % 4 != 0 || d.mat_offset % 4 != 0 || d.mat_offset < 0 ||
            d.p_offset < 0 || d.q_offset < 0)
            return fail(DROT_ERR_INVALID_ARGUMENT,
                        "problem " + std::to_string(i) +
                            ": ld must be >= n_rows and a multiple of 4, offsets multiples of 4");
        if (!(d.params.scale > 0.0f) || d.params.max_iters < 0)
            return fail(DROT_ERR_INVALID_ARGUMENT, "problem " + std::to_string(i) + ": bad params");
        BatchProblem P{};
        P.mat_off = d.mat_offset;
        P.row_off = d.p_offset;
        P.col_off = d.q_offset;
        P.n_rows = d.n_rows;
        P.n_cols = d.n_cols;
        P.ld = d.ld;
        P.n_rt = ceil_div(d.n_rows, kBatchTileRows);
        // The tile's column range plus the load batches' slack of 64 columns
        // must stay under 2^31 bytes (the branch-free tile's buffer offsets).
        P.nstep = tc / tc_min;
        if ((long long)(tc + 64) * d.ld * 4 >= (1LL << 31))
            return fail(DROT_ERR_UNSUPPORTED, "problem " + std::to_string(i) + ": ld too large");
        P.n_ct = ceil_div(d.n_cols, tc);
        P.prow_off = prow_off;
        P.pcol_off = pcol_off;
        P.zf_off = zf_off;
        prow_off += (long long)P.n_ct * d.n_rows;
        pcol_off += (long long)P.n_rt * d.n_cols;
        zf_off += (long long)P.n_rt * d.n_cols;
        P.tile_off = tile_off;
        // one finalize thread per row or column (finalize_rows_block): the
        // batch's rows have at most n_cols / 256 partials; 32 rows per block
        // with eight threads each gave some 17000 mostly idle blocks per
        // launch for the mixed 256-problem workload, 32 us per finalize
        P.rb = ceil_div(d.n_rows, kBlock);
        P.n_fin = P.rb + ceil_div(d.n_cols, kBlock);
        P.fin_off = fin_off;
        P.max_iters = d.params.max_iters;
        P.step = d.params.step_size;
        P.kappa = d.params.scale;
        P.eps = d.params.eps;
        P.warm = d.params.warm_start ? 1 : 0;
        InitArgs ia{};
        warm_constants(d.n_rows, d.n_cols, d.params.step_size, &ia);
        P.v_phi1 = ia.v_phi1;
        P.v_phi2 = ia.v_phi2;
        P.v_a = ia.v_a;
        P.v_b = ia.v_b;
        P.v_alpha = ia.v_alpha;
        B.tile_begin.push_back((int)B.tiles.size());
        for (int ct = 0; ct < P.n_ct; ct++)
            for (int rt = 0; rt < P.n_rt; rt++) B.tiles.push_back(BatchTile{i, rt, ct, 0});
        B.fin_begin.push_back((int)B.fins.size());
        for (int b = 0; b < P.n_fin; b++) B.fins.push_back(BatchFin{i, b});
        tile_off += P.n_rt * P.n_ct;
        fin_off += P.n_fin;
        B.row_len = std::max(B.row_len, d.p_offset + d.n_rows);
        B.col_len = std::max(B.col_len, d.q_offset + d.n_cols);
        B.max_iters = std::max(B.max_iters, P.max_iters);
        B.probs.push_back(P);
    }
    B.tile_begin.push_back((int)B.tiles.size());
    B.fin_begin.push_back((int)B.fins.size());
    B.prow_len = prow_off;
    B.pcol_len = pcol_off;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    B.off_state = take(sizeof(SolverState) * (size_t)n);
    B.off_vec = take(sizeof(float) * 3 * (size_t)(B.row_len + B.col_len));
    B.off_prow = take(8 * (size_t)B.prow_len);
    B.off_pcol = take(8 * (size_t)B.pcol_len);
    B.off_pobj = take(8 * (size_t)tile_off);
    B.off_pfin = take(24 * (size_t)fin_off);
    B.off_probs = take(sizeof(BatchProblem) * (size_t)n);
    B.off_tiles = take(sizeof(BatchTile) * B.tiles.size());
    B.off_fins = take(sizeof(BatchFin) * B.fins.size());
    B.zf_len = flags ? zf_off : 0;
    B.off_zf = take(2 * (size_t)B.zf_len);
    B.off_seg_e = take(B.scr ? 4 * (size_t)B.zf_len : 0);
    B.off_g1 = take(B.scr ? 8 * (size_t)tile_off : 0);
    B.off_f1p = take(B.scr ? 2 * 4 * (size_t)B.row_len : 0);
    B.total = off;
    *out = std::move(B);
    return DROT_OK;
}

// Pinned staging for batch work-lists and state snapshots (per thread).
struct BatchStaging {
    void *buf[2] = {nullptr, nullptr};
    size_t cap = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int device = -1;
};
thread_local BatchStaging g_bstage;

drot_status batch_staging(size_t bytes, BatchStaging **out) {
    int dev = 0;
    DROT_HIP_CHECK(hipGetDevice(&dev));
    BatchStaging &S = g_bstage;
    if (S.device != dev || S.cap < bytes) {
        // buffers of a previous device or size are kept alive (process lifetime)
        S = BatchStaging();
        const size_t cap = std::max<size_t>(bytes, 1 << 20);
        for (int i = 0; i < 2; i++) {
            DROT_HIP_CHECK(hipHostMalloc(&S.buf[i], cap, hipHostMallocDefault));
            DROT_HIP_CHECK(hipEventCreateWithFlags(&S.ev[i], hipEventDisableTiming));
        }
        S.cap = cap;
        S.device = dev;
    }
    *out = &S;
    return DROT_OK;
}

drot_status run_batch(const BatchPlan &B, const float *c, const float *p, const float *q,
                      float *x, void *ws, hipStream_t stream, drot_result *results) {
    const int n = (int)B.probs.size();
    char *base = static_cast<char *>(ws);
    QrBatchArgs A{};
    A.x = x;
    A.c = c;
    A.p = p;
    A.q = q;
    float *vec = reinterpret_cast<float *>(base + B.off_vec);
    A.phi1 = vec;
    A.a = vec + B.row_len;
    A.ahat = vec + 2 * B.row_len;
    A.phi2 = vec + 3 * B.row_len;
    A.b = A.phi2 + B.col_len;
    A.bhat = A.phi2 + 2 * B.col_len;
    A.prow = reinterpret_cast<double *>(base + B.off_prow);
    A.pcol = reinterpret_cast<double *>(base + B.off_pcol);
    A.pobj = reinterpret_cast<double *>(base + B.off_pobj);
    A.pfin = reinterpret_cast<double *>(base + B.off_pfin);
    A.st = reinterpret_cast<SolverState *>(base + B.off_state);
    A.probs = reinterpret_cast<const BatchProblem *>(base + B.off_probs);
    BatchTile *d_tiles = reinterpret_cast<BatchTile *>(base + B.off_tiles);
    BatchFin *d_fins = reinterpret_cast<BatchFin *>(base + B.off_fins);
    A.tiles = d_tiles;
    A.fins = d_fins;
    A.n_tiles = (int)B.tiles.size();
    A.n_fins = (int)B.fins.size();
    A.n_probs = n;
    A.zf = B.zf_len ? reinterpret_cast<unsigned char *>(base + B.off_zf) : nullptr;
    A.zf_len = B.zf_len;
    A.seg_e = B.scr ? reinterpret_cast<float *>(base + B.off_seg_e) : nullptr;
    A.g1 = B.scr ? reinterpret_cast<double *>(base + B.off_g1) : nullptr;
    A.f1pb = B.scr ? reinterpret_cast<float *>(base + B.off_f1p) : nullptr;
    A.row_len = B.row_len;

    const size_t list_bytes =
        sizeof(BatchTile) * B.tiles.size() + sizeof(BatchFin) * B.fins.size();
    const size_t snap_bytes = sizeof(SolverState) * (size_t)n;
    BatchStaging *S = nullptr;
    drot_status ss =
        batch_staging(std::max(list_bytes, sizeof(BatchProblem) * (size_t)n) + snap_bytes + 4096, &S);
    if (ss != DROT_OK) return ss;

    // problem table and full work-lists (synchronous: once per solve)
    DROT_HIP_CHECK(hipMemcpyAsync(base + B.off_probs, B.probs.data(), sizeof(BatchProblem) * n,
                                  hipMemcpyHostToDevice, stream));
    DROT_HIP_CHECK(hipMemcpyAsync(d_tiles, B.tiles.data(), sizeof(BatchTile) * B.tiles.size(),
                                  hipMemcpyHostToDevice, stream));
    DROT_HIP_CHECK(hipMemcpyAsync(d_fins, B.fins.data(), sizeof(BatchFin) * B.fins.size(),
                                  hipMemcpyHostToDevice, stream));
    DROT_HIP_CHECK(launch_batch_init(A, stream));
    DROT_HIP_CHECK(hipStreamSynchronize(stream));

    // iteration chunks; after each chunk an asynchronous snapshot of the
    // problem states, read one chunk later, drives work-list compaction
    long long active_bytes = 0;
    for (const auto &P : B.probs) active_bytes += (long long)P.n_rows * P.n_cols * 10;
    std::vector<char> active(n, 1);
    int n_active = n;
    const long long n_launch = (long long)B.max_iters + 1;
    long long launched = 0;
    int slot = 0;
    bool pending[2] = {false, false};
    SolverState *snap[2];
    for (int h = 0; h < 2; h++)
        snap[h] = reinterpret_cast<SolverState *>(static_cast<char *>(S->buf[h]) + list_bytes +
                                                  4096 - (list_bytes % 256));
    while (launched < n_launch && n_active > 0) {
        const double us = std::max(6.0, (double)active_bytes / 5.0e6);
        int chunk = std::max(2, std::min(64, (int)(1000.0 / us))) & ~1;
        const int cnt = (int)std::min<long long>(chunk, n_launch - launched);
        for (int t = 0; t < cnt; t++) {
            A.k = (int)(launched + t);
            DROT_HIP_CHECK(launch_batch_pass((A.k % 2) == 0, A, stream));
            if (A.k < B.max_iters) DROT_HIP_CHECK(launch_batch_finalize(A, stream));
        }
        launched += cnt;
        DROT_HIP_CHECK(hipMemcpyAsync(snap[slot], A.st, snap_bytes, hipMemcpyDeviceToHost, stream));
        DROT_HIP_CHECK(hipEventRecord(S->ev[slot], stream));
        pending[slot] = true;
        const int other = slot ^ 1;
        if (pending[other]) {
            DROT_HIP_CHECK(hipEventSynchronize(S->ev[other]));
            pending[other] = false;
            bool changed = false;
            for (int i = 0; i < n; i++) {
                if (active[i] && snap[other][i].done) {
                    active[i] = 0;
                    n_active--;
                    changed = true;
                }
            }
            if (changed && n_active > 0) {
                // compact the work-lists to the active problems
                BatchTile *lt = static_cast<BatchTile *>(S->buf[other]);
                int nt = 0;
                active_bytes = 0;
                for (int i = 0; i < n; i++) {
                    if (!active[i]) continue;
                    for (int t = B.tile_begin[i]; t < B.tile_begin[i + 1]; t++) lt[nt++] = B.tiles[t];
                    active_bytes += (long long)B.probs[i].n_rows * B.probs[i].n_cols * 10;
                }
                BatchFin *lf = reinterpret_cast<BatchFin *>(lt + nt);
                int nf = 0;
                for (int i = 0; i < n; i++) {
                    if (!active[i]) continue;
                    for (int f = B.fin_begin[i]; f < B.fin_begin[i + 1]; f++) lf[nf++] = B.fins[f];
                }
                DROT_HIP_CHECK(hipMemcpyAsync(d_tiles, lt, sizeof(BatchTile) * nt,
                                              hipMemcpyHostToDevice, stream));
                DROT_HIP_CHECK(hipMemcpyAsync(d_fins, lf, sizeof(BatchFin) * nf,
                                              hipMemcpyHostToDevice, stream));
                A.n_tiles = nt;
                A.n_fins = nf;
            }
        }
        // Staging half `other` is written again only after the snapshot of the
        // next chunk in that half has been synchronised, and that snapshot is
        // ordered after this upload on the stream, so no extra wait is needed.
        slot = other;
    }
    DROT_HIP_CHECK(launch_batch_fix_x(A, stream));
    DROT_HIP_CHECK(hipMemcpyAsync(snap[0], A.st, snap_bytes, hipMemcpyDeviceToHost, stream));
    DROT_HIP_CHECK(hipStreamSynchronize(stream));
    if (results) {
        for (int i = 0; i < n; i++) {
            const SolverState &h = snap[0][i];
            drot_result &r = results[i];
            memset(&r, 0, sizeof(r));
            r.n_iter = h.done ? h.n_iter : qr_unconverged_n_iter(B.probs[i].max_iters);
            r.converged = h.converged;
            r.iterations_run = h.k_run;
            r.objective = (float)h.obj;
            r.objective_f64 = h.obj;
            r.residual = h.res;
        }
    }
    return DROT_OK;
}


// ------------------------------------------------------------ backward plan

struct BwdPlan {
    std::vector<BwdProblem> probs;
    std::vector<BwdTile> tiles;
    size_t off_probs = 0, off_tiles = 0, off_state = 0, off_tcnt = 0, off_iptr = 0;
    size_t off_idx = 0, off_vec = 0, total = 0;
};

// Default index capacity: the whole matrix up to 2^22 elements, beyond that
// max(2^22, 64 (n_rows + n_cols)) entries (supports of QR plans hold a few
// entries per row; a denser support reports DROT_ERR_WORKSPACE_TOO_SMALL and
// drot_backward_params.nnz_capacity raises the limit).
long long bwd_capacity(const drot_batch_problem &d, const drot_backward_params *bp) {
    const long long mn = (long long)d.n_rows * d.n_cols;
    long long cap;
    if (bp && bp->nnz_capacity > 0) cap = std::min<long long>(bp->nnz_capacity, mn);
    else cap = std::min<long long>(mn, std::max<long long>(1LL << 22, 64LL * (d.n_rows + d.n_cols)));
    return std::min<long long>(cap, 0x7fffffffLL - 1);
}

drot_status plan_bwd(int n, const drot_batch_problem *in, const drot_backward_params *bp,
                     BwdPlan *out) {
    if (n <= 0 || !in) return fail(DROT_ERR_INVALID_ARGUMENT, "empty problem list");
    const double tol = (bp && bp->support_tol > 0.0f) ? (double)bp->support_tol : 0x1p-22;
    const double cg_tol = (bp && bp->cg_tol > 0.0) ? bp->cg_tol : 1e-10;
    BwdPlan B;
    long long iptr = 0, idx = 0, vec = 0;
    for (int i = 0; i < n; i++) {
        const drot_batch_problem &d = in[i];
        const std::string who = "problem " + std::to_string(i);
        if (d.n_rows <= 0 || d.n_cols <= 0) return fail(DROT_ERR_INVALID_ARGUMENT, who + ": bad shape");
        if (d.ld < d.n_rows || d.mat_offset < 0)
            return fail(DROT_ERR_INVALID_ARGUMENT, who + ": ld must be >= n_rows, offsets >= 0");
        if ((long long)d.n_rows + d.n_cols >= (1LL << 30))
            return fail(DROT_ERR_UNSUPPORTED, who + ": too many rows and columns");
        const double step = d.params.step_size, kappa = d.params.scale;
        const double reg = (kappa - 1.0) / step;
        if (!(step > 0.0) || !(reg > 0.0) || !std::isfinite(reg))
            return fail(DROT_ERR_INVALID_ARGUMENT,
                        who + ": the backward needs step_size > 0 and scale > 1 (r_weight > 0)");
        BwdProblem P{};
        P.mat_off = d.mat_offset;
        P.n_rows = d.n_rows;
        P.n_cols = d.n_cols;
        P.ld = d.ld;
        P.cap = bwd_capacity(d, bp);
        P.iptr_off = iptr;
        iptr += (long long)d.n_cols + 1 + 2LL * d.n_rows + 1;
        P.idx_off = idx;
        idx += 2 * P.cap;
        P.vec_off = vec;
        vec += 5LL * (d.n_rows + d.n_cols);
        const long long L = (long long)d.n_rows + d.n_cols;
        P.max_cg = (bp && bp->max_cg_iters > 0) ? bp->max_cg_iters
                                                 : (int)std::min<long long>(4 * L + 100, 1 << 30);
        P.thr = (float)(tol * step);
        P.reg = reg;
        P.cg_tol = cg_tol;
        // column tiles of about 16K elements
        const int cpt = std::max(1, std::min(d.n_cols, 16384 / std::max(1, d.ld)));
        P.tile_begin = (int)B.tiles.size();
        for (int c0 = 0; c0 < d.n_cols; c0 += cpt)
            B.tiles.push_back(BwdTile{i, c0, std::min(d.n_cols, c0 + cpt), 0});
        P.n_tiles = (int)B.tiles.size() - P.tile_begin;
        if (B.tiles.size() >= (size_t)(1u << 31))
            return fail(DROT_ERR_UNSUPPORTED, "too many column tiles");
        B.probs.push_back(P);
    }
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    B.off_probs = take(sizeof(BwdProblem) * B.probs.size());
    B.off_tiles = take(sizeof(BwdTile) * B.tiles.size());
    B.off_state = take(sizeof(BwdState) * B.probs.size());
    B.off_tcnt = take(sizeof(int) * B.tiles.size());
    B.off_iptr = take(sizeof(int) * (size_t)iptr);
    B.off_idx = take(sizeof(int) * (size_t)idx);
    B.off_vec = take(sizeof(double) * (size_t)vec);
    B.total = off;
    *out = std::move(B);
    return DROT_OK;
}

drot_status run_bwd(const BwdPlan &B, const float *c, const float *x, const float *g,
                    const float *gobj, float *grad_c, void *ws, size_t ws_bytes,
                    hipStream_t stream, drot_backward_info *infos) {
    if (!c || !x || !g || !grad_c || !ws) return fail(DROT_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (ws_bytes < B.total)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL, "workspace needs " + std::to_string(B.total) + " bytes");
    char *base = static_cast<char *>(ws);
    const int n = (int)B.probs.size();
    BwdArgs A{};
    A.c = c;
    A.x = x;
    A.g = g;
    A.gobj = gobj;
    A.grad_c = grad_c;
    A.probs = reinterpret_cast<const BwdProblem *>(base + B.off_probs);
    A.tiles = reinterpret_cast<const BwdTile *>(base + B.off_tiles);
    A.st = reinterpret_cast<BwdState *>(base + B.off_state);
    A.tile_cnt = reinterpret_cast<int *>(base + B.off_tcnt);
    A.iptr = reinterpret_cast<int *>(base + B.off_iptr);
    A.idx = reinterpret_cast<int *>(base + B.off_idx);
    A.vec = reinterpret_cast<double *>(base + B.off_vec);
    A.n_tiles = (int)B.tiles.size();
    A.n_probs = n;
    BatchStaging *S = nullptr;
    const size_t tab = sizeof(BwdProblem) * B.probs.size();
    const size_t til = sizeof(BwdTile) * B.tiles.size();
    const size_t stb = sizeof(BwdState) * B.probs.size();
    drot_status ss = batch_staging(align_up(tab, 256) + align_up(til, 256) + stb, &S);
    if (ss != DROT_OK) return ss;
    char *h = static_cast<char *>(S->buf[0]);
    memcpy(h, B.probs.data(), tab);
    memcpy(h + align_up(tab, 256), B.tiles.data(), til);
    BwdState *hs = reinterpret_cast<BwdState *>(h + align_up(tab, 256) + align_up(til, 256));
    DROT_HIP_CHECK(hipMemcpyAsync(base + B.off_probs, h, tab, hipMemcpyHostToDevice, stream));
    DROT_HIP_CHECK(hipMemcpyAsync(base + B.off_tiles, h + align_up(tab, 256), til,
                                  hipMemcpyHostToDevice, stream));
    DROT_HIP_CHECK(launch_bwd_count(A, stream));
    DROT_HIP_CHECK(launch_bwd_fill(A, stream));
    DROT_HIP_CHECK(launch_bwd_cg(A, stream));
    DROT_HIP_CHECK(hipMemcpyAsync(hs, A.st, stb, hipMemcpyDeviceToHost, stream));
    DROT_HIP_CHECK(hipStreamSynchronize(stream));
    int overflow = -1;
    for (int i = 0; i < n; i++) {
        if (hs[i].nnz > B.probs[i].cap && overflow < 0) overflow = i;
        if (infos) {
            drot_backward_info &r = infos[i];
            memset(&r, 0, sizeof(r));
            r.nnz = hs[i].nnz;
            r.cg_iters = hs[i].cg_iters;
            r.converged = hs[i].converged;
            r.cg_residual = hs[i].res;
        }
    }
    if (overflow >= 0)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL,
                    "problem " + std::to_string(overflow) + ": support of " +
                        std::to_string(hs[overflow].nnz) + " entries exceeds the index capacity " +
                        std::to_string(B.probs[overflow].cap) +
                        " (raise drot_backward_params.nnz_capacity)");
    return DROT_OK;
}

drot_batch_problem single_problem(int n_rows, int n_cols, const drot_qr_params *prm) {
    drot_batch_problem d{};
    d.n_rows = n_rows;
    d.n_cols = n_cols;
    d.ld = n_rows;
    d.params = *prm;
    return d;
}


// ---------------------------------------------------------- cost matrices

// Asynchronous upload of small host tables: two pinned buffers used in turn,
// each reused only after the copy that last read it has completed.
struct UploadRing {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    size_t cap = 0;
    int next = 0;
    int device = -1;
};
thread_local UploadRing g_ring;

drot_status upload_async(void *dst, const void *src, size_t bytes, hipStream_t stream) {
    int dev = 0;
    DROT_HIP_CHECK(hipGetDevice(&dev));
    UploadRing &R = g_ring;
    if (R.device != dev || R.cap < bytes) {
        for (int h = 0; h < 2; h++)  // old buffers stay allocated (process lifetime)
            if (R.pending[h]) DROT_HIP_CHECK(hipEventSynchronize(R.ev[h]));
        R = UploadRing();
        R.cap = std::max<size_t>(bytes, 1 << 16);
        for (int h = 0; h < 2; h++) {
            DROT_HIP_CHECK(hipHostMalloc(&R.buf[h], R.cap, hipHostMallocDefault));
            DROT_HIP_CHECK(hipEventCreateWithFlags(&R.ev[h], hipEventDisableTiming));
        }
        R.device = dev;
    }
    const int h = R.next;
    R.next ^= 1;
    if (R.pending[h]) DROT_HIP_CHECK(hipEventSynchronize(R.ev[h]));
    memcpy(R.buf[h], src, bytes);
    DROT_HIP_CHECK(hipMemcpyAsync(dst, R.buf[h], bytes, hipMemcpyHostToDevice, stream));
    DROT_HIP_CHECK(hipEventRecord(R.ev[h], stream));
    R.pending[h] = true;
    return DROT_OK;
}

struct CostPlan {
    std::vector<CostProblem> probs;
    std::vector<CostTile> tiles;
    int max_points = 0;
    size_t off_probs = 0, off_tiles = 0, off_tmax = 0, off_targ = 0, off_xpart = 0;
    size_t off_ypart = 0, off_tdot = 0, table_bytes = 0, total = 0;
};

drot_status plan_cost(int n, const drot_cost_problem *in, int dim, CostPlan *out) {
    if (n <= 0 || !in) return fail(DROT_ERR_INVALID_ARGUMENT, "empty problem list");
    if (dim < 1) return fail(DROT_ERR_INVALID_ARGUMENT, "dim must be >= 1");
    if (dim > kCostMaxDim)
        return fail(DROT_ERR_UNSUPPORTED, "dim " + std::to_string(dim) + " > " +
                                              std::to_string(kCostMaxDim));
    CostPlan B;
    long long xpart = 0, ypart = 0;
    for (int i = 0; i < n; i++) {
        const drot_cost_problem &d = in[i];
        const std::string who = "problem " + std::to_string(i);
        if (d.n_rows <= 0 || d.n_cols <= 0) return fail(DROT_ERR_INVALID_ARGUMENT, who + ": bad shape");
        if (d.ld < d.n_rows || d.mat_offset < 0 || d.src_offset < 0 || d.tgt_offset < 0)
            return fail(DROT_ERR_INVALID_ARGUMENT, who + ": ld must be >= n_rows, offsets >= 0");
        if ((long long)d.n_rows * d.n_cols >= (1LL << 31))
            return fail(DROT_ERR_UNSUPPORTED, who + ": more than 2^31 - 1 elements");
        CostProblem P{};
        P.mat_off = d.mat_offset;
        P.src_off = d.src_offset;
        P.tgt_off = d.tgt_offset;
        P.n_rows = d.n_rows;
        P.n_cols = d.n_cols;
        P.ld = d.ld;
        P.n_it = ceil_div(d.n_rows, kCostTile);
        P.n_jt = ceil_div(d.n_cols, kCostTile);
        P.tile_begin = (int)B.tiles.size();
        P.xpart_off = xpart;
        P.ypart_off = ypart;
        xpart += (long long)P.n_jt * d.n_rows * (dim + 1);
        ypart += (long long)P.n_it * d.n_cols * (dim + 1);
        for (int jt = 0; jt < P.n_jt; jt++)
            for (int it = 0; it < P.n_it; it++) B.tiles.push_back(CostTile{i, it, jt, 0});
        if (B.tiles.size() >= (size_t)(1u << 31)) return fail(DROT_ERR_UNSUPPORTED, "too many tiles");
        B.max_points = std::max(B.max_points, d.n_rows + d.n_cols);
        B.probs.push_back(P);
    }
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    B.off_probs = take(sizeof(CostProblem) * B.probs.size());
    B.off_tiles = take(sizeof(CostTile) * B.tiles.size());
    B.table_bytes = off;
    B.off_tmax = take(sizeof(float) * B.tiles.size());
    B.off_targ = take(sizeof(int) * B.tiles.size());
    B.off_tdot = take(sizeof(double) * B.tiles.size());
    B.off_xpart = take(sizeof(double) * (size_t)xpart);
    B.off_ypart = take(sizeof(double) * (size_t)ypart);
    B.total = off;
    *out = std::move(B);
    return DROT_OK;
}

drot_status cost_args(const CostPlan &B, int dim, void *ws, size_t ws_bytes, hipStream_t stream,
                      CostArgs *A) {
    if (!ws) return fail(DROT_ERR_INVALID_ARGUMENT, "workspace is NULL");
    if (ws_bytes < B.total)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL, "workspace needs " + std::to_string(B.total) + " bytes");
    char *base = static_cast<char *>(ws);
    std::vector<char> tab(B.table_bytes, 0);
    memcpy(tab.data() + B.off_probs, B.probs.data(), sizeof(CostProblem) * B.probs.size());
    memcpy(tab.data() + B.off_tiles, B.tiles.data(), sizeof(CostTile) * B.tiles.size());
    const drot_status st = upload_async(base, tab.data(), B.table_bytes, stream);
    if (st != DROT_OK) return st;
    A->probs = reinterpret_cast<const CostProblem *>(base + B.off_probs);
    A->tiles = reinterpret_cast<const CostTile *>(base + B.off_tiles);
    A->tmax = reinterpret_cast<float *>(base + B.off_tmax);
    A->targ = reinterpret_cast<int *>(base + B.off_targ);
    A->tdot = reinterpret_cast<double *>(base + B.off_tdot);
    A->xpart = reinterpret_cast<double *>(base + B.off_xpart);
    A->ypart = reinterpret_cast<double *>(base + B.off_ypart);
    A->dim = dim;
    A->n_tiles = (int)B.tiles.size();
    A->n_probs = (int)B.probs.size();
    A->max_points = B.max_points;
    return DROT_OK;
}

}  // namespace
}  // namespace drot

extern "C" {

// Leading dimension of drot_batch_layout, in floats: columns start on 128-byte
// lines, so the pass writes whole lines.  With 16-byte alignment only, the
// stores wrote partial lines at every column edge and the mixed-shape batch
// ran about 10% slower (measured with write-through stores).
constexpr int kBatchLdAlign = 32;

drot_status drot_batch_layout(int n_problems, const int *n_rows, const int *n_cols,
                              drot_batch_problem *out, long long *mat_elems,
                              long long *p_elems, long long *q_elems) {
    if (n_problems <= 0 || !n_rows || !n_cols || !out)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_batch_layout");
    long long m = 0, pe = 0, qe = 0;
    for (int i = 0; i < n_problems; i++) {
        if (n_rows[i] <= 0 || n_cols[i] <= 0)
            return fail(DROT_ERR_INVALID_ARGUMENT, "bad shape in drot_batch_layout");
        drot_batch_problem &d = out[i];
        d.n_rows = n_rows[i];
        d.n_cols = n_cols[i];
        d.ld = (n_rows[i] + kBatchLdAlign - 1) / kBatchLdAlign * kBatchLdAlign;
        d.mat_offset = m;
        d.p_offset = pe;
        d.q_offset = qe;
        m += (long long)d.ld * n_cols[i];
        pe += n_rows[i];
        qe += n_cols[i];
    }
    if (mat_elems) *mat_elems = m;
    if (p_elems) *p_elems = pe;
    if (q_elems) *q_elems = qe;
    return DROT_OK;
}

drot_status drot_qr_batch_workspace_size(int n_problems, const drot_batch_problem *problems,
                                         size_t *bytes) {
    if (!bytes) return fail(DROT_ERR_INVALID_ARGUMENT, "bytes is NULL");
    BatchPlan B;
    const drot_status st = plan_batch(n_problems, problems, &B);
    if (st != DROT_OK) return st;
    *bytes = B.total;
    return DROT_OK;
}

drot_status drot_qr_batch_solve_f32(int n_problems, const drot_batch_problem *problems,
                                    const float *c, const float *p, const float *q, float *x,
                                    void *workspace, size_t workspace_bytes, void *stream,
                                    drot_result *results) {
    if (!c || !p || !q || !x || !workspace)
        return fail(DROT_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (((reinterpret_cast<uintptr_t>(c) | reinterpret_cast<uintptr_t>(x)) & 15u) != 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "c and x must be 16-byte aligned");
    BatchPlan B;
    const drot_status st = plan_batch(n_problems, problems, &B);
    if (st != DROT_OK) return st;
    if (workspace_bytes < B.total)
        return fail(DROT_ERR_WORKSPACE_TOO_SMALL,
                    "workspace needs " + std::to_string(B.total) + " bytes");
    return run_batch(B, c, p, q, x, workspace, static_cast<hipStream_t>(stream), results);
}

drot_status drot_qr_backward_workspace_size(int n_rows, int n_cols, const drot_qr_params *params,
                                            const drot_backward_params *bparams, size_t *bytes) {
    if (!bytes || !params) return fail(DROT_ERR_INVALID_ARGUMENT, "null pointer argument");
    const drot_batch_problem d = single_problem(n_rows, n_cols, params);
    BwdPlan B;
    const drot_status st = plan_bwd(1, &d, bparams, &B);
    if (st != DROT_OK) return st;
    *bytes = B.total;
    return DROT_OK;
}

drot_status drot_qr_backward_f32(const float *c, const float *x, const float *grad_x,
                                 const float *grad_obj, int n_rows, int n_cols,
                                 const drot_qr_params *params, const drot_backward_params *bparams,
                                 float *grad_c, void *workspace, size_t workspace_bytes,
                                 void *stream, drot_backward_info *info) {
    if (!params) return fail(DROT_ERR_INVALID_ARGUMENT, "params is NULL");
    const drot_batch_problem d = single_problem(n_rows, n_cols, params);
    BwdPlan B;
    const drot_status st = plan_bwd(1, &d, bparams, &B);
    if (st != DROT_OK) return st;
    return run_bwd(B, c, x, grad_x, grad_obj, grad_c, workspace, workspace_bytes,
                   static_cast<hipStream_t>(stream), info);
}

drot_status drot_qr_batch_backward_workspace_size(int n_problems,
                                                  const drot_batch_problem *problems,
                                                  const drot_backward_params *bparams,
                                                  size_t *bytes) {
    if (!bytes) return fail(DROT_ERR_INVALID_ARGUMENT, "bytes is NULL");
    BwdPlan B;
    const drot_status st = plan_bwd(n_problems, problems, bparams, &B);
    if (st != DROT_OK) return st;
    *bytes = B.total;
    return DROT_OK;
}

drot_status drot_qr_batch_backward_f32(int n_problems, const drot_batch_problem *problems,
                                       const float *c, const float *x, const float *grad_x,
                                       const float *grad_obj, const drot_backward_params *bparams,
                                       float *grad_c, void *workspace, size_t workspace_bytes,
                                       void *stream, drot_backward_info *infos) {
    BwdPlan B;
    const drot_status st = plan_bwd(n_problems, problems, bparams, &B);
    if (st != DROT_OK) return st;
    return run_bwd(B, c, x, grad_x, grad_obj, grad_c, workspace, workspace_bytes,
                   static_cast<hipStream_t>(stream), infos);
}

drot_status drot_sqdist_workspace_size(int n_problems, const drot_cost_problem *problems, int dim,
                                       size_t *bytes) {
    if (!bytes) return fail(DROT_ERR_INVALID_ARGUMENT, "bytes is NULL");
    CostPlan B;
    const drot_status st = plan_cost(n_problems, problems, dim, &B);
    if (st != DROT_OK) return st;
    *bytes = B.total;
    return DROT_OK;
}

drot_status drot_sqdist_f32(int n_problems, const drot_cost_problem *problems, int dim,
                            const float *src, const float *tgt, int normalize, float *c,
                            float *scale, int *argmax, void *workspace, size_t workspace_bytes,
                            void *stream) {
    if (!src || !tgt || !c) return fail(DROT_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (normalize && (!scale || !argmax))
        return fail(DROT_ERR_INVALID_ARGUMENT, "normalize needs the scale and argmax outputs");
    CostPlan B;
    drot_status st = plan_cost(n_problems, problems, dim, &B);
    if (st != DROT_OK) return st;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    CostArgs A{};
    st = cost_args(B, dim, workspace, workspace_bytes, s, &A);
    if (st != DROT_OK) return st;
    A.src = src;
    A.tgt = tgt;
    A.c = c;
    A.scale = scale;
    A.argmax = argmax;
    A.normalize = normalize ? 1 : 0;
    DROT_HIP_CHECK(launch_cost_forward(A, s));
    return DROT_OK;
}

drot_status drot_sqdist_backward_f32(int n_problems, const drot_cost_problem *problems, int dim,
                                     const float *src, const float *tgt, const float *grad_c,
                                     int normalize, const float *scale, const int *argmax,
                                     float *grad_src, float *grad_tgt, void *workspace,
                                     size_t workspace_bytes, void *stream) {
    if (!src || !tgt || !grad_c || !grad_src || !grad_tgt)
        return fail(DROT_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (normalize && (!scale || !argmax))
        return fail(DROT_ERR_INVALID_ARGUMENT, "normalize needs scale and argmax from the forward");
    CostPlan B;
    drot_status st = plan_cost(n_problems, problems, dim, &B);
    if (st != DROT_OK) return st;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    CostArgs A{};
    st = cost_args(B, dim, workspace, workspace_bytes, s, &A);
    if (st != DROT_OK) return st;
    A.src = src;
    A.tgt = tgt;
    A.grad_c = grad_c;
    A.grad_src = grad_src;
    A.grad_tgt = grad_tgt;
    A.scale = const_cast<float *>(scale);
    A.argmax = const_cast<int *>(argmax);
    A.normalize = normalize ? 1 : 0;
    DROT_HIP_CHECK(launch_cost_backward(A, s));
    return DROT_OK;
}

drot_status drot_qr_tile_for(int n_rows, int n_cols, int vec4_ok, drot_qr_tile *out) {
    if (!out || n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_qr_tile_for");
    const Plan P = plan_qr(n_rows, n_cols, vec4_ok != 0);
    out->vec = P.qt.vec;
    out->lanes_per_col = P.qt.lpc;
    out->col_steps = P.nstep;
    out->store_policy = P.qt.nt;
    out->unroll = P.qt.u;
    return DROT_OK;
}

drot_status drot_qr_tile_cache_set(int n_rows, int n_cols, const drot_qr_tile *tile) {
    if (!tile || n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_qr_tile_cache_set");
    const QrTile t{tile->vec, tile->lanes_per_col, tile->store_policy, tile->unroll};
    if (!qr_tile_supported(t) || tile->col_steps < 1 || tile->col_steps > 64 ||
        !qr_tile_fits(t, n_rows, tile->col_steps, true))
        return fail(DROT_ERR_INVALID_ARGUMENT, "unsupported tile");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[{n_rows, n_cols}] = TunedTile{t, tile->col_steps, 0.0};
    return DROT_OK;
}

void drot_qr_tile_cache_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
}

drot_status drot_qr_tile_cache_save(const char *path) {
    if (!path) return fail(DROT_ERR_INVALID_ARGUMENT, "path is NULL");
    std::ofstream f(path);
    if (!f) return fail(DROT_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    f << "# n_rows n_cols vec lanes_per_col col_steps store_policy unroll us_per_iter\n";
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &kv : g_tuned)
        f << kv.first.first << ' ' << kv.first.second << ' ' << kv.second.t.vec << ' '
          << kv.second.t.lpc << ' ' << kv.second.nstep << ' ' << kv.second.t.nt << ' '
          << kv.second.t.u << ' ' << kv.second.us << '\n';
    return f ? DROT_OK : fail(DROT_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
}

drot_status drot_qr_tile_cache_load(const char *path) {
    if (!path) return fail(DROT_ERR_INVALID_ARGUMENT, "path is NULL");
    std::ifstream f(path);
    if (!f) return fail(DROT_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
    std::string line;
    std::map<std::pair<int, int>, TunedTile> got;
    int lineno = 0;
    while (std::getline(f, line)) {
        lineno++;
        if (line.empty() || line[0] == '#') continue;
        int nr, nc, v, l, ns, nt, u;
        double us = 0.0;
        if (sscanf(line.c_str(), "%d %d %d %d %d %d %d %lf", &nr, &nc, &v, &l, &ns, &nt, &u, &us) < 7)
            return fail(DROT_ERR_INVALID_ARGUMENT, std::string(path) + ":" + std::to_string(lineno) +
                                                       ": malformed line");
        const QrTile t{v, l, nt, u};
        // 16-byte tiles serve every row count (a column's partial last chunk
        // is accessed element by element), as drot_qr_tile_cache_set accepts
        if (nr <= 0 || nc <= 0 || !qr_tile_supported(t) || ns < 1 || ns > 64 ||
            !qr_tile_fits(t, nr, ns, true))
            return fail(DROT_ERR_INVALID_ARGUMENT, std::string(path) + ":" + std::to_string(lineno) +
                                                       ": unsupported tile");
        got[{nr, nc}] = TunedTile{t, ns, us};
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &kv : got) g_tuned[kv.first] = kv.second;
    return DROT_OK;
}

drot_status drot_qr_autotune_workspace_size(int n_rows, int n_cols, size_t *bytes) {
    if (!bytes || n_rows <= 0 || n_cols <= 0)
        return fail(DROT_ERR_INVALID_ARGUMENT, "bad arguments to drot_qr_autotune_workspace_size");
    size_t mx = 0;
    for (const auto &cand : tune_candidates(n_rows, n_cols, true))
        mx = std::max(mx, plan_qr_tile(n_rows, n_cols, cand.first, cand.second).total);
    size_t ws = 0;
    drot_qr_workspace_size(n_rows, n_cols, &ws);
    *bytes = std::max(mx, ws);
    return DROT_OK;
}

drot_status drot_qr_autotune_f32(const float *c, const float *p, const float *q, int n_rows,
                                 int n_cols, const drot_qr_params *prm, float *x,
                                 void *workspace, size_t workspace_bytes, void *stream,
                                 int iters, drot_qr_tile *best, double *best_us) {
    drot_status s = check_common(c, p, q, n_rows, n_cols, x, 0);
    if (s != DROT_OK) return s;
    if (!prm || !workspace) return fail(DROT_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (!(prm->scale > 0.0f)) return fail(DROT_ERR_INVALID_ARGUMENT, "scale must be > 0");
    // iterations k_lo..k_hi with k_hi - k_lo sized to a few milliseconds of
    // work, so the difference of the two run times measures those iterations
    // alone.  Both runs share a settling prefix of kSettle iterations: the
    // plan's zero segments (whose X traffic the pass skips) form within the
    // first few dozen iterations, and the rest of a solve runs in that regime.
    constexpr int kSettle = 64;
    const double est_us = 6.0 + (double)n_rows * (double)n_cols * 10.0 / 5.0e6;
    const int k_len = std::max(std::max(8, iters), std::min(4000, (int)(4000.0 / est_us))) & ~1;
    const int k_lo = kSettle + std::max(2, (k_len / 5) & ~1);
    const int k_hi = kSettle + k_len;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const bool v4 = aligned16(c) && aligned16(x);
    double best_t = 1e300;
    std::pair<QrTile, int> arg{};
    bool any = false;
    auto timed = [&](const Plan &P, int k, double *sec) -> drot_status {
        SolveIn in{c, p, q, x, prm->step_size, prm->scale, 0.0f, k, 0.0f, prm->warm_start};
        const auto t0 = std::chrono::steady_clock::now();
        const drot_status r = run(P, in, workspace, st, nullptr);
        *sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return r;
    };
    for (const auto &cand : tune_candidates(n_rows, n_cols, v4)) {
        const Plan P = plan_qr_tile(n_rows, n_cols, cand.first, cand.second);
        if (P.total > workspace_bytes) continue;
        // an untimed run first (a kernel instance's first launch loads its
        // code), then the fastest of three runs at each length: a minimum of
        // per-repetition differences would pick the repetition whose short
        // run was slowed, and underestimate the iteration
        double t_warm = 0.0;
        if ((s = timed(P, k_lo, &t_warm)) != DROT_OK) return s;
        double t_lo = 1e300, t_hi = 1e300;
        for (int rep = 0; rep < 3; rep++) {
            double a = 0.0, b = 0.0;
            if ((s = timed(P, k_lo, &a)) != DROT_OK) return s;
            if ((s = timed(P, k_hi, &b)) != DROT_OK) return s;
            t_lo = std::min(t_lo, a);
            t_hi = std::min(t_hi, b);
        }
        const double d = (t_hi - t_lo) / (double)(k_hi - k_lo);
        const double per = d > 0.0 ? d : 1e300;
        if (per < best_t) {
            best_t = per;
            arg = cand;
            any = true;
        }
    }
    if (!any) return fail(DROT_ERR_WORKSPACE_TOO_SMALL, "no candidate tile fits the workspace");
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[{n_rows, n_cols}] = TunedTile{arg.first, arg.second, best_t * 1e6};
    }
    if (best) {
        const Plan P = plan_qr_tile(n_rows, n_cols, arg.first, arg.second);
        best->vec = P.qt.vec;
        best->lanes_per_col = P.qt.lpc;
        best->col_steps = P.nstep;
        best->store_policy = P.qt.nt;
        best->unroll = P.qt.u;
    }
    if (best_us) *best_us = best_t * 1e6;
    return DROT_OK;
}

}  // extern "C"
