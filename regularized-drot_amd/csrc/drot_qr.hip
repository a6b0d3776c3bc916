// Single-problem QR-DROT pass kernel (gfx950): one launch streams the whole
// matrix once and produces every per-tile partial the dual update needs; the
// per-tile work is qr_tile (drot_qr_tile.h).
// The pass contains no cross-block synchronisation: an agent-scope release
// (buffer_wbl2) per block costs more than the whole pass on this chip, so the
// grid-wide steps of an iteration are the kernel boundaries pass -> finalize
// -> next pass, whose prologue finishes the previous iteration
// (iteration_prologue in drot_device.h).
#include <atomic>

#include "drot_qr_tile.h"

#include <math.h>

namespace drot {

// ST: store policy of X (see Io); ST == 2 needs tc * n_rows * 4 < 2^31.
// ZS: zero-segment flags (A.zin / A.zout) in use; SC: screening as well
// (A.seg_e, A.g1, A.f1_prev / A.f1_out; branch-free tiles only).
template <int VEC, int LPC, bool EVEN, int ST, int U, bool ZS, bool SC>
__global__ __launch_bounds__(kBlock) void qr_pass_kernel(QrPassArgs A) {
    constexpr int CPW = kWave / LPC;
    constexpr int TR = LPC * VEC;
    __shared__ double s_row[kWavesPerBlock * CPW][TR];
    __shared__ double s_red[3 * kWavesPerBlock];

    const unsigned long long t0 = A.ctl.tspan ? (unsigned long long)wall_clock64() : 0ull;
    if (A.ctl.st->done) return;
    // XCD-aware tile order: blocks are dispatched round-robin over the 8 XCDs
    // (block b on XCD b % 8), so give each XCD a contiguous run of tiles in
    // row-tile-major order.  Vertically adjacent tiles then share an L2 and the
    // cache lines straddling their boundary are fetched once.
    const int n_rt = A.n_rt;
    const int n_tiles = n_rt * A.n_ct;
    const int per_xcd = (n_tiles + kXcds - 1) / kXcds;
    const int tile = (int)(blockIdx.x % kXcds) * per_xcd + (int)(blockIdx.x / kXcds);
    if (tile >= n_tiles) return;
    QrTileJob J;
    J.x = A.x;
    J.c = A.c;
    J.phi1 = A.phi1;
    J.phi2 = A.phi2;
    J.ahat = A.ahat;
    J.bhat = A.bhat;
    J.prow = A.prow;
    J.pcol = A.pcol;
    J.pobj = A.pobj + tile;
    J.n_rows = A.n_rows;
    J.n_cols = A.n_cols;
    J.ld = A.n_rows;
    J.rt = tile % n_rt;
    J.ct = tile / n_rt;
    J.nstep = A.nstep;
    J.prio = A.prio;
    J.tc = A.nstep * kWavesPerBlock * CPW;
    J.step = A.step;
    J.kappa = A.kappa;
    J.zin = A.zin;
    J.zout = A.zout;
    J.seg_e = A.seg_e;
    J.g1 = A.g1 ? A.g1 + tile : nullptr;
    J.f1_prev = A.f1_prev;
    J.f1_out = A.f1_out;
    qr_tile<VEC, LPC, EVEN, ST, U, ZS, SC>(J, A.ctl, blockIdx.x == 0, s_row, s_red);
    if (A.ctl.tspan) record_span(A.ctl.tspan, t0);
}

// ----------------------------------------------------------------- launch

// vec 1 and 4: one chunk per lane, 16 to 64 lanes per column, 2 or 4 column
// steps per load batch; vec 8: two 4-row chunks per lane, 8 or 16 lanes per
// column, one step per batch (a batch then holds as many bytes per lane as
// two steps of vec 4), for row counts that are a multiple of 4.
bool qr_tile_supported(const QrTile &t) {
    if (t.nt != 0 && t.nt != 2) return false;
    if (t.vec == 8) return (t.lpc == 8 || t.lpc == 16) && t.u == 1;
    return (t.vec == 1 || t.vec == 4) && (t.lpc == 16 || t.lpc == 32 || t.lpc == 64) &&
           (t.u == 2 || t.u == 4);
}

// Blocks of kernel k the current device holds at once (0 when the query
// fails), cached per kernel instance and device as capacity + 1 (0: not
// queried yet; concurrent first calls store the same value).
constexpr int kMaxDevices = 64;

static int grid_capacity(const void *k, std::atomic<int> *cache) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
    std::atomic<int> *slot = dev < kMaxDevices ? &cache[dev] : nullptr;
    if (slot) {
        const int v = slot->load(std::memory_order_relaxed);
        if (v > 0) return v - 1;
    }
    int n_cu = 0, per_cu = 0;
    int cap = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBlock, 0) == hipSuccess)
        cap = n_cu * per_cu;
    if (slot) slot->store(cap + 1, std::memory_order_relaxed);
    return cap;
}

template <bool EVEN, int VEC, int LPC, int ST, int U, bool ZS, bool SC>
static int pass_capacity() {
    static std::atomic<int> cache[kMaxDevices];  // zero-initialised (static storage)
    return grid_capacity(
        reinterpret_cast<const void *>(qr_pass_kernel<VEC, LPC, EVEN, ST, U, ZS, SC>), cache);
}

template <int VEC, int LPC, int ST, int U, bool ZS, bool SC>
static void launch_pass_z(bool even, int grid, const QrPassArgs &a, hipStream_t s,
                          hipEvent_t e0, hipEvent_t e1) {
    auto k = even ? qr_pass_kernel<VEC, LPC, true, ST, U, ZS, SC>
                  : qr_pass_kernel<VEC, LPC, false, ST, U, ZS, SC>;
    // ranking waves by remaining work pays only when the whole grid is resident
    // (one round of blocks); with later rounds it costs a few percent
    const int cap = even ? pass_capacity<true, VEC, LPC, ST, U, ZS, SC>()
                         : pass_capacity<false, VEC, LPC, ST, U, ZS, SC>();
    QrPassArgs b = a;
    b.prio = grid <= cap ? 1 : 0;
    if (e0 && e1)  // timing: events bracket the kernel's own execution
        hipExtLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, b);
    else
        hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, b);
}

template <int VEC, int LPC, int ST, int U>
static void launch_pass_t(bool even, int grid, const QrPassArgs &a, hipStream_t s,
                          hipEvent_t e0, hipEvent_t e1) {
    if (a.zin && a.zout && a.seg_e)
        launch_pass_z<VEC, LPC, ST, U, true, true>(even, grid, a, s, e0, e1);
    else if (a.zin && a.zout)
        launch_pass_z<VEC, LPC, ST, U, true, false>(even, grid, a, s, e0, e1);
    else
        launch_pass_z<VEC, LPC, ST, U, false, false>(even, grid, a, s, e0, e1);
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

template <int LPC>
static void launch_pass_v8(const QrTile &t, bool even, int grid, const QrPassArgs &a,
                           hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (t.nt == 2) launch_pass_t<8, LPC, 2, 1>(even, grid, a, s, e0, e1);
    else launch_pass_t<8, LPC, 0, 1>(even, grid, a, s, e0, e1);
}

int qr_pass_grid(int n_rt, int n_ct) {
    const int per = (n_rt * n_ct + kXcds - 1) / kXcds;
    return per * kXcds;
}

hipError_t launch_qr_pass(const QrTile &t, bool even, int grid, const QrPassArgs &a,
                          hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (t.vec == 8) {
        if (t.lpc == 8) launch_pass_v8<8>(t, even, grid, a, s, e0, e1);
        else launch_pass_v8<16>(t, even, grid, a, s, e0, e1);
    } else if (t.vec == 4) {
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
