// Dual update of one finalize block, shared by the single-problem and the
// batched finalize kernels.
//
// Row and column halves of quadratic_regularizer_update_aux
// (core/kernel_qr.hpp:405-461, identical in core/kernel_glr.hpp:461-520) plus
// the residual reductions of core/drot_qr.hpp:145-156:
//   rowsum = sum_ct prow ; rs = rowsum - p ; a <- a - rs ; ahat = a - rs
//   colsum = sum_rt pcol ; cs = colsum - q ; b <- b - cs ; bhat = b - cs
// Blocks [0, rb) own kFinRows rows each, the rest kFinRows columns; several
// threads share a row so all partial loads of a row are in flight at once.
// Each block leaves {sum rs^2 or cs^2, sum X', objective slice} in its pfin
// entry; the scalar part (gamma, alpha, residual, stopping test) is formed
// from those by the prologue of the next pass, so there is no grid-wide step
// and no agent-scope fence here.
#pragma once

#include "drot_kernels.h"

namespace drot {

struct FinJob {
    const double *prow, *pcol, *pobj;  // the problem's partials
    const float *p, *q;
    float *a, *b, *ahat, *bhat;        // the problem's dual vectors
    double *pfin;                      // this block's 3 scalars
    const int *done;                   // the problem's SolverState::done
    int n_rows, n_cols, n_rt, n_ct;
    int n_obj;                         // objective partials of the problem
    int rb;                            // row blocks of the problem
    int n_fin;                         // finalize blocks of the problem
    int blk;                           // this block's index within the problem
    int with_obj;
};

// Raw buffer loads for the finalize: an offset past the resource's range
// reads 0, so a lane with nothing to read needs no condition (a load under a
// lane condition becomes a branch that waits for it, and for every load
// issued before it, on the spot).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t fin_rsrc(const void *p, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double fin_load_f64(__amdgpu_buffer_rsrc_t r, unsigned off) {
    typedef unsigned int u2_t __attribute__((ext_vector_type(2)));
    const u2_t d = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    return __longlong_as_double((long long)(((unsigned long long)d.y << 32) | d.x));
}
__device__ __forceinline__ float fin_load_f32(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
constexpr unsigned kFinOob = 0x80000000u;

// The block is latency-bound: every global load (the done flag, the partials,
// p/q, a/b and the objective slice) is issued before the first of them is
// used, so the block pays one memory round trip before its stores when a row
// has at most BATCH * kBlock / kFinRows partials.  BATCH (partial loads in
// flight per thread) does not change the order of the additions.  Partial
// arrays of 2^31 bytes and more take plain loads.
template <int BATCH = 8>
__device__ inline void finalize_block(const FinJob &A, double (*s_part)[kFinRows],
                                      double *s_red) {
    constexpr int kParts = kBlock / kFinRows;
    constexpr int kBatch = BATCH;
    const int N = A.n_rows, M = A.n_cols;
    const int b = A.blk;
    const int part = threadIdx.x / kFinRows;
    const int li = threadIdx.x % kFinRows;
    const bool is_row = b < A.rb;
    const int idx = (is_row ? b : b - A.rb) * kFinRows + li;
    const int len = is_row ? N : M;
    const int cnt = is_row ? A.n_ct : A.n_rt;
    const bool valid = idx < len;
    const double *base = is_row ? A.prow : A.pcol;
    const double *src = base + (valid ? idx : 0);
    const int done = *A.done;
    const bool buf = (long long)cnt * len * 8 < (1LL << 31);  // block-uniform
    double v[kBatch];
    if (buf) {
        const __amdgpu_buffer_rsrc_t pr = fin_rsrc(base, (long long)cnt * len * 8);
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = part + u * kParts;
            v[u] = fin_load_f64(pr, (valid && t < cnt) ? ((unsigned)t * (unsigned)len + (unsigned)idx) * 8u
                                                       : kFinOob);
        }
    } else {
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = part + u * kParts;
            v[u] = (valid && t < cnt) ? src[(size_t)t * len] : 0.0;
        }
    }
    const bool owner = part == 0 && valid;
    const unsigned oo = owner ? (unsigned)idx * 4u : kFinOob;
    const float marg = fin_load_f32(fin_rsrc(is_row ? A.p : A.q, (long long)len * 4), oo);
    const float acc = fin_load_f32(fin_rsrc(is_row ? A.a : A.b, (long long)len * 4), oo);
    double obj = 0.0;
    if (A.with_obj) {  // this block's slice of the pass's objective partials
        const int per = (A.n_obj + A.n_fin - 1) / A.n_fin;
        const int o0 = b * per;
        const int o1 = min(o0 + per, A.n_obj);
        const __amdgpu_buffer_rsrc_t orr = fin_rsrc(A.pobj, (long long)A.n_obj * 8);
        const int t = o0 + (int)threadIdx.x;
        obj += fin_load_f64(orr, t < o1 ? (unsigned)t * 8u : kFinOob);
        for (int t2 = t + kBlock; t2 < o1; t2 += kBlock) obj += A.pobj[t2];
    }
    if (done) return;  // the solve ended: leave the dual state as it is
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < kBatch; u++) sum += v[u];
    for (int t0 = part + kBatch * kParts; t0 < cnt; t0 += kBatch * kParts) {
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = t0 + u * kParts;
            v[u] = (valid && t < cnt) ? src[(size_t)t * len] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) sum += v[u];
    }
    s_part[part][li] = sum;
    __syncthreads();
    double res = 0.0, S = 0.0;
    if (owner) {
        double tot = 0.0;
#pragma unroll
        for (int q = 0; q < kParts; q++) tot += s_part[q][li];
        const double d = tot - (double)marg;
        const float r = (float)d;
        const float an = acc - r;
        if (is_row) {
            A.a[idx] = an;
            A.ahat[idx] = an - r;
        } else {
            A.b[idx] = an;
            A.bhat[idx] = an - r;
            S = tot;
        }
        res = d * d;
    }
    block_sum3(res, S, obj, s_red, false);  // the block ends after these stores
    if (threadIdx.x == 0) {
        A.pfin[0] = res;
        A.pfin[1] = S;
        A.pfin[2] = obj;
    }
}

// A thread's sum of the partials src[t * len], t < cnt, in finalize_block's
// grouping: part q adds t = q, q + kParts, ... in order, then the parts are
// added in order.  CHUNK loads are in flight at once (a multiple of kParts).
template <int CHUNK>
__device__ __forceinline__ double partial_sum_grouped(const double *src, int len, int cnt,
                                                      bool valid) {
    constexpr int kParts = kBlock / kFinRows;
    static_assert(CHUNK % kParts == 0, "whole rounds of the parts");
    double part[kParts];
#pragma unroll
    for (int q = 0; q < kParts; q++) part[q] = 0.0;
    for (int t0 = 0; t0 < cnt; t0 += CHUNK) {
        double v[CHUNK];
#pragma unroll
        for (int u = 0; u < CHUNK; u++)
            v[u] = (valid && t0 + u < cnt) ? src[(size_t)(t0 + u) * len] : 0.0;
#pragma unroll
        for (int u = 0; u < CHUNK; u++) part[u % kParts] += v[u];
    }
    double tot = 0.0;
#pragma unroll
    for (int q = 0; q < kParts; q++) tot += part[q];
    return tot;
}

// The same sum through a buffer resource over the [cnt][len] partials (under
// 2^31 bytes): entry (t, idx).
template <int CHUNK>
__device__ __forceinline__ double partial_sum_grouped_buf(__amdgpu_buffer_rsrc_t r, int len,
                                                          int idx, int cnt, bool valid) {
    constexpr int kParts = kBlock / kFinRows;
    static_assert(CHUNK % kParts == 0, "whole rounds of the parts");
    double part[kParts];
#pragma unroll
    for (int q = 0; q < kParts; q++) part[q] = 0.0;
    for (int t0 = 0; t0 < cnt; t0 += CHUNK) {
        double v[CHUNK];
#pragma unroll
        for (int u = 0; u < CHUNK; u++) {
            const int t = t0 + u;
            v[u] = fin_load_f64(r, (valid && t < cnt) ? ((unsigned)t * (unsigned)len + (unsigned)idx) * 8u
                                                      : kFinOob);
        }
#pragma unroll
        for (int u = 0; u < CHUNK; u++) part[u % kParts] += v[u];
    }
    double tot = 0.0;
#pragma unroll
    for (int q = 0; q < kParts; q++) tot += part[q];
    return tot;
}

// The same dual update with one thread per row or column, kBlock of them per
// block: blocks [0, rb) own kBlock rows each, the rest kBlock columns.  For
// problems whose rows and columns have few partials (the batch: at most 8 row
// partials per 2048 columns, 32 column partials per 2048 rows), where
// finalize_block's eight threads per row would mostly idle and its many small
// blocks would run in several rounds.  A thread adds its partials in
// finalize_block's grouping (part q adds the partials t = q, q + 8, ... in
// order; then the eight parts in order), so a row's or column's sum has the
// same bits as there.  The per-block scalars {sum rs^2 or cs^2, sum X',
// objective slice} cover kBlock rows or columns instead of kFinRows.
__device__ inline void finalize_rows_block(const FinJob &A, double *s_red) {
    constexpr int kParts = kBlock / kFinRows;
    const int N = A.n_rows, M = A.n_cols;
    const int b = A.blk;
    const bool is_row = b < A.rb;
    const int idx = (is_row ? b : b - A.rb) * kBlock + (int)threadIdx.x;
    const int len = is_row ? N : M;
    const int cnt = is_row ? A.n_ct : A.n_rt;
    const bool valid = idx < len;
    const double *base = is_row ? A.prow : A.pcol;
    const double *src = base + (valid ? idx : 0);
    const int done = *A.done;
    // all partial loads of a thread in flight at once: up to kParts of them in
    // the short form (a row of a batch problem of up to 2048 columns), up to
    // 4 kParts in the long one (a column of up to 2048 rows); buffer loads
    // below 2^31 bytes of partials
    const long long pbytes = (long long)cnt * len * 8;
    double tot;
    if (pbytes < (1LL << 31)) {
        const __amdgpu_buffer_rsrc_t pr = fin_rsrc(base, pbytes);
        tot = cnt <= kParts ? partial_sum_grouped_buf<kParts>(pr, len, idx, cnt, valid)
                            : partial_sum_grouped_buf<4 * kParts>(pr, len, idx, cnt, valid);
    } else {
        tot = cnt <= kParts ? partial_sum_grouped<kParts>(src, len, cnt, valid)
                            : partial_sum_grouped<4 * kParts>(src, len, cnt, valid);
    }
    const unsigned oo = valid ? (unsigned)idx * 4u : kFinOob;
    const float marg = fin_load_f32(fin_rsrc(is_row ? A.p : A.q, (long long)len * 4), oo);
    const float acc = fin_load_f32(fin_rsrc(is_row ? A.a : A.b, (long long)len * 4), oo);
    double obj = 0.0;
    if (A.with_obj) {  // this block's slice of the pass's objective partials
        const int per = (A.n_obj + A.n_fin - 1) / A.n_fin;
        const int o0 = b * per;
        const int o1 = min(o0 + per, A.n_obj);
        const __amdgpu_buffer_rsrc_t orr = fin_rsrc(A.pobj, (long long)A.n_obj * 8);
        const int t = o0 + (int)threadIdx.x;
        obj += fin_load_f64(orr, t < o1 ? (unsigned)t * 8u : kFinOob);
        for (int t2 = t + kBlock; t2 < o1; t2 += kBlock) obj += A.pobj[t2];
    }
    if (done) return;  // the solve ended: leave the dual state as it is
    double res = 0.0, S = 0.0;
    if (valid) {
        const double d = tot - (double)marg;
        const float r = (float)d;
        const float an = acc - r;
        if (is_row) {
            A.a[idx] = an;
            A.ahat[idx] = an - r;
        } else {
            A.b[idx] = an;
            A.bhat[idx] = an - r;
            S = tot;
        }
        res = d * d;
    }
    block_sum3(res, S, obj, s_red, false);  // the block ends after these stores
    if (threadIdx.x == 0) {
        A.pfin[0] = res;
        A.pfin[1] = S;
        A.pfin[2] = obj;
    }
}

}  // namespace drot
