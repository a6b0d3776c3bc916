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
    int n_rows, n_cols, n_rt, n_ct;
    int n_obj;                         // objective partials of the problem
    int rb;                            // row blocks of the problem
    int n_fin;                         // finalize blocks of the problem
    int blk;                           // this block's index within the problem
    int with_obj;
};

__device__ inline void finalize_block(const FinJob &A, double (*s_part)[kFinRows],
                                      double *s_red) {
    constexpr int kParts = kBlock / kFinRows;
    const int N = A.n_rows, M = A.n_cols;
    const int b = A.blk;
    const int part = threadIdx.x / kFinRows;
    const int li = threadIdx.x % kFinRows;
    const bool is_row = b < A.rb;
    const int idx = (is_row ? b : b - A.rb) * kFinRows + li;
    const int len = is_row ? N : M;
    const int cnt = is_row ? A.n_ct : A.n_rt;
    const bool valid = idx < len;
    const double *src = (is_row ? A.prow : A.pcol) + (valid ? idx : 0);
    double sum = 0.0;
    if (valid) {
        int t = part;
        for (; t + (8 - 1) * kParts < cnt; t += 8 * kParts) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = src[(size_t)(t + u * kParts) * len];
#pragma unroll
            for (int u = 0; u < 8; u++) sum += v[u];
        }
        for (; t < cnt; t += kParts) sum += src[(size_t)t * len];
    }
    s_part[part][li] = sum;
    __syncthreads();
    double res = 0.0, S = 0.0;
    if (part == 0 && valid) {
        double tot = 0.0;
#pragma unroll
        for (int q = 0; q < kParts; q++) tot += s_part[q][li];
        if (is_row) {
            const double d = tot - (double)A.p[idx];
            const float rs = (float)d;
            const float an = A.a[idx] - rs;
            A.a[idx] = an;
            A.ahat[idx] = an - rs;
            res = d * d;
        } else {
            const double d = tot - (double)A.q[idx];
            const float cs = (float)d;
            const float bn = A.b[idx] - cs;
            A.b[idx] = bn;
            A.bhat[idx] = bn - cs;
            res = d * d;
            S = tot;
        }
    }
    // this block's slice of the pass's objective partials
    double obj = 0.0;
    if (A.with_obj) {
        const int per = (A.n_obj + A.n_fin - 1) / A.n_fin;
        const int o0 = b * per;
        const int o1 = min(o0 + per, A.n_obj);
        for (int t = o0 + threadIdx.x; t < o1; t += kBlock) obj += A.pobj[t];
    }
    block_sum3(res, S, obj, s_red);
    if (threadIdx.x == 0) {
        A.pfin[0] = res;
        A.pfin[1] = S;
        A.pfin[2] = obj;
    }
}

}  // namespace drot
