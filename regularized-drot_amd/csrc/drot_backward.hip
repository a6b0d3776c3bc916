// Implicit-differentiation backward of QR-DROT (gfx950).
//
// The reference solver has no backward (pytorch/src/qr_drot_cuda_kernel.cu:99
// returns a tensor without grad).  For the QR problem
//     min <C,X> + (reg/2)||X||^2  s.t.  X1 = p, X^T 1 = q, X >= 0,
//     reg = (kappa - 1) / step
// the KKT conditions on the support S = {X_ij > 0} read
//     reg X_ij = u_i + v_j - C_ij   (i,j) in S,
// and differentiating the marginal constraints gives
//     K [du; dv] = [(S.dC) 1; (S.dC)^T 1],   K = [[D_r, S], [S^T, D_c]],
// the signless Laplacian of the bipartite support graph (D_r, D_c: support
// counts of rows and columns).  For an upstream gradient G = dL/dX,
//     dL/dC_ij = (w_i + w'_j - G_ij) / reg  on S,   0 off S,
// where [w; w'] solves K [w; w'] = [(S.G) 1; (S.G)^T 1] (a consistent system:
// the right side lies in range(K), so any solution gives the same dL/dC).
// A gradient g_obj of the regularized objective adds g_obj * X (envelope
// theorem: the objective's derivative in C is the plan itself).
//
// Supports of QR-regularized plans are sparse (a few entries per row at the
// usual regularization), so the system is solved on compressed index lists:
//   count  (one pass over X, C; writes grad_c = g_obj * X)   per column tile
//   fill   (second pass: the CSC row-index list, in column-major order)
//   cg     one block per problem: CSR from the CSC list, Jacobi-preconditioned
//          CG in fp64 with fixed-order block reductions, then the scatter of
//          the support entries of dL/dC.
// Everything is deterministic: list order is fixed by the matrix order, row
// lists are sorted by column after the atomic fill, and every sum runs in a
// fixed order.
//
// Support test: element (i,j) is in S when X_ij > 0 and X_ij > thr |C_ij| with
// thr = tol * step.  The solver stores X - step*C on even iterations and adds
// step*C back at the end (fix_x), which leaves a rounding residue of at most
// half an ulp of step*|C_ij| (2^-24 step |C_ij|) where the true plan is zero;
// tol = 2^-22 rejects those residues.
#include "drot_kernels.h"

#include <math.h>

namespace drot {

__device__ __forceinline__ bool in_support(float xv, float cv, float thr) {
    return xv > 0.0f && xv > thr * fabsf(cv);
}

__device__ __forceinline__ unsigned long long lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

// Exclusive block-wide prefix sum of one int per thread; *total receives the
// block total.  s_w holds kWavesPerBlock ints.
__device__ __forceinline__ int block_excl_scan(int v, int *s_w, int *total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    int incl = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int t = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += t;
    }
    if (lane == kWave - 1) s_w[w] = incl;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
        const int t = s_w[k];
        base += (k < w) ? t : 0;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

// ------------------------------------------------------------------- count

// One block per column tile: support size of the tile, and the dense part of
// the gradient, grad_c = g_obj * X (0 in the padding rows n_rows..ld-1).
__global__ __launch_bounds__(kBlock) void bwd_count_kernel(BwdArgs A) {
    __shared__ double s_red[kWavesPerBlock];
    const BwdTile T = A.tiles[blockIdx.x];
    const BwdProblem &P = A.probs[T.prob];
    const float gobj = A.gobj ? A.gobj[T.prob] : 0.0f;
    const int n_rows = P.n_rows, ld = P.ld;
    const float thr = P.thr;
    int cnt = 0;
    for (int j = T.c0; j < T.c1; j++) {
        const long long col = P.mat_off + (long long)j * ld;
        for (int i = threadIdx.x; i < ld; i += kBlock) {
            float gv = 0.0f;
            if (i < n_rows) {
                const float xv = A.x[col + i];
                cnt += in_support(xv, A.c[col + i], thr) ? 1 : 0;
                gv = gobj * xv;
            }
            A.grad_c[col + i] = gv;
        }
    }
    const double t = block_sum((double)cnt, s_red);
    if (threadIdx.x == 0) A.tile_cnt[blockIdx.x] = (int)t;
}

// -------------------------------------------------------------------- fill

// Writes the CSC lists: colptr[j] and the row indices of every support
// element in column-major order, at the offset given by the support counts
// of the problem's earlier tiles.
__global__ __launch_bounds__(kBlock) void bwd_fill_kernel(BwdArgs A) {
    __shared__ double s_red[kWavesPerBlock];
    __shared__ int s_w[kWavesPerBlock];
    const BwdTile T = A.tiles[blockIdx.x];
    const BwdProblem &P = A.probs[T.prob];
    const int n_rows = P.n_rows, ld = P.ld;
    const float thr = P.thr;
    const long long cap = P.cap;
    int *colptr = A.iptr + P.iptr_off;
    int *rowidx = A.idx + P.idx_off;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;

    double before = 0.0;  // exact: counts stay far below 2^53
    for (int t = P.tile_begin + threadIdx.x; t < (int)blockIdx.x; t += kBlock)
        before += (double)A.tile_cnt[t];
    long long pos = (long long)block_sum(before, s_red);

    for (int j = T.c0; j < T.c1; j++) {
        if (threadIdx.x == 0) colptr[j] = (int)(pos < cap ? pos : cap);
        const long long col = P.mat_off + (long long)j * ld;
        for (int i0 = 0; i0 < n_rows; i0 += kBlock) {
            const int i = i0 + threadIdx.x;
            const bool f = i < n_rows && in_support(A.x[col + i], A.c[col + i], thr);
            const unsigned long long m = __ballot(f);
            if (lane == 0) s_w[w] = __popcll(m);
            __syncthreads();
            int wb = 0, tot = 0;
#pragma unroll
            for (int k = 0; k < kWavesPerBlock; k++) {
                const int t = s_w[k];
                wb += (k < w) ? t : 0;
                tot += t;
            }
            __syncthreads();
            if (f) {
                const long long e = pos + wb + __popcll(m & lanes_below(lane));
                if (e < cap) rowidx[e] = i;
            }
            pos += tot;
        }
    }
    if ((int)blockIdx.x == P.tile_begin + P.n_tiles - 1 && threadIdx.x == 0) {
        colptr[P.n_cols] = (int)(pos < cap ? pos : cap);
        A.st[T.prob].nnz = pos;
    }
}

// ---------------------------------------------------------------------- cg

__device__ __forceinline__ int atomic_read(int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One block per problem.
__global__ __launch_bounds__(kBlock) void bwd_cg_kernel(BwdArgs A) {
    __shared__ double s_red[3 * kWavesPerBlock];
    __shared__ int s_w[kWavesPerBlock];
    const int tid = threadIdx.x;
    const BwdProblem &P = A.probs[blockIdx.x];
    BwdState *st = A.st + blockIdx.x;
    const long long nnz_l = st->nnz;
    if (nnz_l > P.cap) {  // index lists incomplete: the host reports the overflow
        if (tid == 0) {
            st->cg_iters = 0;
            st->converged = 0;
            st->res = 0.0;
            st->gnorm = 0.0;
        }
        return;
    }
    const int nnz = (int)nnz_l;
    const int nr = P.n_rows, nc = P.n_cols, L = nr + nc, ld = P.ld;
    const long long mo = P.mat_off;
    int *colptr = A.iptr + P.iptr_off;
    int *rowptr = colptr + nc + 1;
    int *rowcur = rowptr + nr + 1;
    const int *rowidx = A.idx + P.idx_off;
    int *colidx = A.idx + P.idx_off + P.cap;
    double *wv = A.vec + P.vec_off;
    double *rv = wv + L;
    double *zv = rv + L;
    double *pv = zv + L;
    double *qv = pv + L;
    const float *G = A.g;

    // 1. CSR from the CSC list: row counts, offsets, fill, per-row sort
    for (int i = tid; i < nr; i += kBlock) __hip_atomic_store(&rowcur[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int e = tid; e < nnz; e += kBlock) atomicAdd(&rowcur[rowidx[e]], 1);
    __syncthreads();
    int run = 0;
    for (int i0 = 0; i0 < nr; i0 += kBlock) {
        const int i = i0 + tid;
        const int v = i < nr ? atomic_read(&rowcur[i]) : 0;
        int tot = 0;
        const int ex = block_excl_scan(v, s_w, &tot);
        if (i < nr) rowptr[i] = run + ex;
        run += tot;
    }
    if (tid == 0) rowptr[nr] = run;
    for (int i = tid; i < nr; i += kBlock) __hip_atomic_store(&rowcur[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int j = tid; j < nc; j += kBlock) {
        const int e1 = colptr[j + 1];
        for (int e = colptr[j]; e < e1; e++) {
            const int i = rowidx[e];
            const int slot = atomicAdd(&rowcur[i], 1);
            colidx[rowptr[i] + slot] = j;
        }
    }
    __syncthreads();
    for (int i = tid; i < nr; i += kBlock) {  // lists are short: insertion sort
        const int b = rowptr[i], e = rowptr[i + 1];
        for (int a = b + 1; a < e; a++) {
            const int v = colidx[a];
            int k = a - 1;
            while (k >= b && colidx[k] > v) {
                colidx[k + 1] = colidx[k];
                k--;
            }
            colidx[k + 1] = v;
        }
    }
    __syncthreads();

    // 2. right side g = [(S.G) 1; (S.G)^T 1], Jacobi start
    double gg = 0.0, rz = 0.0, unused = 0.0;
    for (int v = tid; v < L; v += kBlock) {
        double acc = 0.0;
        int deg;
        if (v < nr) {
            const int b = rowptr[v], e = rowptr[v + 1];
            deg = e - b;
            for (int k = b; k < e; k++) acc += (double)G[mo + (long long)colidx[k] * ld + v];
        } else {
            const int j = v - nr;
            const int b = colptr[j], e = colptr[j + 1];
            deg = e - b;
            const long long col = mo + (long long)j * ld;
            for (int k = b; k < e; k++) acc += (double)G[col + rowidx[k]];
        }
        const double z = deg > 0 ? acc / (double)deg : 0.0;
        wv[v] = 0.0;
        rv[v] = acc;
        zv[v] = z;
        pv[v] = z;
        gg += acc * acc;
        rz += acc * z;
    }
    block_sum3(gg, rz, unused, s_red);
    const double gnorm = sqrt(gg);
    const double target = P.cg_tol * gnorm;
    double rnorm = gnorm;
    int it = 0;
    bool conv = !(gnorm > target);

    // 3. preconditioned CG (fp64)
    while (!conv && it < P.max_cg) {
        double pq = 0.0, d1 = 0.0, d2 = 0.0;
        for (int v = tid; v < L; v += kBlock) {
            double acc;
            if (v < nr) {
                const int b = rowptr[v], e = rowptr[v + 1];
                acc = (double)(e - b) * pv[v];
                for (int k = b; k < e; k++) acc += pv[nr + colidx[k]];
            } else {
                const int j = v - nr;
                const int b = colptr[j], e = colptr[j + 1];
                acc = (double)(e - b) * pv[v];
                for (int k = b; k < e; k++) acc += pv[rowidx[k]];
            }
            qv[v] = acc;
            pq += pv[v] * acc;
        }
        block_sum3(pq, d1, d2, s_red);
        if (!(pq > 0.0)) break;
        const double alpha = rz / pq;
        double rr = 0.0, rzn = 0.0;
        d1 = 0.0;
        for (int v = tid; v < L; v += kBlock) {
            const int deg = v < nr ? rowptr[v + 1] - rowptr[v] : colptr[v - nr + 1] - colptr[v - nr];
            wv[v] += alpha * pv[v];
            const double r = rv[v] - alpha * qv[v];
            rv[v] = r;
            const double z = deg > 0 ? r / (double)deg : 0.0;
            zv[v] = z;
            rr += r * r;
            rzn += r * z;
        }
        block_sum3(rr, rzn, d1, s_red);
        it++;
        rnorm = sqrt(rr);
        if (rnorm <= target) {
            conv = true;
            break;
        }
        const double beta = rzn / rz;
        rz = rzn;
        for (int v = tid; v < L; v += kBlock) pv[v] = zv[v] + beta * pv[v];
        __syncthreads();
    }
    __syncthreads();

    // 4. support entries of dL/dC = g_obj X + (w_i + w'_j - G_ij) / reg
    const double reg = P.reg;
    const double gobj = A.gobj ? (double)A.gobj[blockIdx.x] : 0.0;
    for (int j = tid; j < nc; j += kBlock) {
        const double wc = wv[nr + j];
        const int e1 = colptr[j + 1];
        const long long col = mo + (long long)j * ld;
        for (int e = colptr[j]; e < e1; e++) {
            const int i = rowidx[e];
            const long long o = col + i;
            const double val = gobj * (double)A.x[o] + (wv[i] + wc - (double)G[o]) / reg;
            A.grad_c[o] = (float)val;
        }
    }
    if (tid == 0) {
        st->cg_iters = it;
        st->converged = conv ? 1 : 0;
        st->res = gnorm > 0.0 ? rnorm / gnorm : 0.0;
        st->gnorm = gnorm;
    }
}

// ----------------------------------------------------------------- launch

hipError_t launch_bwd_count(const BwdArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(bwd_count_kernel, dim3(a.n_tiles), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bwd_fill(const BwdArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(bwd_fill_kernel, dim3(a.n_tiles), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bwd_cg(const BwdArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(bwd_cg_kernel, dim3(a.n_probs), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace drot
