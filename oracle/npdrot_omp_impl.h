/*
 * Body of the OpenMP restatement of the reference's CPU solver, included twice
 * by npdrot_omp.c (R = float / double).  TEST AND BASELINE INFRASTRUCTURE ONLY.
 *
 * It keeps npdrot's data flow, pass for pass, so that its time is the time of
 * the reference's CPU path on the same cores (npdrot runs its prox as a numba
 * prange loop and the sums and updates as BLAS calls):
 *   prox   X = max(X - s C, 0) / (1 + s reg), with the regularizer's value
 *          (trace_nonnegative_quad_prox_nb, numpy/npdrot/proximal/proximal.py:48-72)
 *   a2 = X e, b2 = X^T f      two gemv passes (numpy/npdrot/solver.py:115-116)
 *   t1, t2, c1, c2, yy, xx, a1, b1, r_primal     vector work (solver.py:117-131)
 *   X += (-1/n) yy e^T ; X += (-1/m) f xx^T      two dger passes
 *          (apply_adjoint_operator_and_override, proximal.py:186-210, called at
 *          solver.py:157)
 *   stop when k >= max_iters or r_primal <= eps (solver.py:165-166), then one
 *   more prox (solver.py:172).
 * Layout: X and C are npdrot's (m, n) F-order arrays, element (i, j) at
 * [j * m + i] with m = len(p): the solver layout of the GPU path.
 */

#define NP_CAT_(a, b) a##b
#define NP_CAT(a, b) NP_CAT_(a, b)
#define NPF(name) NP_CAT(name, SFX)

static double NPF(np_prox)(int m, int n, R *x, const R *c, R step, R reg, int nthreads) {
    const size_t total = (size_t)m * (size_t)n;
    const R scale = (R)1 + step * reg;
    double quad = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : quad)
    for (size_t e = 0; e < total; e++) {
        R v = x[e] - step * c[e];
        v = (v > (R)0 ? v : (R)0) / scale;
        x[e] = v;
        quad += 0.5 * (double)reg * (double)v * (double)v;
    }
    return quad * (double)reg;  /* the numba kernel's value (proximal.py:71) */
}

/* a2 = X e: threads own row ranges and sweep the columns (gemv "N") */
static void NPF(np_rowsum)(int m, int n, const R *x, R *a2, int nthreads) {
#pragma omp parallel num_threads(nthreads)
    {
        const int nt = omp_get_num_threads(), t = omp_get_thread_num();
        const int i0 = (int)((long long)m * t / nt), i1 = (int)((long long)m * (t + 1) / nt);
        for (int i = i0; i < i1; i++) a2[i] = (R)0;
        for (int j = 0; j < n; j++) {
            const R *col = x + (size_t)j * m;
            for (int i = i0; i < i1; i++) a2[i] += col[i];
        }
    }
}

/* b2 = X^T f (gemv "T") */
static void NPF(np_colsum)(int m, int n, const R *x, R *b2, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int j = 0; j < n; j++) {
        const R *col = x + (size_t)j * m;
        R s = (R)0;
        for (int i = 0; i < m; i++) s += col[i];
        b2[j] = s;
    }
}

/* X += alpha u v^T (dger) */
static void NPF(np_ger)(int m, int n, R alpha, const R *u, const R *v, R *x, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int j = 0; j < n; j++) {
        R *col = x + (size_t)j * m;
        const R av = alpha * v[j];
        for (int i = 0; i < m; i++) col[i] += av * u[i];
    }
}

/* x: the initial iterate on entry (Y0), the final prox output on return.
 * Returns the iterations executed; *r_out = last primal residual. */
int NPF(npdrot_qr_solve)(int m, int n, const R *c, const R *p, const R *q, R step, R reg,
                         int max_iters, double eps, R *x, double *r_out, int nthreads) {
    R *e = (R *)malloc(sizeof(R) * (size_t)n);
    R *f = (R *)malloc(sizeof(R) * (size_t)m);
    R *a1 = (R *)malloc(sizeof(R) * (size_t)m);
    R *a2 = (R *)malloc(sizeof(R) * (size_t)m);
    R *b1 = (R *)malloc(sizeof(R) * (size_t)n);
    R *b2 = (R *)malloc(sizeof(R) * (size_t)n);
    R *yy = (R *)malloc(sizeof(R) * (size_t)m);
    R *xx = (R *)malloc(sizeof(R) * (size_t)n);
    int k = -1;
    if (!e || !f || !a1 || !a2 || !b1 || !b2 || !yy || !xx) goto out;
    if (nthreads < 1) nthreads = 1;
    for (int j = 0; j < n; j++) e[j] = (R)1;
    for (int i = 0; i < m; i++) f[i] = (R)1;
    /* a1 = x e, b1 = x^T f (solver.py:92-93) */
    NPF(np_rowsum)(m, n, x, a1, nthreads);
    NPF(np_colsum)(m, n, x, b1, nthreads);
    k = 0;
    double r = 0.0;
    for (;;) {
        (void)NPF(np_prox)(m, n, x, c, step, reg, nthreads);
        NPF(np_rowsum)(m, n, x, a2, nthreads);
        NPF(np_colsum)(m, n, x, b2, nthreads);
        R s1 = (R)0, s2 = (R)0;
        for (int i = 0; i < m; i++) {
            yy[i] = (R)2 * a2[i] - a1[i] - p[i]; /* t1 */
            s1 += yy[i];
        }
        for (int j = 0; j < n; j++) {
            xx[j] = (R)2 * b2[j] - b1[j] - q[j]; /* t2 */
            s2 += xx[j];
        }
        const R c1 = s1 / (R)(m + n), c2 = s2 / (R)(m + n);
        double rr = 0.0;
        for (int i = 0; i < m; i++) {
            yy[i] -= c1;
            a1[i] = a2[i] - yy[i] - c2;
            const double d = (double)a2[i] - (double)p[i];
            rr += d * d;
        }
        for (int j = 0; j < n; j++) {
            xx[j] -= c2;
            b1[j] = b2[j] - xx[j] - c1;
            const double d = (double)b2[j] - (double)q[j];
            rr += d * d;
        }
        r = sqrt(rr);
        NPF(np_ger)(m, n, (R)(-1.0 / n), yy, e, x, nthreads);
        NPF(np_ger)(m, n, (R)(-1.0 / m), f, xx, x, nthreads);
        k++;
        if (k >= max_iters || r <= eps) break;
    }
    (void)NPF(np_prox)(m, n, x, c, step, reg, nthreads);
    *r_out = r;
out:
    free(e); free(f); free(a1); free(a2); free(b1); free(b2); free(yy); free(xx);
    return k;
}

#undef NPF
