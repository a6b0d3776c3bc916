/*
 * Body of the CPU oracle, included twice by drot_oracle.c:
 *   R = float,  SFX = _f32  -> the fp32 trajectory the GPU path must reproduce
 *   R = double, SFX = _f64  -> the FP64 oracle run (objective / residual reporting)
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path links or calls this.
 *
 * Arithmetic contract shared with the HIP kernels (and restating the reference):
 *   - per-element work in R, with a fixed operation order and an explicit fused
 *     multiply-add where the reference's nvcc build contracts `t - s*c`
 *     (core/kernel_qr.hpp:90-91, :241-245, :479; core/kernel_glr.hpp:109,233);
 *   - every reduction (row sums, column sums, total sum, objective, residual,
 *     group norms) is accumulated in double.  The reference accumulates in fp32
 *     with atomics in a nondeterministic order; double accumulation is the
 *     order-independent limit of the same sums, so two implementations that
 *     both accumulate in double agree after the final rounding to R.
 *
 * Layout: column-major as in the reference (core/kernel_qr.hpp:3):
 *   x[j * n_rows + i], i < n_rows = len(p) (contiguous), j < n_cols = len(q).
 */

#define OR_CAT_(a, b) a##b
#define OR_CAT(a, b) OR_CAT_(a, b)
#define ORF(name) OR_CAT(name, SFX)

/* Warm start, pytorch/src/qr_drot_cuda_kernel.cu:45-71 and core/drot_qr.hpp:321-347
 * (identical formulas in the GL files, glr_drot_cuda_kernel.cu:52-78).
 * Evaluated in R on the host exactly as written there.  X = 0. */
void ORF(oracle_warm_start)(int n_rows, int n_cols, R step, R *x, R *phi1,
                            R *phi2, R *a, R *b, R *alpha) {
    const R n = (R)n_rows;
    const R m = (R)n_cols;
    const R k = (R)1.0 * step * n * m - (R)2.0;
    const R v_phi1 = (k + (R)2) / n / (m + n);
    const R v_phi2 = (k + (R)2) / m / (m + n);
    const R v_a = (k + (R)1) / n;
    const R v_b = (k + (R)1) / m;
    for (int i = 0; i < n_rows; i++) { phi1[i] = v_phi1; a[i] = v_a; }
    for (int j = 0; j < n_cols; j++) { phi2[j] = v_phi2; b[j] = v_b; }
    *alpha = k + (R)1;
    memset(x, 0, sizeof(R) * (size_t)n_rows * (size_t)n_cols);
}

/* Cold start, core/drot_qr.hpp:308-320 with init_x (core/kernel_qr.hpp:7-40):
 * X = p q^T, phi = a = b = 0.  The reference leaves alpha (aux[4]) unset on this
 * path (core/drot_qr.hpp:191 clears aux[0..3] only); the oracle sets it to 0. */
void ORF(oracle_cold_start)(int n_rows, int n_cols, const R *p, const R *q, R *x,
                            R *phi1, R *phi2, R *a, R *b, R *alpha) {
    for (int j = 0; j < n_cols; j++)
        for (int i = 0; i < n_rows; i++)
            x[(size_t)j * n_rows + i] = p[i] * q[j];
    for (int i = 0; i < n_rows; i++) { phi1[i] = 0; a[i] = 0; }
    for (int j = 0; j < n_cols; j++) { phi2[j] = 0; b[j] = 0; }
    *alpha = 0;
}

/* Dual / aux update shared by QR and GL:
 * core/kernel_qr.hpp:405-467 (quadratic_regularizer_update_aux, update_alpha),
 * core/kernel_glr.hpp:461-526 (identical).  gamma uses alpha BEFORE its update.
 *   gamma = (2 (S - 1) - alpha) / (n_rows + n_cols)           [in double, -> R]
 *   rs = r - p ; a <- a - rs ; phi1 <- ((a - rs) + gamma) / n_cols
 *   cs = c - q ; b <- b - cs ; phi2 <- ((b - cs) + gamma) / n_rows
 *   alpha <- (alpha + 1) - S                                 [in double, -> R]
 * aux_out[2], aux_out[3] receive sum(rs^2), sum(cs^2) in double. */
static void ORF(dual_update)(int n_rows, int n_cols, const double *rowsum,
                             const double *colsum, double xsum, const R *p,
                             const R *q, R *phi1, R *phi2, R *a, R *b, R *alpha,
                             double *aux_out) {
    const R gamma = (R)((2.0 * (xsum - 1.0) - (double)(*alpha)) /
                        (double)(n_rows + n_cols));
    double res_r = 0.0, res_c = 0.0;
    for (int i = 0; i < n_rows; i++) {
        const double d = rowsum[i] - (double)p[i];
        const R rs = (R)d;
        const R an = a[i] - rs;
        a[i] = an;
        phi1[i] = ((an - rs) + gamma) / (R)n_cols;
        res_r += d * d;
    }
    for (int j = 0; j < n_cols; j++) {
        const double d = colsum[j] - (double)q[j];
        const R cs = (R)d;
        const R bn = b[j] - cs;
        b[j] = bn;
        phi2[j] = ((bn - cs) + gamma) / (R)n_rows;
        res_c += d * d;
    }
    *alpha = (R)(((double)(*alpha) + 1.0) - xsum);
    aux_out[2] = res_r;
    aux_out[3] = res_c;
}

/* One QR-DROT iteration k (core/drot_qr.hpp:52-157, core/kernel_qr.hpp:50-376):
 *   X' = max(X + phi1 1^T + 1 phi2^T - s C, 0) / kappa
 *   even k: stored X := X' - s C, objective = <C, X'>
 *   odd  k: stored X := X' (C is not read; the -sC is already folded in)
 * Row sums, column sums, total sum and objective are taken over X'.
 * aux_out = {obj (even k only, else 0), sum X', sum rs^2, sum cs^2}. */
int ORF(oracle_qr_step)(int k, int n_rows, int n_cols, R *x, const R *c,
                        const R *p, const R *q, R *phi1, R *phi2, R *a, R *b,
                        R *alpha, R step, R kappa, double *aux_out,
                        int nthreads) {
    const int even = (k % 2) == 0;
    double *rowsum = (double *)calloc((size_t)n_rows, sizeof(double));
    double *colsum = (double *)calloc((size_t)n_cols, sizeof(double));
    if (!rowsum || !colsum) { free(rowsum); free(colsum); return -1; }
    double obj = 0.0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : obj)
    {
        double *rs_local = (double *)calloc((size_t)n_rows, sizeof(double));
#pragma omp for schedule(static)
        for (int j = 0; j < n_cols; j++) {
            R *xc = x + (size_t)j * n_rows;
            const R *cc = c + (size_t)j * n_rows;
            const R f2 = phi2[j];
            double cs = 0.0;
            for (int i = 0; i < n_rows; i++) {
                const R t = (xc[i] + phi1[i]) + f2;
                R y;
                if (even) {
                    y = OR_FMA(-step, cc[i], t);
                } else {
                    y = t;
                }
                y = (y > (R)0) ? y : (R)0;
                const R xp = y / kappa;
                if (even) {
                    xc[i] = OR_FMA(-step, cc[i], xp);
                    obj += (double)xp * (double)cc[i];
                } else {
                    xc[i] = xp;
                }
                rs_local[i] += (double)xp;
                cs += (double)xp;
            }
            colsum[j] = cs;
        }
#pragma omp critical
        for (int i = 0; i < n_rows; i++) rowsum[i] += rs_local[i];
        free(rs_local);
    }
    double xsum = 0.0;
    for (int j = 0; j < n_cols; j++) xsum += colsum[j];
    aux_out[0] = even ? obj : 0.0;
    aux_out[1] = xsum;
    ORF(dual_update)(n_rows, n_cols, rowsum, colsum, xsum, p, q, phi1, phi2, a, b,
                     alpha, aux_out);
    free(rowsum);
    free(colsum);
    return 0;
}

/* fix_x on exit, core/kernel_qr.hpp:469-482: X = max(X + s C, 0). */
void ORF(oracle_fix_x)(int n_rows, int n_cols, R *x, const R *c, R step) {
    const size_t total = (size_t)n_rows * (size_t)n_cols;
    for (size_t e = 0; e < total; e++) {
        const R y = OR_FMA(step, c[e], x[e]);
        x[e] = (y > (R)0) ? y : (R)0;
    }
}

/* One GL-DROT iteration (core/drot_glr.hpp:87-240, core/kernel_glr.hpp:59-399):
 *   X' = max(X + phi1 + phi2 - s C, 0)                       (no kappa)
 *   for each column j and group g = rows [bounds[g], bounds[g+1]):
 *       scale = max(1 - lambda / (||X'_{g,j}||_2 + 1e-9), 0)   [double, -> R]
 *       X_{g,j} := X'_{g,j} * scale                            [in R]
 * Sums and objective are taken over the stored (scaled) X.
 * aux_out = {<C,X>, sum X, sum rs^2, sum cs^2}. */
int ORF(oracle_gl_step)(int n_rows, int n_cols, int n_groups, const int *bounds,
                        R *x, const R *c, const R *p, const R *q, R *phi1,
                        R *phi2, R *a, R *b, R *alpha, R step, R lambda,
                        double *aux_out, int nthreads) {
    double *rowsum = (double *)calloc((size_t)n_rows, sizeof(double));
    double *colsum = (double *)calloc((size_t)n_cols, sizeof(double));
    if (!rowsum || !colsum) { free(rowsum); free(colsum); return -1; }
    double obj = 0.0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : obj)
    {
        double *rs_local = (double *)calloc((size_t)n_rows, sizeof(double));
#pragma omp for schedule(static)
        for (int j = 0; j < n_cols; j++) {
            R *xc = x + (size_t)j * n_rows;
            const R *cc = c + (size_t)j * n_rows;
            const R f2 = phi2[j];
            double cs = 0.0;
            for (int g = 0; g < n_groups; g++) {
                const int i0 = bounds[g], i1 = bounds[g + 1];
                double nrm2 = 0.0;
                for (int i = i0; i < i1; i++) {
                    const R t = (xc[i] + phi1[i]) + f2;
                    R y = OR_FMA(-step, cc[i], t);
                    y = (y > (R)0) ? y : (R)0;
                    xc[i] = y;
                    nrm2 += (double)y * (double)y;
                }
                double sc = 1.0 - (double)lambda / (sqrt(nrm2) + 1e-9);
                const R scale = (R)(sc > 0.0 ? sc : 0.0);
                for (int i = i0; i < i1; i++) {
                    const R v = xc[i] * scale;
                    xc[i] = v;
                    rs_local[i] += (double)v;
                    cs += (double)v;
                    obj += (double)v * (double)cc[i];
                }
            }
            colsum[j] = cs;
        }
#pragma omp critical
        for (int i = 0; i < n_rows; i++) rowsum[i] += rs_local[i];
        free(rs_local);
    }
    double xsum = 0.0;
    for (int j = 0; j < n_cols; j++) xsum += colsum[j];
    aux_out[0] = obj;
    aux_out[1] = xsum;
    ORF(dual_update)(n_rows, n_cols, rowsum, colsum, xsum, p, q, phi1, phi2, a, b,
                     alpha, aux_out);
    free(rowsum);
    free(colsum);
    return 0;
}

/* QR solve loop, core/drot_qr.hpp:159-234 with the warm start of
 * pytorch/src/qr_drot_cuda_kernel.cu:45-71 (or the cold start of
 * core/drot_qr.hpp:308-320):
 *   - the stopping test runs on even k only: residual < eps (strict);
 *     on success fix_x is applied and n_iter = k + 1;
 *   - otherwise the loop ends at max_iters with n_iter = max_iters + 1 and X as
 *     stored (shifted by -sC if the last executed iteration was even, as in the
 *     reference);
 *   - objective = obj of the last even iteration.
 * out[0] = residual (sqrt(sum rs^2 + sum cs^2)) of the last even iteration,
 * out[1] = objective, out[2] = converged flag. */
int ORF(oracle_qr_solve)(int n_rows, int n_cols, const R *c, const R *p,
                         const R *q, R step, R kappa, int max_iters, R eps,
                         int warm_start, R *x, R *phi1, R *phi2, R *a, R *b,
                         R *alpha, int *n_iter_out, double *out, int nthreads) {
    if (warm_start)
        ORF(oracle_warm_start)(n_rows, n_cols, step, x, phi1, phi2, a, b, alpha);
    else
        ORF(oracle_cold_start)(n_rows, n_cols, p, q, x, phi1, phi2, a, b, alpha);
    double aux[4] = {0, 0, 0, 0};
    double residual = 1e9, objective = 0.0;
    int converged = 0;
    int k = 0;
    while (k < max_iters) {
        if (ORF(oracle_qr_step)(k, n_rows, n_cols, x, c, p, q, phi1, phi2, a, b,
                                alpha, step, kappa, aux, nthreads) != 0)
            return -1;
        if (k % 2 == 0) {
            residual = sqrt(aux[2] + aux[3]);
            objective = aux[0];
            if (residual < (double)eps) {
                ORF(oracle_fix_x)(n_rows, n_cols, x, c, step);
                converged = 1;
                break;
            }
        }
        k++;
    }
    *n_iter_out = k + 1;
    out[0] = residual;
    out[1] = objective;
    out[2] = (double)converged;
    return 0;
}

/* GL solve loop, core/drot_glr.hpp:242-355: the stopping test runs every
 * iteration, residual <= eps; n_iter counts executed iterations.
 * out = {residual, objective, converged}. */
int ORF(oracle_gl_solve)(int n_rows, int n_cols, int n_groups, const int *bounds,
                         const R *c, const R *p, const R *q, R step, R lambda,
                         int max_iters, R eps, int warm_start, R *x, R *phi1,
                         R *phi2, R *a, R *b, R *alpha, int *n_iter_out,
                         double *out, int nthreads) {
    if (warm_start)
        ORF(oracle_warm_start)(n_rows, n_cols, step, x, phi1, phi2, a, b, alpha);
    else
        ORF(oracle_cold_start)(n_rows, n_cols, p, q, x, phi1, phi2, a, b, alpha);
    double aux[4] = {0, 0, 0, 0};
    double residual = 1e9, objective = 0.0;
    int converged = 0;
    int n_iter = 0;
    while (n_iter < max_iters) {
        if (ORF(oracle_gl_step)(n_rows, n_cols, n_groups, bounds, x, c, p, q, phi1,
                                phi2, a, b, alpha, step, lambda, aux,
                                nthreads) != 0)
            return -1;
        residual = sqrt(aux[2] + aux[3]);
        objective = aux[0];
        n_iter++;
        if (residual <= (double)eps) { converged = 1; break; }
    }
    *n_iter_out = n_iter;
    out[0] = residual;
    out[1] = objective;
    out[2] = (double)converged;
    return 0;
}

#undef ORF
