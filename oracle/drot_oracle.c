/*
 * CPU oracle for the Douglas-Rachford regularized-OT hot path (QR-DROT and
 * GL-DROT).  TEST INFRASTRUCTURE ONLY: it is loaded by tests/, by
 * __graft_entry__.smoke() as the checker and by bench.py's cpu_baseline leg.
 * The product library (libdrot.so) never links or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's CUDA core and
 * PyTorch binding:
 *   core/drot_qr.hpp:52-347, core/kernel_qr.hpp:7-482          (QR-DROT)
 *   core/drot_glr.hpp:87-474, core/kernel_glr.hpp:59-526        (GL-DROT)
 *   pytorch/src/qr_drot_cuda_kernel.cu:12-78                    (step, kappa, warm start)
 *   pytorch/src/glr_drot_cuda_kernel.cu:13-86                   (step, lambda, warm start)
 * Each function cites the lines it follows.  The iteration body lives in
 * drot_oracle_impl.h and is compiled twice: in float (the trajectory the GPU
 * path reproduces) and in double (the FP64 oracle run).
 *
 * How each function is pinned (tests/test_oracle.py): closed forms for the
 * first iterations from the warm start, the KKT form of the converged QR plan,
 * an independent scipy solution of the same QP/LP on small problems, special
 * cases with known plans, and the fp32 / fp64 variants against each other.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- float instance ---------------------------------------------------- */
#define R float
#define SFX _f32
#define OR_FMA fmaf
#include "drot_oracle_impl.h"
#undef R
#undef SFX
#undef OR_FMA

/* ---- double instance --------------------------------------------------- */
#define R double
#define SFX _f64
#define OR_FMA fma
#include "drot_oracle_impl.h"
#undef R
#undef SFX
#undef OR_FMA

/* ---- parameter conventions (host-side fp32 arithmetic of the reference) -- */

/* PyTorch binding, pytorch/src/qr_drot_cuda_kernel.cu:38-40:
 *   step  = rho / (float(n_rows) + float(n_cols))
 *   kappa = 1 + step * r_weight * float(n_rows + n_cols) */
void oracle_qr_params_torch(float rho, float r_weight, int n_rows, int n_cols,
                            float *step, float *kappa) {
    const float s = rho / ((float)n_rows + (float)n_cols);
    *step = s;
    *kappa = 1 + s * r_weight * (float)(n_rows + n_cols);
}

/* Core wrapper, core/drot_qr.hpp:262 (step_size is passed in already normalised,
 * core/singleexp_qr.cu:71, core/multiexp_qr.cu:110):
 *   kappa = 1 + step * T(n_rows + n_cols) * r_weight */
float oracle_qr_kappa_core(float step, float r_weight, int n_rows, int n_cols) {
    return (1 + step * (float)(n_rows + n_cols) * r_weight);
}

/* GL lambda, PyTorch binding (pytorch/src/glr_drot_cuda_kernel.cu:47):
 *   lambda = r_weight * sqrt(n_rows / NGROUPS)   (integer division, double sqrt) */
float oracle_gl_lambda_torch(float r_weight, int n_rows, int n_groups) {
    return (float)((double)r_weight * sqrt((double)(n_rows / n_groups)));
}

/* GL lambda, core wrapper (core/drot_glr.hpp:385):
 *   lambda = r_weight / sqrt(n_rows / NGROUPS) */
float oracle_gl_lambda_core(float r_weight, int n_rows, int n_groups) {
    return (float)((double)r_weight / sqrt((double)(n_rows / n_groups)));
}

/* GL step size, pytorch/src/glr_drot_cuda_kernel.cu:45. */
float oracle_step_torch(float rho, int n_rows, int n_cols) {
    return rho / ((float)n_rows + (float)n_cols);
}

/* Group row bounds.  The reference computes them in float with two different
 * rounding offsets depending on the kernel that runs:
 *   groups of <= 256 rows (CEILDIV(n_rows, G) <= GROUP_THREASHOLD,
 *   core/drot_glr.hpp:280): int(T(n_rows)/T(G)*T(g) + 0.49)
 *   (core/kernel_glr.hpp:79-80); otherwise
 *   int(T(n_rows)/G*T(g) + 0.49999) (core/kernel_glr.hpp:185-186, :344-345).
 * bounds must hold n_groups + 1 entries. */
void oracle_group_bounds(int n_rows, int n_groups, int *bounds) {
    const int small = ((n_rows + n_groups - 1) / n_groups) <= 256;
    const double off = small ? 0.49 : 0.49999;
    const float per = (float)n_rows / (float)n_groups;
    for (int g = 0; g <= n_groups; g++) {
        const float v = per * (float)g;
        bounds[g] = (int)((double)v + off);
    }
}
