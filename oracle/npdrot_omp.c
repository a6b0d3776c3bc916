/*
 * OpenMP restatement of the reference's CPU solver (numpy/npdrot: regdrot with
 * the quadratic prox), pass for pass, for the CPU baseline.  TEST AND BASELINE
 * INFRASTRUCTURE ONLY: bench.py's CPU legs time it and tests/ check it against
 * the numpy restatement (oracle/npdrot.py); the product library never links
 * it.  npdrot itself cannot run here (it imports numba, which is absent).
 *
 * Compiled in float (the benchmark's fp32 workload) and double (npdrot's own
 * precision: its BLAS dger updates are the double routine).  See
 * npdrot_omp_impl.h for the passes and the reference lines they follow.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>

#define R float
#define SFX _f32
#include "npdrot_omp_impl.h"
#undef R
#undef SFX

#define R double
#define SFX _f64
#include "npdrot_omp_impl.h"
#undef R
#undef SFX
