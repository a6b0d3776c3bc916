"""The GPU's converged results against an FP64 oracle run.

The parity tests compare the fp32 GPU trajectory with the fp32 oracle (same
element arithmetic).  Here the reference point is the oracle in double
precision, converged to the same tolerance: the objective the GPU reports, its
plan and the residual of that plan (evaluated in fp64 from the returned X)
must agree with it to the accuracy an fp32 iteration can reach.  Bars, from
the fp32-vs-fp64 oracle differences on these problems (QR: objective up to
3.3e-7, plan up to 5.9e-5 relative, n_iter within 0.05%; GL: objective 5e-7,
plan 1e-3, n_iter within 6%): QR objective 1e-6, plan 2e-4, n_iter within 1%;
GL objective 1e-5, plan 5e-3, n_iter within 10%; the fp64 residual of the
returned plan within 1.1 x the tolerance.
"""
import numpy as np
import pytest

import oracle as O
from reg_drot import problems

pytestmark = pytest.mark.gpu

EPS = 1e-6


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / np.abs(b).max())


def _residual(x, p, q):
    x = x.astype(np.float64)
    return float(np.sqrt(((p - x.sum(0)) ** 2).sum() + ((q - x.sum(1)) ** 2).sum()))


@pytest.mark.parametrize("n_rows,n_cols,r_weight", [(200, 150, 1e-2), (300, 300, 1e-3),
                                                    (640, 480, 1e-4)])
def test_qr_converged_against_fp64_oracle(device, n_rows, n_cols, r_weight):
    from test_gpu_qr import gpu_qr
    from reg_drot import _abi
    c = problems.reference_test_cost(n_rows, n_cols, seed=41 + n_rows)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi.qr_params_from_rho(2.0, r_weight, n_rows, n_cols, 200000, EPS)
    print(f"[fp64 qr] {n_rows}x{n_cols} r_weight={r_weight}", flush=True)
    x, r, *_ = gpu_qr(c, p, q, prm, device)
    f64 = np.float64
    ref = O.qr_solve(c.astype(f64), p.astype(f64), q.astype(f64), f64(np.float32(prm.step_size)),
                     f64(np.float32(prm.scale)), 200000, EPS, dtype=f64)
    assert r.converged == 1 and ref.converged
    assert abs(r.n_iter - ref.n_iter) <= 0.01 * ref.n_iter + 2
    assert abs(r.objective_f64 - ref.objective) <= 1e-6 * abs(ref.objective)
    assert _rel(x, ref.x) <= 2e-4
    assert _residual(x, p.astype(f64), q.astype(f64)) <= 1.1 * EPS


@pytest.mark.parametrize("n_rows,n_cols,n_groups,r_weight", [(100, 100, 5, 1e-2),
                                                             (300, 200, 2, 1e-3)])
def test_gl_converged_against_fp64_oracle(device, n_rows, n_cols, n_groups, r_weight):
    from test_gpu_gl import gpu_gl
    from reg_drot import _abi
    c = problems.reference_test_cost(n_rows, n_cols, seed=7 + n_rows)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi.glr_params_from_rho(2.0, r_weight, n_rows, n_cols, n_groups, 200000, EPS)
    print(f"[fp64 gl] {n_rows}x{n_cols} G={n_groups}", flush=True)
    x, r, *_ = gpu_gl(c, p, q, prm, device)
    f64 = np.float64
    ref = O.gl_solve(c.astype(f64), p.astype(f64), q.astype(f64), n_groups,
                     f64(np.float32(prm.step_size)), f64(np.float32(prm.lambda_)), 200000, EPS,
                     dtype=f64)
    assert r.converged == 1 and ref.converged
    assert abs(r.n_iter - ref.n_iter) <= 0.1 * ref.n_iter
    assert abs(r.objective_f64 - ref.objective) <= 1e-5 * abs(ref.objective)
    assert _rel(x, ref.x) <= 5e-3
    assert _residual(x, p.astype(f64), q.astype(f64)) <= 1.1 * EPS
