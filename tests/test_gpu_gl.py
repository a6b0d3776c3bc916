"""GPU parity of GL-DROT against the CPU oracle, through the C ABI (libdrot.so).

Bar: max|X_gpu - X_oracle| <= 1e-5 * max|X_oracle| after the same number of
iterations, identical n_iter for full solves.
"""
import numpy as np
import pytest

import oracle as O
from reg_drot import problems

pytestmark = pytest.mark.gpu

PLAN_RTOL = 1e-5


def gpu_gl(c_np, p_np, q_np, prm, device):
    import torch
    from reg_drot import _abi as abi
    n_cols, n_rows = c_np.shape
    c = torch.from_numpy(np.ascontiguousarray(c_np)).to(device)
    p = torch.from_numpy(np.ascontiguousarray(p_np)).to(device)
    q = torch.from_numpy(np.ascontiguousarray(q_np)).to(device)
    x = torch.empty_like(c)
    wsb = abi.glr_workspace_size(n_rows, n_cols, prm.n_groups)
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    stream = torch.cuda.current_stream(device).cuda_stream
    r = abi.glr_solve(c.data_ptr(), p.data_ptr(), q.data_ptr(), n_rows, n_cols, prm,
                      x.data_ptr(), ws.data_ptr(), wsb, stream)
    phi1 = torch.empty(n_rows, dtype=torch.float32, device=device)
    phi2 = torch.empty(n_cols, dtype=torch.float32, device=device)
    alpha = abi.export_duals(1, n_rows, n_cols, prm.n_groups, ws.data_ptr(), wsb,
                             phi1.data_ptr(), phi2.data_ptr(), stream)
    torch.cuda.synchronize(device)
    return x.cpu().numpy(), r, phi1.cpu().numpy(), phi2.cpu().numpy(), alpha


def rel_err(a, b):
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) / scale


# (n_rows, n_cols, n_groups): the reference test's group layouts
# (pytorch/test.py:21-29), small (<= 256 rows per group) and large groups.
CASES = [(2, 4, 2), (10, 20, 2), (100, 100, 5), (20, 10, 2), (257, 64, 3),
         (1500, 1000, 2), (1000, 500, 5), (5000, 300, 10), (600, 200, 600)]


@pytest.mark.parametrize("n_rows,n_cols,n_groups", CASES)
@pytest.mark.parametrize("iters", [1, 5, 30])
def test_gl_trajectory_matches_oracle(device, n_rows, n_cols, n_groups, iters):
    c = problems.reference_test_cost(n_rows, n_cols, seed=n_rows + 3 * n_cols)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    from reg_drot import _abi as abi
    prm = abi.glr_params_from_rho(2.0, 1e-2, n_rows, n_cols, n_groups, iters, 0.0)
    print(f"[gl traj] {n_rows}x{n_cols} G={n_groups} iters={iters}", flush=True)
    x, r, phi1, phi2, alpha = gpu_gl(c, p, q, prm, device)
    ref = O.gl_solve(c, p, q, n_groups, np.float32(prm.step_size), np.float32(prm.lambda_),
                     iters, 0.0)
    assert r.iterations_run == iters == ref.n_iter
    assert rel_err(x, ref.x) <= PLAN_RTOL
    assert rel_err(phi1, ref.state.phi1) <= PLAN_RTOL
    assert rel_err(phi2, ref.state.phi2) <= PLAN_RTOL


@pytest.mark.parametrize("n_rows,n_cols,n_groups,r_weight", [
    (10, 20, 2, 1e-3), (100, 100, 5, 1e-2), (1500, 1000, 2, 1e-3)])
def test_gl_full_solve_matches_oracle(device, n_rows, n_cols, n_groups, r_weight):
    c = problems.reference_test_cost(n_rows, n_cols, seed=99 + n_rows)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    from reg_drot import _abi as abi
    prm = abi.glr_params_from_rho(2.0, r_weight, n_rows, n_cols, n_groups, 20000, 1e-4)
    print(f"[gl solve] {n_rows}x{n_cols} G={n_groups}", flush=True)
    x, r, *_ = gpu_gl(c, p, q, prm, device)
    ref = O.gl_solve(c, p, q, n_groups, np.float32(prm.step_size), np.float32(prm.lambda_),
                     20000, np.float32(1e-4))
    assert r.converged == int(ref.converged) == 1
    assert r.n_iter == ref.n_iter
    assert rel_err(x, ref.x) <= PLAN_RTOL
    assert abs(r.objective_f64 - ref.objective) <= 1e-6 * abs(ref.objective)
