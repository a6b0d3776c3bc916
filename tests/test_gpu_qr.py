"""GPU parity of QR-DROT against the CPU oracle, through the C ABI (libdrot.so).

Bar (written per test): the fp32 trajectory is reproduced to
max|X_gpu - X_oracle| <= 1e-5 * max|X_oracle| after the same number of
iterations, with identical n_iter for full solves.  Both sides accumulate every
reduction in fp64 and apply the same fp32 element operations, so in practice
the trajectories agree to the last bit or within a few ulps.
"""
import numpy as np
import pytest

import oracle as O
from reg_drot import problems

pytestmark = pytest.mark.gpu

PLAN_RTOL = 1e-5


def _abi():
    from reg_drot import _abi
    return _abi


def gpu_qr(c_np, p_np, q_np, prm, device):
    """Run drot_qr_solve_f32 on device copies; return x, result, phi1, phi2, alpha."""
    import torch
    abi = _abi()
    n_cols, n_rows = c_np.shape
    c = torch.from_numpy(np.ascontiguousarray(c_np)).to(device)
    p = torch.from_numpy(np.ascontiguousarray(p_np)).to(device)
    q = torch.from_numpy(np.ascontiguousarray(q_np)).to(device)
    x = torch.empty_like(c)
    wsb = abi.qr_workspace_size(n_rows, n_cols)
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    stream = torch.cuda.current_stream(device).cuda_stream
    r = abi.qr_solve(c.data_ptr(), p.data_ptr(), q.data_ptr(), n_rows, n_cols, prm,
                     x.data_ptr(), ws.data_ptr(), wsb, stream)
    phi1 = torch.empty(n_rows, dtype=torch.float32, device=device)
    phi2 = torch.empty(n_cols, dtype=torch.float32, device=device)
    alpha = abi.export_duals(0, n_rows, n_cols, 1, ws.data_ptr(), wsb, phi1.data_ptr(),
                             phi2.data_ptr(), stream)
    torch.cuda.synchronize(device)
    return x.cpu().numpy(), r, phi1.cpu().numpy(), phi2.cpu().numpy(), alpha


def rel_err(a, b):
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) / scale


SHAPES = [(2, 3), (10, 20), (20, 10), (64, 64), (100, 100), (257, 129), (130, 300),
          (1, 50), (50, 1), (1000, 1000), (1500, 1000)]


@pytest.mark.parametrize("n_rows,n_cols", SHAPES)
@pytest.mark.parametrize("iters", [1, 2, 7, 40])
def test_qr_trajectory_matches_oracle(device, n_rows, n_cols, iters):
    """Same iterates after `iters` iterations (eps=0 so the loop runs them all)."""
    c = problems.reference_test_cost(n_rows, n_cols, seed=n_rows * 7 + n_cols)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    abi = _abi()
    prm = abi.qr_params_from_rho(2.0, 1e-2, n_rows, n_cols, iters, 0.0)
    print(f"[qr traj] {n_rows}x{n_cols} iters={iters}", flush=True)
    x, r, phi1, phi2, alpha = gpu_qr(c, p, q, prm, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), iters, 0.0)
    assert r.iterations_run == iters
    assert r.n_iter == ref.n_iter == iters + 1
    assert rel_err(x, ref.x) <= PLAN_RTOL
    assert rel_err(phi1, ref.state.phi1) <= PLAN_RTOL
    assert rel_err(phi2, ref.state.phi2) <= PLAN_RTOL
    assert abs(alpha - ref.state.alpha) <= 1e-5 * max(1.0, abs(ref.state.alpha))


@pytest.mark.parametrize("n_rows,n_cols,r_weight", [
    (2, 3, 1e-2), (10, 20, 1e-3), (100, 100, 1e-2), (20, 10, 10.0), (257, 129, 1e-3),
    (1000, 1000, 1e-3), (1000, 1000, 1e-4)])
def test_qr_full_solve_matches_oracle(device, n_rows, n_cols, r_weight):
    """Full solve to 1e-4: same n_iter, plan within 1e-5 relative, same objective."""
    c = problems.reference_test_cost(n_rows, n_cols, seed=12345 + n_rows)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    abi = _abi()
    prm = abi.qr_params_from_rho(2.0, r_weight, n_rows, n_cols, 20000, 1e-4)
    print(f"[qr solve] {n_rows}x{n_cols} r_weight={r_weight}", flush=True)
    x, r, *_ = gpu_qr(c, p, q, prm, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 20000,
                     np.float32(1e-4))
    assert r.converged == int(ref.converged) == 1
    assert r.n_iter == ref.n_iter
    assert rel_err(x, ref.x) <= PLAN_RTOL
    assert abs(r.objective_f64 - ref.objective) <= 1e-6 * max(1e-12, abs(ref.objective))
    assert abs(r.residual - ref.residual) <= 1e-3 * ref.residual + 1e-12
    # the reference test's own acceptance check (pytorch/test.py:53-59)
    xs = x.astype(np.float64)
    res = np.sqrt(((p - xs.sum(0)) ** 2).sum() + ((q - xs.sum(1)) ** 2).sum())
    assert res <= 1e-4 * 1.05


def test_qr_cold_start_matches_oracle(device):
    """Cold start X = p q^T (core/drot_qr.hpp:308-320) through the core-convention params."""
    n_rows, n_cols = 300, 200
    c = problems.reference_test_cost(n_rows, n_cols, seed=3)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    abi = _abi()
    step = np.float32(2.0) / np.float32(n_rows + n_cols)
    prm = abi.qr_params_from_step(float(step), 1e-3, n_rows, n_cols, 30, 0.0, warm=False)
    x, r, phi1, phi2, alpha = gpu_qr(c, p, q, prm, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 30, 0.0,
                     warm=False)
    assert rel_err(x, ref.x) <= PLAN_RTOL
    assert rel_err(phi1, ref.state.phi1) <= PLAN_RTOL


def test_qr_odd_max_iters_returns_shifted_iterate(device):
    """An odd max_iters ends on an even iteration without fix_x, as in the reference."""
    n_rows, n_cols = 64, 48
    c = problems.reference_test_cost(n_rows, n_cols, seed=5)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi().qr_params_from_rho(2.0, 1e-2, n_rows, n_cols, 5, 0.0)
    x, r, *_ = gpu_qr(c, p, q, prm, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 5, 0.0)
    assert r.n_iter == 6 and not r.converged
    assert rel_err(x, ref.x) <= PLAN_RTOL
    assert (x < 0).any()  # shifted by -sC


@pytest.mark.parametrize("n", [3000])
def test_qr_large_first_iterations(device, n):
    """At the benchmark size: 6 iterations on the cloud cost, oracle in fp32."""
    c = problems.cloud_cost(n, n, dim=16, seed=0)
    p, q = problems.uniform_marginals(n, n)
    prm = _abi().qr_params_from_rho(2.0, 1e-3, n, n, 6, 0.0)
    x, r, phi1, phi2, alpha = gpu_qr(c, p, q, prm, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 6, 0.0)
    assert rel_err(x, ref.x) <= PLAN_RTOL
    assert rel_err(phi1, ref.state.phi1) <= PLAN_RTOL


@pytest.mark.parametrize("n_rows", [256, 255])
@pytest.mark.parametrize("r_weight", [1e-4, 1e-1, 30.0])
def test_qr_first_iteration_bit_exact(device, r_weight, n_rows):
    """The first iteration involves no reduction (the warm start fixes phi), so
    its element arithmetic must give the oracle's bits exactly: the prox
    max(X + phi1 + phi2 - sC, 0)/scale with a correctly rounded division, then
    the stored X' - sC.  Costs of both signs spanning 1e-3..3e38 drive the
    quotients over the whole exponent range, including above 2^120."""
    n_cols = 96
    rng = np.random.default_rng(11)
    mag = 10.0 ** rng.uniform(-3.0, 38.5, size=(n_cols, n_rows))
    c = (mag * rng.choice([-1.0, 1.0], size=mag.shape)).astype(np.float32)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi().qr_params_from_rho(2.0, r_weight, n_rows, n_cols, 1, 0.0)
    x, r, *_ = gpu_qr(c, p, q, prm, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 1, 0.0)
    assert r.iterations_run == 1
    want = np.asarray(ref.x, dtype=np.float32)
    assert np.isfinite(want).all()
    assert np.array_equal(x.view(np.uint32), want.view(np.uint32))


def test_qr_int_max_iters(device):
    """max_iters = INT_MAX (the reference's `while (k < max_iters)` allows it):
    the solve runs to convergence with the oracle's n_iter."""
    n_rows, n_cols = 120, 90
    c = problems.reference_test_cost(n_rows, n_cols, seed=21)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    big = 2 ** 31 - 1
    prm = _abi().qr_params_from_rho(2.0, 1e-2, n_rows, n_cols, big, 1e-4)
    x, r, *_ = gpu_qr(c, p, q, prm, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 20000,
                     np.float32(1e-4))
    assert r.converged == 1 and ref.converged
    assert r.n_iter == ref.n_iter
    assert rel_err(x, ref.x) <= PLAN_RTOL
