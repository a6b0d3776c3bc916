"""Batched QR-DROT (drot_qr_batch_solve_f32) against the single-problem solver
and the oracle: every problem of a mixed batch must give the plan, n_iter,
convergence flag and objective of solving it alone (fp64 reductions, same
per-element arithmetic; bar: max|dX| <= 1e-6 max|X|, identical n_iter)."""
import numpy as np
import pytest

import oracle as O
from reg_drot import problems

pytestmark = pytest.mark.gpu


def _single(c_np, p_np, q_np, prm, device):
    import torch
    from reg_drot import _abi
    n_cols, n_rows = c_np.shape
    c = torch.from_numpy(c_np).to(device)
    p = torch.from_numpy(p_np).to(device)
    q = torch.from_numpy(q_np).to(device)
    x = torch.empty_like(c)
    wsb = _abi.qr_workspace_size(n_rows, n_cols)
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    r = _abi.qr_solve(c.data_ptr(), p.data_ptr(), q.data_ptr(), n_rows, n_cols, prm,
                      x.data_ptr(), ws.data_ptr(), wsb,
                      torch.cuda.current_stream(device).cuda_stream)
    torch.cuda.synchronize(device)
    return x.cpu().numpy(), r


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def _run_batch(shapes, r_weights, max_iters, eps, device, seed=0):
    import torch
    from reg_drot import batch
    cs, ps, qs, nps = [], [], [], []
    for i, (n, m) in enumerate(shapes):
        c = problems.reference_test_cost(n, m, seed=seed + 31 * i)
        p, q = problems.uniform_marginals(n, m)
        nps.append((c, p, q))
        cs.append(torch.from_numpy(c).to(device))
        ps.append(torch.from_numpy(p).to(device))
        qs.append(torch.from_numpy(q).to(device))
    pb = batch.pack(cs, ps, qs, 2.0, r_weights, max_iters, eps)
    res = pb.solve()
    torch.cuda.synchronize(device)
    plans = [pb.plan(i).cpu().numpy() for i in range(len(shapes))]
    return nps, plans, res, pb


SHAPES = [(2, 3), (64, 64), (130, 7), (7, 130), (1, 50), (50, 1), (257, 129), (300, 200),
          (1000, 800), (96, 2048), (2048, 96), (513, 511)]


def test_batch_matches_single_solves(device):
    rws = [1e-2, 1e-3, 1e-2, 1e-1, 1e-2, 1e-2, 1e-3, 5e-3, 1e-3, 1e-2, 1e-2, 1e-3]
    print(f"[batch] {len(SHAPES)} problems", flush=True)
    nps, plans, res, pb = _run_batch(SHAPES, rws, 20000, 1e-4, device)
    from reg_drot import _abi
    for i, ((c, p, q), xb, r) in enumerate(zip(nps, plans, res)):
        prm = _abi.qr_params_from_rho(2.0, rws[i], len(p), len(q), 20000, 1e-4)
        xs, rs = _single(c, p, q, prm, device)
        assert r.n_iter == rs.n_iter and r.converged == rs.converged, (i, r.n_iter, rs.n_iter)
        assert r.iterations_run == rs.iterations_run
        assert _rel(xb, xs) <= 1e-6, (i, SHAPES[i])
        assert abs(r.objective_f64 - rs.objective_f64) <= 1e-9 * max(1.0, abs(rs.objective_f64))


def test_batch_trajectories_match_oracle(device):
    """Fixed iteration counts (eps = 0) per problem, against the fp32 oracle."""
    shapes = [(64, 48), (130, 70), (257, 129), (33, 200)]
    from reg_drot import _abi
    for iters in (1, 6, 25):
        nps, plans, res, pb = _run_batch(shapes, 1e-2, iters, 0.0, device, seed=5)
        for i, ((c, p, q), xb, r) in enumerate(zip(nps, plans, res)):
            prm = _abi.qr_params_from_rho(2.0, 1e-2, len(p), len(q), iters, 0.0)
            ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), iters, 0.0)
            assert r.iterations_run == iters and r.n_iter == ref.n_iter
            assert _rel(xb, ref.x) <= 1e-5, (iters, shapes[i])


def test_batch_per_problem_limits(device):
    """Problems stop independently: one hits its own max_iters, others converge."""
    import torch
    from reg_drot import _abi, batch
    shapes = [(100, 80), (120, 90), (64, 64)]
    cs, ps, qs, nps = [], [], [], []
    for i, (n, m) in enumerate(shapes):
        c = problems.reference_test_cost(n, m, seed=70 + i)
        p, q = problems.uniform_marginals(n, m)
        nps.append((c, p, q))
        cs.append(torch.from_numpy(c).to(device))
        ps.append(torch.from_numpy(p).to(device))
        qs.append(torch.from_numpy(q).to(device))
    pb = batch.pack(cs, ps, qs, 2.0, 1e-2, 20000, 1e-4)
    pb.probs[1].params = _abi.qr_params_from_rho(2.0, 1e-2, 120, 90, 17, 1e-4)
    res = pb.solve()
    torch.cuda.synchronize(device)
    assert res[1].n_iter == 18 and not res[1].converged and res[1].iterations_run == 17
    assert res[0].converged and res[2].converged
    c, p, q = nps[1]
    ref = O.qr_solve(c, p, q, np.float32(pb.probs[1].params.step_size),
                     np.float32(pb.probs[1].params.scale), 17, np.float32(1e-4))
    assert _rel(pb.plan(1).cpu().numpy(), ref.x) <= 1e-5


def test_batch_random_mixed_sizes(device):
    """A config-4-like batch (m, n uniform in [64, 700], log-uniform alpha)."""
    rng = np.random.default_rng(11)
    shapes = [(int(a), int(b)) for a, b in rng.integers(64, 701, size=(24, 2))]
    rws = list(10 ** rng.uniform(-4, -2, size=24))
    nps, plans, res, pb = _run_batch(shapes, rws, 20000, 1e-4, device, seed=9)
    from reg_drot import _abi
    for i in range(0, 24, 3):  # spot-check every third against its single solve
        c, p, q = nps[i]
        prm = _abi.qr_params_from_rho(2.0, float(rws[i]), len(p), len(q), 20000, 1e-4)
        xs, rs = _single(c, p, q, prm, device)
        assert res[i].n_iter == rs.n_iter
        assert _rel(plans[i], xs) <= 1e-6
    assert all(r.converged for r in res)


def test_batch_torch_api(device):
    import torch
    import reg_drot
    from reg_drot import batch
    torch.manual_seed(4)
    cs, ps, qs = [], [], []
    for n, m in [(40, 30), (200, 150)]:
        c = torch.rand(m, n, device=device) + 1e-3
        cs.append(c / c.max())
        ps.append(torch.full((n,), 1.0 / n, device=device))
        qs.append(torch.full((m,), 1.0 / m, device=device))
    plans, infos = batch.quadratic_drot_batch(cs, ps, qs, 2.0, 1e-2, 5000, 1e-4)
    for c, p, q, xb, info in zip(cs, ps, qs, plans, infos):
        xs = reg_drot.quadratic_drot(c, p, q, 2.0, 1e-2, 5000, 1e-4)
        assert torch.equal(xb, xs)
        assert info["converged"]
    with pytest.raises(RuntimeError, match="CUDA"):
        batch.pack([cs[0].cpu()], [ps[0]], [qs[0]], 2.0, 1e-2, 10, 1e-4)
