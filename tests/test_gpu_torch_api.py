"""The reference's own test suite (pytorch/test.py:11-90) run against the
drop-in `reg_drot` module: same seed, same problem construction, same cases,
same acceptance check (primal residual of the returned plan <= eps), plus the
error behaviour of the binding (pytorch/src/drot_cuda.cpp:22-24,
qr_drot_cuda_kernel.cu:92-93)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OT_TESTSETS = [
    (2, 3, 2.0, 1e-2, 1e-4, 20000),
    (10, 20, 2.0, 1e-3, 1e-4, 20000),
    (100, 100, 2.0, 1e-2, 1e-4, 20000),
    (20, 10, 2.0, 10., 1e-4, 20000),
    (5000, 5000, 2.0, 1e-5, 1e-4, 20000),
    (1000, 5000, 2.0, 1e-5, 1e-4, 20000),
    (5000, 1000, 2.0, 0.0, 1e-4, 20000),
]

DA_TESTSETS = [
    (2, 4, 2, 2.0, 1e-2, 1e-4, 20000),
    (10, 20, 2, 2.0, 1e-3, 1e-4, 20000),
    (100, 100, 5, 2.0, 1e-2, 1e-4, 20000),
    (20, 10, 2, 2.0, 10., 1e-4, 20000),
    (5000, 5000, 10, 2.0, 1e-5, 1e-4, 20000),
    (1000, 5000, 5, 2.0, 1e-5, 1e-4, 20000),
    (5000, 1000, 2, 2.0, 1e-5, 1e-4, 20000),
]


def _problem(torch, n, m, device):
    p = torch.ones([n], dtype=torch.float32) / n
    q = torch.ones([m], dtype=torch.float32) / m
    c = torch.rand([m, n], dtype=torch.float32) + 1e-3
    c = c / torch.max(c)
    return c.to(device), p.to(device), q.to(device)


def _residual(torch, x, p, q):
    return float(torch.sqrt(torch.sum(torch.square(p - torch.sum(x, dim=0))) +
                            torch.sum(torch.square(q - torch.sum(x, dim=1)))))


def test_reference_test_suite(device):
    import torch
    import reg_drot
    torch.manual_seed(12345)
    for n, m, rho, r_weight, eps, max_iter in OT_TESTSETS:
        c, p, q = _problem(torch, n, m, device)
        print(f"[torch api qr] N={n} M={m} r_weight={r_weight}", flush=True)
        x, info = reg_drot.quadratic_drot_ex(c, p, q, rho, r_weight, max_iter, eps)
        assert x.shape == c.shape and x.dtype == torch.float32 and x.device == c.device
        res = _residual(torch, x, p, q)
        assert info["converged"], (n, m, info)
        # fp32 re-summation of the returned plan may differ from the fp64
        # residual the solver tested by a few fp32 ulps
        assert res <= eps * 1.01, (n, m, res, info)
    for n, m, n_class, rho, r_weight, eps, max_iter in DA_TESTSETS:
        c, p, q = _problem(torch, n, m, device)
        print(f"[torch api gl] N={n} M={m} C={n_class}", flush=True)
        x, info = reg_drot.group_lasso_drot_ex(c, p, q, n_class, rho, r_weight, max_iter, eps)
        res = _residual(torch, x, p, q)
        assert info["converged"], (n, m, n_class, info)
        assert res <= eps * 1.01, (n, m, n_class, res, info)


def test_plain_functions_return_plan_only(device):
    import torch
    import reg_drot
    torch.manual_seed(1)
    c, p, q = _problem(torch, 64, 32, device)
    x = reg_drot.quadratic_drot(c, p, q, 2.0, 1e-2, 1000, 1e-4)
    assert isinstance(x, torch.Tensor) and x.shape == (32, 64)
    x = reg_drot.group_lasso_drot(c, p, q, 2, 2.0, 1e-2, 1000, 1e-4)
    assert isinstance(x, torch.Tensor) and x.shape == (32, 64)


def test_runs_on_current_stream(device):
    import torch
    import reg_drot
    torch.manual_seed(2)
    c, p, q = _problem(torch, 300, 200, device)
    ref = reg_drot.quadratic_drot(c, p, q, 2.0, 1e-3, 500, 1e-4)
    s = torch.cuda.Stream(device)
    with torch.cuda.stream(s):
        x = reg_drot.quadratic_drot(c, p, q, 2.0, 1e-3, 500, 1e-4)
    s.synchronize()
    assert torch.equal(x, ref)


def test_binding_errors(device):
    import torch
    import reg_drot
    c = torch.rand(3, 2, device=device)
    p = torch.ones(2, device=device) / 2
    q = torch.ones(3, device=device) / 3
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        reg_drot.quadratic_drot(c.cpu(), p, q, 2.0, 0.1, 10, 1e-4)
    with pytest.raises(RuntimeError, match="must be contiguous"):
        reg_drot.quadratic_drot(torch.rand(2, 3, device=device).t(), p, q, 2.0, 0.1, 10, 1e-4)
    with pytest.raises(RuntimeError, match="does not match with q.numel"):
        reg_drot.quadratic_drot(c, p, torch.ones(4, device=device), 2.0, 0.1, 10, 1e-4)
    with pytest.raises(RuntimeError, match="does not match with p.numel"):
        reg_drot.group_lasso_drot(c, torch.ones(5, device=device), q, 2, 2.0, 0.1, 10, 1e-4)
    with pytest.raises(RuntimeError, match="float32"):
        reg_drot.quadratic_drot(c.double(), p, q, 2.0, 0.1, 10, 1e-4)


def test_deterministic(device):
    """Fixed-order fp64 reductions: two runs give bit-identical plans."""
    import torch
    import reg_drot
    torch.manual_seed(3)
    c, p, q = _problem(torch, 1000, 1000, device)
    a = reg_drot.quadratic_drot(c, p, q, 2.0, 1e-3, 2000, 1e-4)
    b = reg_drot.quadratic_drot(c, p, q, 2.0, 1e-3, 2000, 1e-4)
    assert torch.equal(a, b)
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
