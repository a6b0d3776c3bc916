"""Implicit-differentiation backward on the GPU (drot_qr_backward_f32, the batch
variant, and the torch.autograd wrappers) against the CPU restatement
(oracle/implicit.py) and fp64 finite differences.

Bars: same plan in, the GPU gradient matches the dense fp64 lstsq gradient to
max|d| <= 1e-5 max|g|; batched == single; the autograd gradient of
<W, X*> + F matches central differences of fp64 oracle solves to 1e-4
(relative 2-norm over the checked coordinates)."""
import numpy as np
import pytest

import oracle as O
from oracle import implicit as I
from reg_drot import problems

pytestmark = pytest.mark.gpu


def _maxrel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


def _tensors(c, p, q, device):
    import torch
    return (torch.from_numpy(np.ascontiguousarray(c, np.float32)).to(device),
            torch.from_numpy(np.ascontiguousarray(p, np.float32)).to(device),
            torch.from_numpy(np.ascontiguousarray(q, np.float32)).to(device))


CASES = [(64, 48, 1e-2), (257, 129, 1e-2), (130, 7, 1e-1), (1, 50, 1e-2), (50, 1, 1e-2),
         (300, 200, 1e-3), (96, 2048, 1e-2), (513, 511, 1e-4)]


@pytest.mark.parametrize("n_rows,n_cols,rw", CASES)
def test_backward_matches_oracle(n_rows, n_cols, rw, device):
    import torch
    import reg_drot
    print(f"[bwd] {n_rows}x{n_cols} rw={rw}", flush=True)
    c, p, q = problems.reference_test_cost(n_rows, n_cols, seed=n_rows + n_cols), \
        *problems.uniform_marginals(n_rows, n_cols)
    ct, pt, qt = _tensors(c, p, q, device)
    x, info = reg_drot.quadratic_drot_ex(ct, pt, qt, 2.0, rw, 50000, 1e-6)
    g = np.random.default_rng(n_rows).standard_normal(c.shape).astype(np.float32)
    gt = torch.from_numpy(g).to(device)
    gobj = torch.tensor(0.5, device=device)
    gc, binfo = reg_drot._C.quadratic_drot_backward(ct, pt, qt, x, gt, gobj, 2.0, rw)
    torch.cuda.synchronize(device)
    xn = x.cpu().numpy()
    s, k = info["step_size"], info["scale"]
    S = I.support(xn, c, s)
    assert binfo["nnz"] == int(S.sum())
    assert binfo["cg_converged"], binfo
    want = I.implicit_grad_c(xn, c, g, s, k, gobj=0.5, S=S)
    assert _maxrel(gc.cpu().numpy(), want) <= 1e-5, binfo


def test_backward_without_objective_gradient(device):
    import torch
    import reg_drot
    c = problems.reference_test_cost(80, 60, seed=1)
    p, q = problems.uniform_marginals(80, 60)
    ct, pt, qt = _tensors(c, p, q, device)
    x, info = reg_drot.quadratic_drot_ex(ct, pt, qt, 2.0, 1e-2, 20000, 1e-6)
    g = np.random.default_rng(2).standard_normal(c.shape).astype(np.float32)
    gc, _ = reg_drot._C.quadratic_drot_backward(ct, pt, qt, x, torch.from_numpy(g).to(device),
                                                None, 2.0, 1e-2)
    xn = x.cpu().numpy()
    want = I.implicit_grad_c(xn, c, g, info["step_size"], info["scale"])
    out = gc.cpu().numpy()
    assert _maxrel(out, want) <= 1e-5
    assert np.all(out[~I.support(xn, c, info["step_size"])] == 0.0)


def test_batch_backward_matches_single(device):
    import torch
    import reg_drot
    from reg_drot import batch
    shapes = [(64, 48), (130, 70), (7, 300), (257, 129), (1, 40), (200, 1)]
    rws = [1e-2, 1e-3, 1e-2, 5e-3, 1e-2, 1e-2]
    cs, ps, qs = [], [], []
    for i, (n, m) in enumerate(shapes):
        c, p, q = _tensors(problems.reference_test_cost(n, m, seed=40 + i),
                           *problems.uniform_marginals(n, m), device)
        cs.append(c)
        ps.append(p)
        qs.append(q)
    for c in cs:
        c.requires_grad_(True)
    plans, objs = reg_drot.quadratic_drot_batch_diff(cs, ps, qs, 2.0, rws, 20000, 1e-6)
    rng = np.random.default_rng(5)
    ws = [torch.from_numpy(rng.standard_normal(tuple(c.shape)).astype(np.float32)).to(device)
          for c in cs]
    loss = sum((w * x).sum() for w, x in zip(ws, plans)) + 0.25 * objs.sum()
    loss.backward()
    for i, (c, p, q) in enumerate(zip(cs, ps, qs)):
        ci = c.detach().clone().requires_grad_(True)
        x, F = reg_drot.quadratic_drot_diff(ci, p, q, 2.0, rws[i], 20000, 1e-6,
                                            return_objective=True)
        assert torch.equal(x, plans[i].detach()), i
        assert abs(F.item() - objs[i].item()) <= 1e-12 * max(1.0, abs(F.item()))
        ((ws[i] * x).sum() + 0.25 * F).backward()
        assert _maxrel(c.grad.cpu().numpy(), ci.grad.cpu().numpy()) <= 1e-6, shapes[i]


@pytest.mark.parametrize("n_rows,n_cols,rw,seed", [(24, 18, 3e-2, 1), (40, 32, 1e-2, 2),
                                                   (33, 47, 1e-1, 3)])
def test_autograd_matches_finite_differences(n_rows, n_cols, rw, seed, device):
    """dL/dC of L = <W, X*> + F from a tightly solved fp32 forward, against fp64
    central differences (h = 1e-6) of the oracle on 10 coordinates."""
    import torch
    import reg_drot
    rng = np.random.default_rng(seed)
    c = problems.reference_test_cost(n_rows, n_cols, seed=seed)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    ct, pt, qt = _tensors(c, p, q, device)
    ct.requires_grad_(True)
    x, F = reg_drot.quadratic_drot_diff(ct, pt, qt, 2.0, rw, 200000, 1e-7,
                                        return_objective=True)
    w = rng.standard_normal(c.shape)
    ((torch.from_numpy(w.astype(np.float32)).to(device) * x).sum() + F).backward()
    grad = ct.grad.cpu().numpy()
    s, k = O.qr_params_torch(2.0, rw, n_rows, n_cols)
    x64 = I.tight_solve(c.astype(np.float64), p, q, s, k, iters=30000)
    sj, si = np.nonzero(x64 > 1e-9 * x64.max())
    pick = rng.choice(len(sj), 7, replace=False)
    coords = [(int(sj[t]), int(si[t])) for t in pick]
    coords += [(int(rng.integers(n_cols)), int(rng.integers(n_rows))) for _ in range(3)]
    fd = I.fd_grad(c.astype(np.float64), p, q, s, k, w, 1.0, coords, h=1e-6, iters=30000)
    got = np.array([grad[j, i] for j, i in coords])
    rel = float(np.linalg.norm(got - fd) / np.linalg.norm(fd))
    print(f"[fd] {n_rows}x{n_cols} rel={rel:.2e}", flush=True)
    assert rel <= 1e-4


def test_backward_errors(device):
    import torch
    import reg_drot
    c, p, q = _tensors(problems.reference_test_cost(30, 20, seed=0),
                       *problems.uniform_marginals(30, 20), device)
    x = reg_drot.quadratic_drot(c, p, q, 2.0, 1e-2, 5000, 1e-5)
    g = torch.ones_like(c)
    with pytest.raises(RuntimeError, match="r_weight > 0"):
        reg_drot._C.quadratic_drot_backward(c, p, q, x, g, None, 2.0, 0.0)
    with pytest.raises(RuntimeError, match="shape"):
        reg_drot._C.quadratic_drot_backward(c, p, q, x[:, :10].contiguous(), g, None, 2.0, 1e-2)
    from reg_drot import _abi, batch
    pb = batch.pack([c], [p], [q], 2.0, 1e-2, 5000, 1e-5)
    pb.solve()
    bp = _abi.backward_params(nnz_capacity=3)
    nbytes = _abi.batch_backward_workspace_size(pb.probs, bp)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    gx = torch.ones_like(pb.x)
    gc = torch.empty_like(pb.x)
    with pytest.raises(_abi.DrotError, match="capacity"):
        _abi.batch_backward(pb.probs, pb.c.data_ptr(), pb.x.data_ptr(), gx.data_ptr(), 0, bp,
                            gc.data_ptr(), ws.data_ptr(), nbytes,
                            torch.cuda.current_stream(device).cuda_stream)
