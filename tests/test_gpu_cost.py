"""Point-cloud cost kernels (drot_sqdist_f32 / drot_sqdist_backward_f32) and the
differentiable points -> C -> plan -> loss pipeline.

Forward against fp64 numpy distances (bar: max|d| <= 2e-6 max C); backward
against torch fp64 autograd of the same expression on the CPU (bar 1e-5); the
whole pipeline against fp64 central differences of oracle solves (bar 1e-4)."""
import numpy as np
import pytest

import oracle as O
from oracle import implicit as I

pytestmark = pytest.mark.gpu


def _ref_cost(x, y, normalize):
    d = ((y[:, None, :].astype(np.float64) - x[None, :, :]) ** 2).sum(-1)
    return d / d.max() if normalize and d.max() > 0 else d


SHAPES = [(1, 1, 3), (7, 130, 2), (64, 64, 16), (65, 63, 8), (300, 200, 16), (129, 70, 33),
          (50, 40, 64), (200, 3, 1)]


@pytest.mark.parametrize("n,m,dim", SHAPES)
@pytest.mark.parametrize("normalize", [True, False])
def test_sqdist_forward(n, m, dim, normalize, device):
    import torch
    from reg_drot import cost
    rng = np.random.default_rng(n * 7 + m)
    x = rng.standard_normal((n, dim)).astype(np.float32)
    y = rng.standard_normal((m, dim)).astype(np.float32)
    c = cost.sqdist(torch.from_numpy(x).to(device), torch.from_numpy(y).to(device), normalize)
    torch.cuda.synchronize(device)
    want = _ref_cost(x, y, normalize)
    got = c.cpu().numpy()
    assert got.shape == (m, n)
    assert np.abs(got - want).max() <= 2e-6 * max(want.max(), 1e-30)


@pytest.mark.parametrize("n,m,dim", SHAPES)
@pytest.mark.parametrize("normalize", [True, False])
def test_sqdist_backward(n, m, dim, normalize, device):
    import torch
    from reg_drot import cost
    rng = np.random.default_rng(n + 3 * m + dim)
    x = rng.standard_normal((n, dim)).astype(np.float32)
    y = rng.standard_normal((m, dim)).astype(np.float32)
    g = rng.standard_normal((m, n)).astype(np.float32)
    xt = torch.from_numpy(x).to(device).requires_grad_(True)
    yt = torch.from_numpy(y).to(device).requires_grad_(True)
    (cost.sqdist(xt, yt, normalize) * torch.from_numpy(g).to(device)).sum().backward()
    xr = torch.from_numpy(x).double().requires_grad_(True)
    yr = torch.from_numpy(y).double().requires_grad_(True)
    d = ((yr[:, None, :] - xr[None, :, :]) ** 2).sum(-1)
    if normalize:
        d = d / d.max()
    (d * torch.from_numpy(g).double()).sum().backward()
    for got, want in ((xt.grad, xr.grad), (yt.grad, yr.grad)):
        got = got.cpu().numpy()
        want = want.numpy()
        assert np.abs(got - want).max() <= 1e-5 * max(np.abs(want).max(), 1e-30)


def test_sqdist_packed_layout(device):
    import torch
    from reg_drot import batch, cost
    shapes = [(5, 9), (64, 64), (130, 7), (33, 70)]
    rng = np.random.default_rng(3)
    xs = [torch.from_numpy(rng.standard_normal((n, 8)).astype(np.float32)).to(device)
          for n, _ in shapes]
    ys = [torch.from_numpy(rng.standard_normal((m, 8)).astype(np.float32)).to(device)
          for _, m in shapes]
    ps = [torch.full((n,), 1.0 / n, device=device) for n, _ in shapes]
    qs = [torch.full((m,), 1.0 / m, device=device) for _, m in shapes]
    pb = batch.layout(ps, qs, 2.0, 1e-2, 100, 1e-4)
    cp = cost.sqdist_packed(xs, ys, pb)
    for i in range(len(shapes)):
        assert torch.equal(pb.block(cp, i), cost.sqdist(xs[i], ys[i]))
        pr = pb.probs[i]
        full = cp[pr.mat_offset: pr.mat_offset + pr.ld * pr.n_cols].view(pr.n_cols, pr.ld)
        assert torch.all(full[:, pr.n_rows:] == 0)


def test_pipeline_matches_finite_differences(device):
    """points -> normalized cost -> batched plan -> <W, X*> + F, gradient to the
    points against fp64 central differences of oracle solves."""
    import torch
    from reg_drot import batch, cost, autograd
    rng = np.random.default_rng(21)
    shapes = [(24, 20), (18, 26)]
    dim, rw = 4, 3e-2
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    xs_np = [f32(rng.standard_normal((n, dim))) for n, _ in shapes]
    ys_np = [f32(rng.standard_normal((m, dim))) for _, m in shapes]
    ws_np = [rng.standard_normal((m, n)) for n, m in shapes]
    xs = [torch.tensor(a, dtype=torch.float32, device=device, requires_grad=True) for a in xs_np]
    ys = [torch.tensor(a, dtype=torch.float32, device=device, requires_grad=True) for a in ys_np]
    ps = [torch.full((n,), 1.0 / n, device=device) for n, _ in shapes]
    qs = [torch.full((m,), 1.0 / m, device=device) for _, m in shapes]
    pb = batch.layout(ps, qs, 2.0, rw, 200000, 1e-7)
    cp = cost.sqdist_packed(xs, ys, pb)
    xp, objs = autograd.batch_solve_diff(cp, pb)
    loss = objs.sum()
    for i in range(len(shapes)):
        loss = loss + (torch.tensor(ws_np[i], dtype=torch.float32, device=device)
                       * pb.block(xp, i)).sum()
    loss.backward()

    def f64_loss(i, x, y):
        n, m = shapes[i]
        c = _ref_cost(x, y, True)
        p = np.full(n, 1.0 / n)
        q = np.full(m, 1.0 / m)
        s, k = O.qr_params_torch(2.0, rw, n, m)
        return I.loss(c, p, q, s, k, ws_np[i], 1.0, iters=30000)

    h = 1e-6
    got, fd = [], []
    for i in range(len(shapes)):
        for which in ("x", "y"):
            base = xs_np[i] if which == "x" else ys_np[i]
            grad = (xs[i] if which == "x" else ys[i]).grad.cpu().numpy()
            for _ in range(3):
                a, b = int(rng.integers(base.shape[0])), int(rng.integers(dim))
                pert = []
                for sgn in (1, -1):
                    arr = base.copy()
                    arr[a, b] += sgn * h
                    xx, yy = (arr, ys_np[i]) if which == "x" else (xs_np[i], arr)
                    pert.append(f64_loss(i, xx, yy))
                fd.append((pert[0] - pert[1]) / (2 * h))
                got.append(grad[a, b])
    got, fd = np.array(got), np.array(fd)
    rel = float(np.linalg.norm(got - fd) / np.linalg.norm(fd))
    print(f"[pipeline fd] rel={rel:.2e}", flush=True)
    assert rel <= 1e-4
