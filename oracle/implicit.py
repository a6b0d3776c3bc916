"""CPU restatement of the QR-DROT backward, for tests only.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU baseline as the checker; the product path never imports it.

The reference has no backward (pytorch/src/qr_drot_cuda_kernel.cu:99 returns a
plan without gradient), so this module restates the mathematics the HIP
backward implements, three independent ways:

  implicit_grad_c  dense fp64 implicit differentiation of the KKT conditions
                   on the support: K [w; w'] = [(S.G)1; (S.G)^T 1] solved by
                   least squares (numpy.linalg.lstsq), then
                   dL/dC = S o (w 1^T + 1 w'^T - G) / reg + g_F X;
  unrolled_plan    the DR iteration of the oracle (oracle/drot_oracle_impl.h,
                   following core/kernel_qr.hpp:50-482 and
                   core/drot_qr.hpp:52-347) written with torch fp64 operations,
                   so that autograd through many iterations gives the
                   derivative of the iterate;
  fd_grad          central finite differences of L(C) = <W, X*(C)> + g_F F(C)
                   through tight fp64 solves of the C oracle;
  newton_solve     the exact fp64 optimum by semi-smooth Newton on the dual,
                   for finite differences at sizes where DR to 1e-14 is slow.

Arrays use the solver layout: C, X, G have shape (n_cols, n_rows).
"""
from __future__ import annotations

import numpy as np

SUPPORT_TOL = 2.0 ** -22


def reg_of(step, kappa) -> float:
    """Quadratic weight of the problem the iteration solves: (kappa - 1)/step,
    evaluated from the fp32 parameters the solver uses."""
    return (float(np.float32(kappa)) - 1.0) / float(np.float32(step))


def support(x, c, step, tol=SUPPORT_TOL):
    """x_ij > 0 and x_ij > tol * step * |c_ij| (fp32, as the HIP backward)."""
    x = np.asarray(x, np.float32)
    thr = np.float32(tol * float(np.float32(step)))
    return (x > 0) & (x > thr * np.abs(np.asarray(c, np.float32)))


def laplacian(S):
    """Signless bipartite Laplacian [[D_r, S^T], [S, D_c]] of a (n_cols, n_rows)
    support, rows (p side) first."""
    St = S.T.astype(np.float64)  # (n_rows, n_cols)
    n, m = St.shape
    K = np.zeros((n + m, n + m))
    K[:n, :n] = np.diag(St.sum(1))
    K[n:, n:] = np.diag(St.sum(0))
    K[:n, n:] = St
    K[n:, :n] = St.T
    return K


def implicit_grad_c(x, c, g, step, kappa, gobj=0.0, S=None, tol=SUPPORT_TOL):
    """Dense fp64 dL/dC (shape of c) for upstream gradient g = dL/dX and
    gobj = dL/dF."""
    if S is None:
        S = support(x, c, step, tol)
    reg = reg_of(step, kappa)
    St = S.T
    Gt = np.asarray(g, np.float64).T
    n = St.shape[0]
    rhs = np.concatenate([(St * Gt).sum(1), (St * Gt).sum(0)])
    w = np.linalg.lstsq(laplacian(S), rhs, rcond=None)[0]
    grad_t = np.where(St, (w[:n, None] + w[None, n:] - Gt) / reg, 0.0)
    return grad_t.T + float(gobj) * np.asarray(x, np.float64)


def double_centering(g):
    """G - row means - column means + grand mean (the projection that the
    implicit gradient reduces to when the support is full)."""
    g = np.asarray(g, np.float64)
    return g - g.mean(0, keepdims=True) - g.mean(1, keepdims=True) + g.mean()


def unrolled_plan(c, p, q, step, kappa, iters):
    """Plan after `iters` DR iterations from the warm start, as a torch fp64
    expression of c (a torch tensor, (n_cols, n_rows)); differentiable.

    Same iteration as oracle_qr_step (X' = max(X'_prev + phi1 + phi2 - sC, 0)/kappa,
    the even/odd storage of X' - sC being an exact rewriting) and dual_update
    (gamma from alpha before its update; a, b, phi1, phi2, alpha)."""
    import torch

    n_cols, n_rows = c.shape
    dt = torch.float64
    s = float(np.float32(step))
    k = float(np.float32(kappa))
    p = torch.as_tensor(np.asarray(p, np.float64), dtype=dt)
    q = torch.as_tensor(np.asarray(q, np.float64), dtype=dt)
    n, m = float(n_rows), float(n_cols)
    kk = s * n * m - 2.0
    phi1 = torch.full((n_rows,), (kk + 2) / n / (m + n), dtype=dt)
    phi2 = torch.full((n_cols,), (kk + 2) / m / (m + n), dtype=dt)
    a = torch.full((n_rows,), (kk + 1) / n, dtype=dt)
    b = torch.full((n_cols,), (kk + 1) / m, dtype=dt)
    alpha = torch.tensor(kk + 1, dtype=dt)
    xp = torch.zeros((n_cols, n_rows), dtype=dt)
    for _ in range(iters):
        xp = torch.clamp(xp + phi1[None, :] + phi2[:, None] - s * c, min=0.0) / k
        S = xp.sum()
        gamma = (2.0 * (S - 1.0) - alpha) / (n + m)
        rs = xp.sum(0) - p
        cs = xp.sum(1) - q
        a = a - rs
        phi1 = ((a - rs) + gamma) / m
        b = b - cs
        phi2 = ((b - cs) + gamma) / n
        alpha = (alpha + 1.0) - S
    return xp


def tight_solve(c, p, q, step, kappa, iters=400000, eps=1e-14):
    """fp64 oracle solve to (near) machine precision; returns the plan."""
    import oracle as O

    r = O.qr_solve(np.asarray(c, np.float64), np.asarray(p, np.float64),
                   np.asarray(q, np.float64), float(np.float32(step)),
                   float(np.float32(kappa)), iters, eps, dtype=np.float64)
    return r.x


def loss(c, p, q, step, kappa, w, gobj, **kw):
    x = tight_solve(c, p, q, step, kappa, **kw)
    reg = reg_of(step, kappa)
    F = float((c * x).sum() + 0.5 * reg * (x * x).sum())
    return float((w * x).sum()) + gobj * F


def fd_grad(c, p, q, step, kappa, w, gobj, coords, h=1e-6, **kw):
    """Central differences of loss() in the listed (col, row) coordinates of C."""
    c = np.asarray(c, np.float64)
    out = []
    for (j, i) in coords:
        cp = c.copy()
        cm = c.copy()
        cp[j, i] += h
        cm[j, i] -= h
        out.append((loss(cp, p, q, step, kappa, w, gobj, **kw)
                    - loss(cm, p, q, step, kappa, w, gobj, **kw)) / (2 * h))
    return np.array(out)


# ------------------------------------------------------------------ exact solve

def dual_value(u, v, c, p, q, reg):
    """QR dual D(u, v) = u.p + v.q - (1/(2 reg)) sum [u_i + v_j - C_ji]_+^2
    (the dual of numpy/npdrot/solver.py:243-260 with the reference's layout)."""
    y = np.maximum(u[None, :] + v[:, None] - c, 0.0)
    return float(u @ p + v @ q - 0.5 / reg * (y * y).sum())


def newton_solve(c, p, q, reg, u0, v0, gtol=1e-15, max_iter=200):
    """Exact optimum of min <C,X> + (reg/2)||X||^2 over X >= 0, X 1 = p, X^T 1 = q
    by semi-smooth Newton ascent on the dual, in fp64.

    c: (n_cols, n_rows); u0 (n_rows), v0 (n_cols): starting potentials (the DR
    potentials phi1/s, phi2/s of a converged solve).  X = [u + v - C]_+ / reg;
    the dual gradient is (p - X^T 1, q - X 1) and its generalized Hessian
    -K_S/reg, K_S the signless bipartite Laplacian of the support
    S = {u + v - C > 0}.  Each step solves (K_S + delta I) d = reg * grad (a
    small ridge fixes the gauge directions u + t, v - t of each connected
    component), then halves the step until the gradient norm decreases.
    Converges to machine precision in a few dozen steps from a DR start.  The
    marginals must balance (sum p = sum q, e.g. fp64 1/n): otherwise no plan
    meets both and the gradient stalls at |sum p - sum q|.
    Returns (X, u, v, gradient norm)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    c = np.asarray(c, np.float64)
    p = np.asarray(p, np.float64)
    q = np.asarray(q, np.float64)
    n_cols, n_rows = c.shape
    u = np.asarray(u0, np.float64).copy()
    v = np.asarray(v0, np.float64).copy()

    def grad(u_, v_):
        y = u_[None, :] + v_[:, None] - c
        x = np.maximum(y, 0.0) / reg
        return x, y > 0, np.concatenate([p - x.sum(0), q - x.sum(1)])

    x, S, g = grad(u, v)
    gn = float(np.linalg.norm(g))
    scale = float(np.linalg.norm(np.concatenate([p, q])))
    for _ in range(max_iter):
        if gn <= gtol * scale:
            break
        jj, ii = np.nonzero(S)  # (col, row) entries of the support
        n = n_rows
        rows = np.concatenate([ii, n + jj, np.arange(n), n + np.arange(n_cols)])
        cols = np.concatenate([n + jj, ii, np.arange(n), n + np.arange(n_cols)])
        deg = np.concatenate([S.sum(0), S.sum(1)]).astype(np.float64)
        delta = 1e-12 * max(1.0, float(deg.max()))
        vals = np.concatenate([np.ones(2 * len(ii)), deg + delta])
        K = sp.csc_matrix((vals, (rows, cols)), shape=(n + n_cols, n + n_cols))
        d = spla.spsolve(K, reg * g)
        t = 1.0
        while True:
            un, vn = u + t * d[:n], v + t * d[n:]
            xn, Sn, gnew = grad(un, vn)
            gnn = float(np.linalg.norm(gnew))
            if gnn < gn * (1.0 - 1e-4 * t) or t < 1e-10:
                break
            t *= 0.5
        if gnn >= gn:
            break  # no further progress at fp64 resolution
        u, v, x, S, g, gn = un, vn, xn, Sn, gnew, gnn
    return x, u, v, gn


def exact_plan(c, p, q, step, kappa, iters=20000, eps=1e-9):
    """The exact fp64 optimum of the problem the QR iteration solves (reg from
    the fp32 step and scale): an fp64 oracle DR solve for the starting
    potentials, then newton_solve.  Returns (X, u, v)."""
    import oracle as O

    c = np.asarray(c, np.float64)
    r = O.qr_solve(c, np.asarray(p, np.float64), np.asarray(q, np.float64),
                   float(np.float32(step)), float(np.float32(kappa)), iters, eps,
                   dtype=np.float64)
    s = float(np.float32(step))
    x, u, v, _ = newton_solve(c, p, q, reg_of(step, kappa), r.state.phi1 / s,
                              r.state.phi2 / s)
    return x, u, v


def objective(c, x, reg):
    """<C, X> + (reg/2)||X||^2 in fp64."""
    x = np.asarray(x, np.float64)
    return float((np.asarray(c, np.float64) * x).sum() + 0.5 * reg * (x * x).sum())
