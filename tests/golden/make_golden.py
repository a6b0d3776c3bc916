"""Write the golden fixtures of tests/golden/ from solvers independent of the
oracle and of the GPU path.

The reference holds no golden vectors for this path (its tests only assert
residual <= eps, pytorch/test.py:53-59), so every fixture here is fixed by the
mathematics of the problem the reference solves:

  qr_dual_*.npz   quadratically regularized OT
                      min <C,X> + (reg/2)||X||^2  s.t. X1 = p, X^T 1 = q, X >= 0
                  with reg = (n_rows + n_cols) * r_weight (the scaling of
                  pytorch/README.md:44-47), solved through its smooth concave
                  dual with L-BFGS-B; X* = [u_i + v_j - C_ij]_+ / reg.
  lp_*.npz        unregularized OT (r_weight = 0) solved by HiGHS linprog.

Inputs use the cost construction of pytorch/test.py:41-42 with fixed seeds.
Arrays follow the solver layout: C and X have shape (n_cols, n_rows).
Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
from scipy.optimize import linprog, minimize

HERE = os.path.dirname(os.path.abspath(__file__))


def problem(n_rows, n_cols, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n_cols, n_rows), dtype=np.float32) + np.float32(1e-3)
    c = (c / c.max()).astype(np.float32)
    p = (np.ones(n_rows, np.float32) / np.float32(n_rows)).astype(np.float32)
    q = (np.ones(n_cols, np.float32) / np.float32(n_cols)).astype(np.float32)
    return c, p, q


def qr_dual_solve(c, p, q, reg):
    """Maximise D(u,v) = u.p + v.q - (1/2reg) sum [u_i+v_j-C_ij]_+^2."""
    C = c.T.astype(np.float64)  # (n_rows, n_cols)
    n, m = C.shape
    pd, qd = p.astype(np.float64), q.astype(np.float64)

    def negdual(z):
        u, v = z[:n], z[n:]
        t = np.maximum(u[:, None] + v[None, :] - C, 0.0)
        f = u @ pd + v @ qd - 0.5 / reg * (t * t).sum()
        gu = pd - t.sum(1) / reg
        gv = qd - t.sum(0) / reg
        return -f, -np.concatenate([gu, gv])

    res = minimize(negdual, np.zeros(n + m), jac=True, method="L-BFGS-B",
                   options=dict(maxiter=200000, maxfun=400000, gtol=1e-15, ftol=1e-18,
                                maxcor=50))
    u, v = res.x[:n], res.x[n:]
    x = np.maximum(u[:, None] + v[None, :] - C, 0.0) / reg
    return x.T.copy(), u, v, -res.fun


def lp_solve(c, p, q):
    C = c.T.astype(np.float64)
    n, m = C.shape
    a_eq = np.zeros((n + m, n * m))
    for i in range(n):
        a_eq[i, i * m:(i + 1) * m] = 1.0
    for j in range(m):
        a_eq[n + j, j::m] = 1.0
    b_eq = np.concatenate([p, q]).astype(np.float64)
    res = linprog(C.reshape(-1), A_eq=a_eq, b_eq=b_eq, bounds=(0, None), method="highs")
    assert res.status == 0, res.message
    return res.x.reshape(n, m).T.copy(), float(res.fun)


QR_CASES = [  # n_rows, n_cols, r_weight, seed
    (2, 3, 1e-2, 11), (10, 20, 1e-3, 12), (20, 10, 10.0, 13), (30, 20, 1e-2, 14),
    (24, 24, 1e-3, 15), (40, 33, 5e-3, 16),
]
LP_CASES = [(3, 4, 21), (8, 8, 22), (12, 9, 23), (16, 20, 24)]


def _save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez(path, **arrays)
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def main():
    manifest = {}
    for n_rows, n_cols, rw, seed in QR_CASES:
        c, p, q = problem(n_rows, n_cols, seed)
        reg = (n_rows + n_cols) * rw
        x, u, v, dual = qr_dual_solve(c, p, q, reg)
        name = f"qr_dual_{n_rows}x{n_cols}.npz"
        manifest[name] = _save(name, c=c, p=p, q=q, r_weight=np.float64(rw),
                               reg=np.float64(reg), x=x, u=u, v=v, dual=np.float64(dual))
    for n_rows, n_cols, seed in LP_CASES:
        c, p, q = problem(n_rows, n_cols, seed)
        x, obj = lp_solve(c, p, q)
        name = f"lp_{n_rows}x{n_cols}.npz"
        manifest[name] = _save(name, c=c, p=p, q=q, x=x, obj=np.float64(obj))
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
