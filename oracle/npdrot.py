"""numpy restatement of the reference's CPU solver (npdrot), for tests and the
CPU baseline only.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU baseline; the
product path never imports it.

npdrot itself cannot run here (its modules import numba at top level,
numpy/npdrot/solver.py:3 and numpy/npdrot/proximal/proximal.py:3, and numba is
not installed), so this module restates its algorithm in plain numpy:

  regdrot        numpy/npdrot/solver.py:7-189 -- DR splitting that keeps the
                 iterate Y in place of X: prox, row/column sums, the scalar
                 corrections c1, c2, and the rank-2 update
                 Y <- X' - (1/n) yy 1^T - (1/m) 1 xx^T
                 (apply_adjoint_operator_and_override, proximal.py:186-210,
                 two BLAS dger calls there; two broadcast subtractions here);
  quad_prox      trace_nonnegative_quad_prox, proximal.py:75-96:
                 X' = max(Y - step C, 0) / (1 + step reg).

Deliberate differences (the reference's documented slips are not copied):
  * the quadratic term is returned as (reg/2)||X||^2 (proximal.py:71 returns
    reg^2/2 ||X||^2 on the numba path);
  * the log's "gap" entry holds the duality gap, not r_primal (solver.py:177).
The iteration is the same map as the CUDA solver's (core/kernel_qr.hpp,
core/drot_qr.hpp) in another bookkeeping: from Y0 = step 1 1^T both produce the
same iterates X'_k, which tests/test_npdrot_oracle.py checks against the C
oracle.  Arrays use npdrot's orientation: X is (m, n) = (len(p), len(q)).
"""
from __future__ import annotations

import numpy as np


def quad_prox(y, c, step, reg):
    """In place X' = max(Y - step C, 0) / (1 + step reg); returns (reg/2)||X'||^2."""
    np.subtract(y, step * c, out=y)
    np.maximum(y, 0.0, out=y)
    y /= 1.0 + step * reg
    return 0.5 * reg * float(np.vdot(y, y))


def regdrot(init, c, p, q, reg, max_iters=100, eps_abs=1e-6, step=None, dtype=np.float64,
            log_every=0):
    """Quadratically regularized DR-OT (npdrot.quaddrot -> regdrot).

    init, c: (m, n) arrays; p: (m,), q: (n,).  Stops when the primal residual
    ||[X'1; X'^T 1] - [p; q]|| <= eps_abs or after max_iters iterations
    (solver.py:144-175).  Returns (X, info) where X is the final prox output
    (solver.py:180 applies one more prox to the last iterate) and info holds
    num_iters, r_primal (per iteration) and the dual potentials (-yy/n, -xx/m)
    of the last iteration (solver.py:176)."""
    m, n = c.shape
    step = 2.0 / (m + n) if step is None else float(step)
    c = np.asfortranarray(c, dtype)
    y = np.array(init, dtype=dtype, order="F", copy=True)
    p = np.asarray(p, dtype)
    q = np.asarray(q, dtype)
    a1 = y.sum(1)
    b1 = y.sum(0)
    r_primal = []
    k = 0
    yy = np.zeros(m, dtype)
    xx = np.zeros(n, dtype)
    while True:
        quad_prox(y, c, step, reg)
        a2 = y.sum(1)
        b2 = y.sum(0)
        t1 = 2 * a2 - a1 - p
        t2 = 2 * b2 - b1 - q
        c1 = t1.sum() / (m + n)
        c2 = t2.sum() / (m + n)
        yy = t1 - c1
        xx = t2 - c2
        a1 = a2 - yy - c2
        b1 = b2 - xx - c1
        r = float(np.sqrt(np.sum((a2 - p) ** 2) + np.sum((b2 - q) ** 2)))
        r_primal.append(r)
        y -= (yy / n)[:, None]
        y -= (xx / m)[None, :]
        k += 1
        if k >= max_iters or r <= eps_abs:
            break
    x = y
    quad_prox(x, c, step, reg)
    return x, {"num_iters": k, "r_primal": np.array(r_primal),
               "dual_sol": (-yy / n, -xx / m), "step": step}
