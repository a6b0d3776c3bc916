"""Pins the CPU restatement of the QR-DROT backward (oracle/implicit.py).

The reference has no backward, so the implicit gradient is checked against
routes that do not share its linear algebra:
  * full support: the implicit map reduces to -(double centering)/reg;
  * a one-to-one support: every perturbation of C on S is absorbed by the
    potentials, so dL/dC = 0;
  * central finite differences of fp64 oracle solves;
  * autograd through the unrolled fp64 DR iteration;
  * the support test separates the fp32 rounding residue of fix_x from the
    true support of a tight fp64 solve.
A dropped term (the potentials w, the -G, the envelope X), a sign error or a
transposed support fails at least one of these.
"""
import numpy as np
import pytest

import oracle as O
from oracle import implicit as I


def _problem(n_rows, n_cols, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n_cols, n_rows)) + 1e-3
    c = (c / c.max()).astype(np.float32).astype(np.float64)
    p = np.full(n_rows, 1.0 / n_rows)
    q = np.full(n_cols, 1.0 / n_cols)
    return c, p, q


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def test_full_support_is_double_centering():
    c, p, q = _problem(20, 10, 13)
    s, k = O.qr_params_torch(2.0, 10.0, 20, 10)  # strong regularization: X > 0 everywhere
    x = I.tight_solve(c, p, q, s, k, iters=20000)
    S = I.support(x, c, s)
    assert S.all()
    g = np.random.default_rng(1).standard_normal(c.shape)
    got = I.implicit_grad_c(x, c, g, s, k, gobj=0.7)
    want = -I.double_centering(g) / I.reg_of(s, k) + 0.7 * x
    assert _rel(got, want) < 1e-10


def test_one_to_one_support_gives_zero():
    n = 9
    c, _, _ = _problem(n, n, 2)
    S = np.eye(n, dtype=bool)[np.random.default_rng(3).permutation(n)]
    g = np.random.default_rng(4).standard_normal(c.shape)
    s, k = O.qr_params_torch(2.0, 1e-2, n, n)
    got = I.implicit_grad_c(np.zeros_like(c), c, g, s, k, S=S)
    assert np.abs(got).max() < 1e-12 * np.abs(g).max() / I.reg_of(s, k)


def test_laplacian_null_space():
    """K annihilates (1 on rows, -1 on columns) of every connected component."""
    S = np.zeros((4, 5), bool)
    S[0, 0] = S[0, 1] = S[1, 1] = True  # component A: rows 0,1 / cols 0,1
    S[2, 3] = S[3, 4] = S[3, 3] = True  # component B: rows 3,4 / cols 2,3
    K = I.laplacian(S)
    v = np.zeros(9)
    v[[0, 1]] = 1.0
    v[[5, 6]] = -1.0
    assert np.abs(K @ v).max() == 0.0
    assert np.linalg.matrix_rank(K) == 9 - 3  # two components + the isolated row 2


@pytest.mark.parametrize("n_rows,n_cols,rw,seed", [(16, 16, 1e-1, 3), (10, 8, 1e-2, 5),
                                                   (14, 11, 3e-2, 6)])
def test_matches_finite_differences(n_rows, n_cols, rw, seed):
    c, p, q = _problem(n_rows, n_cols, seed)
    s, k = O.qr_params_torch(2.0, rw, n_rows, n_cols)
    x = I.tight_solve(c, p, q, s, k, iters=30000)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(c.shape)
    g = I.implicit_grad_c(x, c, w, s, k, gobj=1.0, S=I.support(x, c, s, tol=2.0 ** -40))
    sj, si = np.nonzero(I.support(x, c, s, tol=2.0 ** -40))
    pick = rng.choice(len(sj), 4, replace=False)
    coords = [(int(sj[t]), int(si[t])) for t in pick]
    coords += [(int(rng.integers(n_cols)), int(rng.integers(n_rows))) for _ in range(2)]
    fd = I.fd_grad(c, p, q, s, k, w, 1.0, coords, h=1e-6, iters=30000)
    got = np.array([g[j, i] for j, i in coords])
    assert _rel(got, fd) < 1e-6


@pytest.mark.parametrize("n_rows,n_cols,rw", [(12, 9, 1e-2), (16, 16, 1e-1)])
def test_matches_unrolled_autograd(n_rows, n_cols, rw):
    torch = pytest.importorskip("torch")
    c, p, q = _problem(n_rows, n_cols, 8)
    s, k = O.qr_params_torch(2.0, rw, n_rows, n_cols)
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    xp = I.unrolled_plan(ct, p, q, s, k, 6000)
    w = np.random.default_rng(9).standard_normal(c.shape)
    reg = I.reg_of(s, k)
    F = (ct * xp).sum() + 0.5 * reg * (xp * xp).sum()
    ((torch.tensor(w) * xp).sum() + 0.3 * F).backward()
    x = xp.detach().numpy()
    g = I.implicit_grad_c(x, c, w, s, k, gobj=0.3, S=I.support(x, c, s, tol=2.0 ** -40))
    assert _rel(ct.grad.numpy(), g) < 1e-8


def test_unrolled_plan_is_the_oracle_iteration():
    torch = pytest.importorskip("torch")
    c, p, q = _problem(9, 7, 10)
    s, k = O.qr_params_torch(2.0, 1e-2, 9, 7)
    for iters in (1, 2, 7):
        xt = I.unrolled_plan(torch.tensor(c), p, q, s, k, iters).numpy()
        st = O.warm_start(9, 7, s, dtype=np.float64)
        for it in range(iters):
            O.qr_step(st, it, c, p, q, float(np.float32(s)), float(np.float32(k)))
        xo = st.x.copy()
        if iters % 2 == 1:  # the last iteration was even: the store holds X' - sC
            xo = xo + float(np.float32(s)) * c
        assert _rel(xt, xo) < 1e-12


def test_support_rejects_fix_x_residue():
    """The fp32 plan's positive entries include the rounding residue of adding
    step*C back; the support test keeps exactly the tight fp64 support."""
    c, p, q = _problem(40, 30, 12)
    c32 = c.astype(np.float32)
    s, k = O.qr_params_torch(2.0, 1e-2, 40, 30)
    r32 = O.qr_solve(c32, p.astype(np.float32), q.astype(np.float32), s, k, 200000, 1e-7)
    x64 = I.tight_solve(c, p, q, s, k, iters=60000)
    s64 = x64 > 1e-9 * x64.max()
    assert (r32.x > 0).sum() > s64.sum()  # residues are there
    assert np.array_equal(I.support(r32.x, c32, s), s64)


@pytest.mark.parametrize("n_rows,n_cols,rw", [(24, 20, 3e-2), (100, 80, 1e-3), (60, 90, 1e-4)])
def test_newton_solve_is_the_exact_optimum(n_rows, n_cols, rw):
    """newton_solve against two independent routes: the fp64 DR oracle run to
    1e-14 (same plan to 1e-10 relative) and strong duality (primal objective
    equals the dual value D(u, v) to 1e-12), with primal feasibility to 1e-13."""
    from reg_drot import problems
    c = problems.cloud_cost(n_rows, n_cols, 16, seed=n_rows + n_cols).astype(np.float64)
    p = np.full(n_rows, 1.0 / n_rows)
    q = np.full(n_cols, 1.0 / n_cols)
    s, k = O.qr_params_torch(2.0, rw, n_rows, n_cols)
    x, u, v = I.exact_plan(c, p, q, s, k)
    reg = I.reg_of(s, k)
    assert np.sqrt(((p - x.sum(0)) ** 2).sum() + ((q - x.sum(1)) ** 2).sum()) <= 1e-13
    assert (x >= 0).all()
    primal = I.objective(c, x, reg)
    assert abs(primal - I.dual_value(u, v, c, p, q, reg)) <= 1e-12 * abs(primal)
    xt = I.tight_solve(c, p, q, s, k, iters=400000, eps=1e-14)
    assert np.abs(xt - x).max() <= 1e-10 * np.abs(x).max()


def test_newton_solve_detects_a_wrong_start_cost():
    """Starting Newton from the potentials of C and solving for C' (C with one
    support entry raised) moves the plan: the solver does not just return its
    start, and the new plan is feasible."""
    from reg_drot import problems
    n_rows, n_cols = 40, 30
    c = problems.cloud_cost(n_rows, n_cols, 16, seed=5).astype(np.float64)
    p = np.full(n_rows, 1.0 / n_rows)
    q = np.full(n_cols, 1.0 / n_cols)
    # a strong regularization gives a dense support (a tree-shaped support
    # would fix the plan by the marginals alone)
    s, k = O.qr_params_torch(2.0, 0.5, n_rows, n_cols)
    x, u, v = I.exact_plan(c, p, q, s, k)
    assert (x > 0).mean() > 0.5
    c2 = c.copy()
    j, i = np.unravel_index(int(np.argmax(x)), x.shape)  # an entry of the support
    c2[j, i] += 1e-3
    x2, *_ = I.newton_solve(c2, p, q, I.reg_of(s, k), u, v)
    assert np.abs(x2 - x).max() > 1e-8
    assert np.sqrt(((p - x2.sum(0)) ** 2).sum() + ((q - x2.sum(1)) ** 2).sum()) <= 1e-13
