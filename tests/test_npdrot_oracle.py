"""The numpy restatement of the reference's CPU solver (oracle/npdrot.py,
following numpy/npdrot/solver.py:7-189 and proximal.py:75-96,186-210) against
the C oracle (the CUDA path's iteration) and the golden dual solutions.

npdrot keeps the DR iterate Y and applies rank-2 updates; the CUDA path keeps
X' and dual accumulators (a, b, alpha).  From Y0 = step 1 1^T with marginals
whose sums agree exactly, the two bookkeepings produce the same iterates to
rounding, which pins the oracle's dual update independently of its own code.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from oracle import npdrot as N

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cost(n_rows, n_cols, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n_cols, n_rows)) + 1e-3
    return c / c.max()


def _oracle_iterate(c, p, q, s, k, iters):
    st = O.warm_start(len(p), len(q), s, dtype=np.float64)
    for it in range(iters):
        O.qr_step(st, it, c, p, q, s, k)
    x = st.x.copy()
    if iters % 2 == 1:  # last iteration even: the store holds X' - sC
        x = x + s * c
    return x, st


@pytest.mark.parametrize("n_rows,n_cols,rw", [(32, 16, 1e-2), (64, 8, 1e-3), (16, 16, 1e-1),
                                              (8, 128, 1e-2)])
def test_trajectory_matches_cuda_bookkeeping(n_rows, n_cols, rw):
    """Power-of-two sizes make the uniform marginals exact (sum p == sum q)."""
    c = _cost(n_rows, n_cols, n_rows + n_cols)
    p = np.full(n_rows, 1.0 / n_rows)
    q = np.full(n_cols, 1.0 / n_cols)
    s, k = (float(v) for v in O.qr_params_torch(2.0, rw, n_rows, n_cols))
    reg = (k - 1.0) / s
    for iters in (2, 3, 17, 250):
        xo, _ = _oracle_iterate(c, p, q, s, k, iters)
        xn, info = N.regdrot(np.full((n_rows, n_cols), s), c.T, p, q, reg,
                             max_iters=iters - 1, eps_abs=-1.0, step=s)
        assert np.abs(xn - xo.T).max() <= 1e-13 * np.abs(xo).max(), iters


def test_residual_sequence_matches_oracle():
    n_rows, n_cols = 32, 16
    c = _cost(n_rows, n_cols, 5)
    p = np.full(n_rows, 1.0 / n_rows)
    q = np.full(n_cols, 1.0 / n_cols)
    s, k = (float(v) for v in O.qr_params_torch(2.0, 1e-2, n_rows, n_cols))
    _, info = N.regdrot(np.full((n_rows, n_cols), s), c.T, p, q, (k - 1.0) / s,
                        max_iters=40, eps_abs=-1.0, step=s)
    st = O.warm_start(n_rows, n_cols, s, dtype=np.float64)
    res = []
    for it in range(40):
        aux = O.qr_step(st, it, c, p, q, s, k)
        res.append(np.sqrt(aux[2] + aux[3]))
    assert np.allclose(info["r_primal"], res, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "qr_dual_*.npz"))))
def test_converges_to_dual_solution(path):
    g = np.load(path)
    c = g["c"].astype(np.float64)
    n_cols, n_rows = c.shape
    s, k = (float(v) for v in O.qr_params_torch(2.0, float(g["r_weight"]), n_rows, n_cols))
    x, info = N.regdrot(np.full((n_rows, n_cols), s), c.T, g["p"], g["q"], (k - 1.0) / s,
                        max_iters=300000, eps_abs=1e-12, step=s)
    assert np.abs(x.T - g["x"]).max() <= 2e-5 * np.abs(g["x"]).max()
    # the dual potentials the reference logs are the KKT potentials
    u, v = info["dual_sol"]
    u, v = u / s, v / s
    xk = np.maximum(u[:, None] + v[None, :] - c.T, 0.0) / ((k - 1.0) / s)
    assert np.abs(xk - x).max() <= 2e-5 * np.abs(x).max()


def test_random_marginals_same_solution_as_oracle():
    n_rows, n_cols = 24, 18
    rng = np.random.default_rng(3)
    c = _cost(n_rows, n_cols, 9)
    p = rng.random(n_rows) + 0.1
    p /= p.sum()
    q = rng.random(n_cols) + 0.1
    q /= q.sum()
    s, k = (float(v) for v in O.qr_params_torch(2.0, 3e-2, n_rows, n_cols))
    x, _ = N.regdrot(np.full((n_rows, n_cols), s), c.T, p, q, (k - 1.0) / s,
                     max_iters=200000, eps_abs=1e-13, step=s)
    ref = O.qr_solve(c, p, q, s, k, 200000, 1e-13, dtype=np.float64)
    assert np.abs(x.T - ref.x).max() <= 1e-8 * np.abs(ref.x).max()


def test_quad_prox_value():
    y = np.array([[0.5, -1.0], [2.0, 0.25]])
    c = np.array([[1.0, 0.0], [0.5, 0.25]])
    val = N.quad_prox(y, c, 0.5, 2.0)
    want = np.maximum(np.array([[0.0, -1.0], [1.75, 0.125]]), 0.0) / 2.0
    assert np.array_equal(y, want)
    assert val == pytest.approx(0.5 * 2.0 * (want ** 2).sum())


@pytest.mark.parametrize("n_rows,n_cols,rw", [(64, 48, 1e-2), (300, 200, 1e-3), (129, 77, 1e-1)])
@pytest.mark.parametrize("iters", [1, 2, 40, 400])
def test_openmp_restatement_matches_numpy(n_rows, n_cols, rw, iters):
    """oracle/npdrot_omp.c (the CPU baseline: npdrot's passes in C with
    OpenMP) against the numpy restatement, fp64: same plan to 1e-12 after the
    same iterations, same residual sequence end; the fp32 instance follows the
    fp64 one to fp32 accuracy."""
    c = _cost(n_rows, n_cols, 3 * n_rows + n_cols)
    p = np.full(n_rows, 1.0 / n_rows)
    q = np.full(n_cols, 1.0 / n_cols)
    s, k = (float(v) for v in O.qr_params_torch(2.0, rw, n_rows, n_cols))
    reg = (k - 1.0) / s
    xn, info = N.regdrot(np.full((n_rows, n_cols), s), c.T, p, q, reg, max_iters=iters,
                         eps_abs=-1.0, step=s)
    y0 = np.full((n_cols, n_rows), s)  # solver layout (n_cols, n_rows)
    xc, kc, rc = O.npdrot_solve(y0, c, p, q, s, reg, iters, -1.0, dtype=np.float64, nthreads=4)
    assert kc == info["num_iters"] == iters
    assert np.abs(xc - xn.T).max() <= 1e-12 * np.abs(xn).max()
    assert abs(rc - info["r_primal"][-1]) <= 1e-9 * info["r_primal"][-1] + 1e-15
    x32, k32, _ = O.npdrot_solve(y0, c, p, q, s, reg, iters, -1.0, dtype=np.float32,
                                 nthreads=3)
    assert k32 == iters
    assert np.abs(x32 - xc).max() <= 1e-4 * np.abs(xc).max()


def test_openmp_restatement_stops_on_the_residual():
    """The stop rule of solver.py:165-166 (r_primal <= eps_abs) and the extra
    prox after the loop, against the numpy restatement; thread counts do not
    change the result beyond rounding."""
    n_rows, n_cols = 256, 128
    c = _cost(n_rows, n_cols, 5)
    p = np.full(n_rows, 1.0 / n_rows)
    q = np.full(n_cols, 1.0 / n_cols)
    s, k = (float(v) for v in O.qr_params_torch(2.0, 1e-2, n_rows, n_cols))
    reg = (k - 1.0) / s
    xn, info = N.regdrot(np.full((n_rows, n_cols), s), c.T, p, q, reg, max_iters=20000,
                         eps_abs=1e-6, step=s)
    y0 = np.full((n_cols, n_rows), s)
    for nt in (1, 5):
        xc, kc, rc = O.npdrot_solve(y0, c, p, q, s, reg, 20000, 1e-6, nthreads=nt)
        assert kc == info["num_iters"]
        assert rc <= 1e-6
        assert np.abs(xc - xn.T).max() <= 1e-11 * np.abs(xn).max()
