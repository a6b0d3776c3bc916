"""Pins the CPU oracle (oracle/drot_oracle.c) to what the mathematics fixes.

The reference ships no golden vectors for this path, so each oracle function is
checked against an independent route:
  * converged QR plans vs an L-BFGS-B solve of the QR dual (tests/golden/qr_dual_*)
    and the KKT form X = [u + v - C]_+ / reg with u = phi1/s, v = phi2/s;
  * unregularized plans vs HiGHS linprog (tests/golden/lp_*) and the assignment
    solution of scipy.optimize.linear_sum_assignment;
  * special cases with known plans (constant cost, one row / one column,
    1-D monotone coupling);
  * the first iteration from the warm start in closed form (Y0 = s 11^T);
  * GL: the block KKT conditions of the group-lasso problem, lambda = 0 and
    singleton groups reducing to unregularized OT, and GL(lambda=0) == QR(r=0);
  * the stopping-rule and n_iter conventions by stepping the iteration by hand;
  * group bounds vs exact rational rounding; parameter conventions.
A plausible slip (a sign, a transposed sum, a dropped gamma or alpha term, a
wrong divisor) breaks convergence or moves the fixed point and fails here.
"""
import glob
import os
from fractions import Fraction

import numpy as np
import pytest
from scipy.optimize import linear_sum_assignment

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tight_qr(c, p, q, r_weight, iters=300000, eps=1e-10):
    n_cols, n_rows = c.shape
    s, k = O.qr_params_torch(2.0, r_weight, n_rows, n_cols)
    return O.qr_solve(c, p, q, float(s), float(k), iters, eps, dtype=np.float64), s, k


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


# ------------------------------------------------------------------ QR

@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "qr_dual_*.npz"))))
def test_qr_converges_to_dual_solution(path):
    g = np.load(path)
    res, s, k = tight_qr(g["c"], g["p"], g["q"], float(g["r_weight"]))
    assert rel(res.x, g["x"]) < 2e-5, os.path.basename(path)
    # the oracle's duals are the DR potentials: u = phi1 / s, v = phi2 / s
    u = res.state.phi1 / float(s)
    v = res.state.phi2 / float(s)
    C = g["c"].T.astype(np.float64)
    reg = float(g["reg"])
    x_kkt = np.maximum(u[:, None] + v[None, :] - C, 0.0) / reg
    assert rel(x_kkt.T, res.x) < 2e-5
    # zero duality gap: primal value equals the dual value
    xp = res.x.T
    primal = float((C * xp).sum() + 0.5 * reg * (xp * xp).sum())
    assert abs(primal - float(g["dual"])) < 1e-6 * max(1.0, abs(primal))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "lp_*.npz"))))
def test_qr_unregularized_matches_linprog(path):
    g = np.load(path)
    res, _, _ = tight_qr(g["c"], g["p"], g["q"], 0.0, iters=400000, eps=1e-10)
    obj = float((g["c"].astype(np.float64) * res.x).sum())
    assert abs(obj - float(g["obj"])) < 1e-6
    assert rel(res.x, g["x"]) < 1e-4  # random costs: the LP vertex is unique


def test_qr_unregularized_square_is_assignment():
    n = 12
    rng = np.random.default_rng(4)
    c = rng.random((n, n)).astype(np.float32)
    p = np.full(n, 1 / n, np.float32)
    res, _, _ = tight_qr(c, p, p, 0.0, iters=400000)
    rows, cols = linear_sum_assignment(c.astype(np.float64))
    perm = np.zeros((n, n))
    perm[rows, cols] = 1.0 / n
    assert rel(res.x, perm) < 1e-4


def test_qr_constant_cost_gives_product_plan():
    n_rows, n_cols = 7, 5
    c = np.full((n_cols, n_rows), 0.5, np.float32)
    p = np.full(n_rows, 1 / n_rows, np.float32)
    q = np.full(n_cols, 1 / n_cols, np.float32)
    res, _, _ = tight_qr(c, p, q, 1e-2)
    assert rel(res.x, np.full((n_cols, n_rows), 1 / (n_rows * n_cols))) < 1e-6


@pytest.mark.parametrize("n_rows,n_cols", [(1, 9), (9, 1)])
def test_qr_single_row_or_column_is_forced(n_rows, n_cols):
    rng = np.random.default_rng(5)
    c = rng.random((n_cols, n_rows)).astype(np.float32)
    p = rng.random(n_rows).astype(np.float32) + 0.1
    q = rng.random(n_cols).astype(np.float32) + 0.1
    p /= p.sum()
    q /= q.sum()
    res, _, _ = tight_qr(c, p, q, 1e-2)
    forced = (q[:, None] if n_rows == 1 else p[None, :]) * np.ones((n_cols, n_rows))
    assert rel(res.x, forced) < 1e-6


def test_qr_1d_sorted_points_give_monotone_coupling():
    n = 10
    x = np.sort(np.random.default_rng(6).random(n))
    y = np.sort(np.random.default_rng(7).random(n))
    c = ((y[:, None] - x[None, :]) ** 2).astype(np.float32)
    c /= c.max()
    p = np.full(n, 1 / n, np.float32)
    res, _, _ = tight_qr(c, p, p, 0.0, iters=400000)
    assert rel(res.x, np.eye(n) / n) < 1e-4


def test_warm_start_is_the_implied_iterate():
    """Warm start == DR iterate Y0 = s 11^T (pytorch/src/qr_drot_cuda_kernel.cu:45-71)."""
    n_rows, n_cols = 11, 7
    s = O.step_torch(2.0, n_rows, n_cols)
    st = O.warm_start(n_rows, n_cols, float(s), dtype=np.float64)
    y0 = np.full((n_rows, n_cols), float(s))
    p = np.full(n_rows, 1 / n_rows)
    q = np.full(n_cols, 1 / n_cols)
    np.testing.assert_allclose(st.phi1 + st.phi2[0], float(s), rtol=1e-12)
    np.testing.assert_allclose(st.a, y0.sum(1) - p, rtol=1e-12)
    np.testing.assert_allclose(st.b, y0.sum(0) - q, rtol=1e-12)
    assert abs(st.alpha - (y0.sum() - 1.0)) < 1e-12
    assert not st.x.any()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
def test_first_iteration_closed_form(dtype, tol):
    """From Y0 = s 11^T the first prox is X' = s max(1 - C, 0) / kappa; k=0 is
    even, so the stored iterate is X' - s C (core/kernel_qr.hpp:241-245)."""
    n_rows, n_cols = 40, 30
    rng = np.random.default_rng(8)
    c = rng.random((n_cols, n_rows)).astype(np.float32) * 1.5  # some C > 1: clamp active
    p = np.full(n_rows, 1 / n_rows, np.float32)
    q = np.full(n_cols, 1 / n_cols, np.float32)
    s, k = O.qr_params_torch(2.0, 1e-2, n_rows, n_cols)
    st = O.warm_start(n_rows, n_cols, float(s), dtype=dtype)
    alpha0 = st.alpha
    aux = O.qr_step(st, 0, c, p, q, float(s), float(k))
    sd, kd = float(s), float(k)
    c64 = c.astype(np.float64)
    xp = sd * np.maximum(1.0 - c64, 0.0) / kd
    assert rel(st.x + sd * c64, xp) < tol
    assert abs(aux[0] - (c.astype(np.float64) * xp).sum()) < 10 * tol * abs(aux[0])
    assert abs(aux[1] - xp.sum()) < 10 * tol * xp.sum()
    assert abs(st.alpha - (alpha0 + 1.0 - xp.sum())) < 10 * tol * max(1.0, abs(alpha0))
    rs = xp.sum(0) - p
    cs = xp.sum(1) - q
    assert abs(aux[2] - (rs ** 2).sum()) < 1e-5 * (rs ** 2).sum()
    assert abs(aux[3] - (cs ** 2).sum()) < 1e-5 * (cs ** 2).sum()


def test_cold_start_first_iteration():
    """Cold start X = p q^T, phi = 0 (core/drot_qr.hpp:308-320): X' = max(pq^T - sC, 0)/kappa."""
    n_rows, n_cols = 9, 6
    rng = np.random.default_rng(9)
    c = (rng.random((n_cols, n_rows)) * 1e-3).astype(np.float32)
    p = np.full(n_rows, 1 / n_rows)
    q = np.full(n_cols, 1 / n_cols)
    s, k = O.qr_params_torch(2.0, 1e-2, n_rows, n_cols)
    st = O.cold_start(p, q, dtype=np.float64)
    np.testing.assert_allclose(st.x, np.outer(q, p))
    O.qr_step(st, 0, c, p, q, float(s), float(k))
    c64 = c.astype(np.float64)
    xp = np.maximum(np.outer(q, p) - float(s) * c64, 0.0) / float(k)
    assert rel(st.x + float(s) * c64, xp) < 1e-12


def test_qr_stop_rule_and_n_iter():
    """Test on even k only, residual < eps; n_iter = k + 1; else max_iters + 1
    (core/drot_qr.hpp:190-233)."""
    n_rows, n_cols = 30, 20
    rng = np.random.default_rng(10)
    c = rng.random((n_cols, n_rows)).astype(np.float32)
    p = np.full(n_rows, 1 / n_rows, np.float32)
    q = np.full(n_cols, 1 / n_cols, np.float32)
    s, k = O.qr_params_torch(2.0, 1e-2, n_rows, n_cols)
    eps = np.float32(1e-4)
    st = O.warm_start(n_rows, n_cols, s)
    first_even = None
    for it in range(5000):
        aux = O.qr_step(st, it, c, p, q, s, k)
        if it % 2 == 0 and np.sqrt(aux[2] + aux[3]) < float(eps):
            first_even = it
            break
    res = O.qr_solve(c, p, q, s, k, 5000, eps)
    assert first_even is not None and res.converged
    assert res.n_iter == first_even + 1
    # after fix_x the plan is the un-shifted X' of that iteration
    fixed = np.maximum(st.x.astype(np.float64) + float(s) * c.astype(np.float64), 0)
    assert rel(res.x, fixed) < 1e-6
    short = O.qr_solve(c, p, q, s, k, 10, eps)
    assert short.n_iter == 11 and not short.converged


def test_qr_fp32_tracks_fp64():
    n_rows, n_cols = 100, 80
    c = np.random.default_rng(11).random((n_cols, n_rows)).astype(np.float32)
    p = np.full(n_rows, 1 / n_rows, np.float32)
    q = np.full(n_cols, 1 / n_cols, np.float32)
    s, k = O.qr_params_torch(2.0, 1e-2, n_rows, n_cols)
    a = O.qr_solve(c, p, q, s, k, 600, 0.0)
    b = O.qr_solve(c, p, q, float(s), float(k), 600, 0.0, dtype=np.float64)
    assert rel(a.x, b.x) < 1e-3


# ------------------------------------------------------------------ GL

def gl_tight(c, p, q, n_groups, reg, iters=100000, eps=0.0, bounds=None):
    # the +1e-9 in the reference's norm (core/kernel_glr.hpp:125,320) leaves a
    # residual floor near 1e-8, so run a fixed number of iterations
    n_cols, n_rows = c.shape
    s = O.step_torch(2.0, n_rows, n_cols)
    lam = reg * float(s)
    return O.gl_solve(c, p, q, n_groups, float(s), lam, iters, eps, dtype=np.float64,
                      bounds=bounds), float(s)


@pytest.mark.parametrize("n_rows,n_cols,n_groups,reg", [(12, 8, 3, 0.02), (20, 10, 2, 0.05),
                                                         (15, 6, 5, 0.01)])
def test_gl_block_kkt(n_rows, n_cols, n_groups, reg):
    """Group-lasso OT optimality: with u = phi1/s, v = phi2/s and W = [u+v-C]_+,
    every block X_{g,j} is parallel to W_{g,j} with ||W_{g,j}|| = reg when
    active, and ||W_{g,j}|| <= reg when X_{g,j} = 0."""
    rng = np.random.default_rng(n_rows)
    c = rng.random((n_cols, n_rows)).astype(np.float32)
    p = np.full(n_rows, 1 / n_rows, np.float32)
    q = np.full(n_cols, 1 / n_cols, np.float32)
    res, s = gl_tight(c, p, q, n_groups, reg)
    assert res.residual < 1e-7
    b = O.group_bounds(n_rows, n_groups)
    X = res.x.T
    W = np.maximum(res.state.phi1[:, None] / s + res.state.phi2[None, :] / s - c.T, 0.0)
    n_active = 0
    for j in range(n_cols):
        for g in range(n_groups):
            w = W[b[g]:b[g + 1], j]
            x = X[b[g]:b[g + 1], j]
            nx = np.linalg.norm(x)
            if nx > 1e-7:
                n_active += 1
                assert abs(np.linalg.norm(w) - reg) < 1e-3 * reg
                cos = float(w @ x) / (np.linalg.norm(w) * nx)
                assert cos > 1 - 1e-5
            else:
                assert np.linalg.norm(w) <= reg * (1 + 1e-3)
    assert n_active > 0


def test_gl_lambda_zero_equals_qr_unregularized():
    """GL with lambda = 0 and QR with r_weight = 0 generate the same X' sequence."""
    n_rows, n_cols = 16, 12
    c = np.random.default_rng(12).random((n_cols, n_rows)).astype(np.float32)
    p = np.full(n_rows, 1 / n_rows, np.float32)
    q = np.full(n_cols, 1 / n_cols, np.float32)
    s = float(O.step_torch(2.0, n_rows, n_cols))
    gl = O.gl_solve(c, p, q, 4, s, 0.0, 200, 0.0, dtype=np.float64)
    qr = O.qr_solve(c, p, q, s, 1.0, 200, 0.0, dtype=np.float64)  # 200: ends odd, X'
    assert rel(gl.x, qr.x) < 1e-12
    np.testing.assert_allclose(gl.state.phi1, qr.state.phi1, rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "lp_*.npz")))[:3])
def test_gl_singleton_groups_solve_the_lp(path):
    """With one row per group the penalty is reg * sum X = reg on the feasible
    set, so the solution is the unregularized LP plan."""
    g = np.load(path)
    n_cols, n_rows = g["c"].shape
    res, _ = gl_tight(g["c"], g["p"], g["q"], n_rows, 0.05, iters=400000)
    obj = float((g["c"].astype(np.float64) * res.x).sum())
    assert abs(obj - float(g["obj"])) < 1e-6


def test_gl_stop_rule_every_iteration():
    n_rows, n_cols = 20, 10
    c = np.random.default_rng(13).random((n_cols, n_rows)).astype(np.float32)
    p = np.full(n_rows, 1 / n_rows, np.float32)
    q = np.full(n_cols, 1 / n_cols, np.float32)
    s = O.step_torch(2.0, n_rows, n_cols)
    lam = O.gl_lambda_torch(1e-2, n_rows, 2)
    b = O.group_bounds(n_rows, 2)
    eps = np.float32(1e-4)
    st = O.warm_start(n_rows, n_cols, s)
    hit = None
    for it in range(20000):
        aux = O.gl_step(st, b, c, p, q, s, lam)
        if np.sqrt(aux[2] + aux[3]) <= float(eps):
            hit = it
            break
    res = O.gl_solve(c, p, q, 2, s, lam, 20000, eps)
    assert hit is not None and res.n_iter == hit + 1 and res.converged
    assert O.gl_solve(c, p, q, 2, s, lam, 7, eps).n_iter == 7


# ------------------------------------------------------------------ conventions

def test_group_bounds_match_exact_rounding():
    """bounds[g] = floor(g*N/G + 0.49) (<= 256 rows per group) or + 0.49999,
    computed in fp32 by the reference; for these sizes the fp32 evaluation
    equals the exact rational one."""
    for n in range(1, 300):
        for g in (1, 2, 3, 5, 7, 10):
            if g > n:
                continue
            b = O.group_bounds(n, g)
            small = -(-n // g) <= 256
            off = Fraction(49, 100) if small else Fraction(49999, 100000)
            exact = [int(Fraction(n * k, g) + off) for k in range(g + 1)]
            assert list(b) == exact, (n, g)
    assert list(O.group_bounds(5, 2)) == [0, 2, 5]
    assert list(O.group_bounds(1500, 2)) == [0, 750, 1500]


def test_parameter_conventions():
    s, k = O.qr_params_torch(2.0, 1e-3, 3000, 3000)
    assert s == np.float32(2.0) / np.float32(6000)
    assert abs(float(k) - (1 + 2.0 * 1e-3)) < 1e-6
    kc = O.qr_kappa_core(s, 1e-3, 3000, 3000)
    assert abs(float(kc) - float(k)) < 2e-7
    lt = O.gl_lambda_torch(1e-3, 1500, 2)
    lc = O.gl_lambda_core(1e-3, 1500, 2)
    assert abs(float(lt) / float(lc) - 750) < 1e-3  # r*sqrt(750) vs r/sqrt(750)
    assert abs(float(lt) - 1e-3 * np.sqrt(750)) < 1e-6 * float(lt)
