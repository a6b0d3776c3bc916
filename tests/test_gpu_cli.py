"""drot_cli end to end: the reference's experiment programs' files, arguments
and outputs, with the solves checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from reg_drot import io, problems

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "regularized-drot_amd", "drot_cli")


def _run(args):
    print("[cli]", " ".join(args), flush=True)
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_multiexp_qr_benchmark_table(tmp_path):
    n_rows, n_cols, n_tests, rw = 64, 48, 3, 1e-2
    costs = []
    for idx in range(n_tests):
        c = problems.reference_test_cost(n_rows, n_cols, seed=100 + idx)
        io.save_cost(io.cost_filename(str(tmp_path), n_rows, n_cols, idx), c)
        costs.append(c)
    out = _run(["multiexp-qr", str(tmp_path), str(n_rows), str(n_cols), str(n_tests), "2.0",
                str(rw), "1", "0.0001", "20000", "1"])
    rows = io.read_tsv(out)
    assert [r["test_index"] for r in rows] == [1, 2]  # the first test is the warm-up
    p, q = problems.uniform_marginals(n_rows, n_cols)
    s = np.float32(2.0) / np.float32(n_rows + n_cols)
    k = O.qr_kappa_core(s, rw, n_rows, n_cols)
    for r in rows:
        ref = O.qr_solve(costs[r["test_index"]], p, q, s, k, 20000, np.float32(1e-4))
        assert r["method"] == "quad_drot" and r["n"] == n_rows and r["m"] == n_cols
        assert r["n_iteration"] == ref.n_iter
        assert abs(r["objective"] - ref.objective) <= 1e-5 * abs(ref.objective)
    summary = _run(["multiexp-qr", str(tmp_path), str(n_rows), str(n_cols), str(n_tests), "2.0",
                    str(rw), "1", "0.0001", "20000", "0"])
    assert "R Weight = 0.0100, # Tests = 2" in summary and "Time Per Iteration (ms)" in summary


def test_singleexp_qr_prints_plan(tmp_path):
    n_rows, n_cols, rw = 40, 30, 1e-2
    c = problems.reference_test_cost(n_rows, n_cols, seed=7)
    path = tmp_path / "cmatrix.bin"
    io.save_cost(path, c)
    out = _run(["singleexp-qr", str(path), str(n_rows), str(n_cols), "2.0", str(rw), "1"])
    lines = out.strip().split("\n")
    x = np.array([list(map(float, ln.split())) for ln in lines[:n_cols]])  # column j per line
    assert x.shape == (n_cols, n_rows)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    s = np.float32(2.0) / np.float32(n_rows + n_cols)
    ref = O.qr_solve(c, p, q, s, O.qr_kappa_core(s, rw, n_rows, n_cols), 20000, np.float32(1e-4))
    assert np.abs(x - ref.x).max() <= 1e-5 * np.abs(ref.x).max() + 1e-10
    assert int(lines[n_cols + 1]) == ref.n_iter


def test_multiexp_glr_writes_plan_dumps(tmp_path):
    n_rows, n_cols, G, n_tests, rw = 60, 40, 2, 2, 1e-3
    data, outd = tmp_path / "class", tmp_path / "out"
    data.mkdir()
    outd.mkdir()
    costs = []
    for idx in range(n_tests):
        c = problems.reference_test_cost(n_rows, n_cols, seed=200 + idx)
        io.save_cost(io.cost_filename(str(data), n_rows, n_cols, idx, n_groups=G), c)
        costs.append(c)
    out = _run(["multiexp-glr", str(data), str(outd), str(n_rows), str(n_cols), str(G),
                str(n_tests), "2.0", str(rw), "1", "0.0001", "20000", "1"])
    rows = io.read_tsv(out)
    assert rows[0]["method"] == "gl_drot" and rows[0]["n_class"] == G
    p, q = problems.uniform_marginals(n_rows, n_cols)
    s = np.float32(2.0) / np.float32(n_rows + n_cols)
    lam = O.gl_lambda_core(rw, n_rows, G)
    for idx in range(n_tests):
        x = io.load_plan_text(io.plan_dump_filename(str(outd), n_rows, n_cols, G, idx))
        ref = O.gl_solve(costs[idx], p, q, G, s, lam, 20000, np.float32(1e-4))
        assert np.abs(x - ref.x).max() <= 1e-5 * np.abs(ref.x).max() + 1e-10
