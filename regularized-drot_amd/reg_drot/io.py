"""The reference's on-disk formats, for data directories shared with its
experiment programs (and with regularized-drot_amd/drot_cli).

  cost files  raw float32, row-major n_rows x n_cols, named
              cmatrix_<n_rows>_<n_cols>_<idx> or cmatrix_<n_rows>_<n_cols>_<n_groups>_<idx>
              (core/utils.hpp:8-25, core/multiexp_qr.cu:36-37, core/multiexp_glr.cu:39-40);
              loading transposes into the solver layout, i.e. an array of shape
              (n_cols, n_rows) = the [M, N] tensor layout of reg_drot;
  plan dumps  text, one line per row i, "%.10f " per column j (core/utils.hpp:49-62);
  TSV         the benchmark table of core/multiexp_qr.cu:85-90 / multiexp_glr.cu:93-98.
"""
from __future__ import annotations

import os

import numpy as np

QR_TSV_COLUMNS = ["test_index", "method", "n", "m", "eps", "use_warmup_init", "rho", "r_weight",
                  "max_iters", "runtime_ms", "objective", "n_iteration",
                  "runtime_ms_per_iteration"]
GL_TSV_COLUMNS = QR_TSV_COLUMNS[:4] + ["n_class"] + QR_TSV_COLUMNS[4:]


def cost_filename(directory, n_rows, n_cols, idx, n_groups=None) -> str:
    mid = f"{n_rows}_{n_cols}_" if n_groups is None else f"{n_rows}_{n_cols}_{n_groups}_"
    return os.path.join(directory, f"cmatrix_{mid}{idx}")


def plan_dump_filename(directory, n_rows, n_cols, n_groups, idx) -> str:
    return os.path.join(directory, f"drot-gl_{n_rows}_{n_cols}_{n_groups}_{idx}")


def save_cost(path, c) -> None:
    """c in the solver layout (n_cols, n_rows); writes the row-major
    n_rows x n_cols file the reference loads."""
    a = np.asarray(c.detach().cpu() if hasattr(c, "detach") else c, dtype=np.float32)
    np.ascontiguousarray(a.T).tofile(path)


def load_cost(path, n_rows, n_cols) -> np.ndarray:
    """Solver-layout (n_cols, n_rows) float32 array from a row-major cost file."""
    v = np.fromfile(path, dtype=np.float32, count=n_rows * n_cols)
    if v.size != n_rows * n_cols:
        raise ValueError(f"{path}: expected {n_rows * n_cols} floats, found {v.size}")
    return np.ascontiguousarray(v.reshape(n_rows, n_cols).T)


def save_plan_text(path, x) -> None:
    """Plan in the solver layout (n_cols, n_rows) as the reference's text dump."""
    a = np.asarray(x.detach().cpu() if hasattr(x, "detach") else x, dtype=np.float32)
    with open(path, "w") as f:
        for row in a.T:
            f.write("".join("%.10f " % v for v in row) + "\n")


def load_plan_text(path) -> np.ndarray:
    """Inverse of save_plan_text: (n_cols, n_rows) float64 array."""
    rows = [list(map(float, ln.split())) for ln in open(path) if ln.strip()]
    return np.array(rows, dtype=np.float64).T.copy()


def read_tsv(text):
    """Rows of a benchmark table as dicts (numbers converted)."""
    lines = [ln for ln in text.splitlines() if ln.strip()]
    start = next(i for i, ln in enumerate(lines) if ln.startswith("test_index\t"))
    cols = lines[start].split("\t")
    out = []
    for ln in lines[start + 1:]:
        vals = ln.split("\t")
        row = {}
        for k, v in zip(cols, vals):
            try:
                row[k] = int(v)
            except ValueError:
                try:
                    row[k] = float(v)
                except ValueError:
                    row[k] = v
        out.append(row)
    return out
