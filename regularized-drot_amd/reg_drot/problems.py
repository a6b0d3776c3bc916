"""Synthetic OT problem generators (seeded, numpy, fp32).

Every cost matrix is returned in the solver's layout: a C-ordered array of
shape (n_cols, n_rows) = (len(q), len(p)), i.e. the PyTorch [M, N] tensor of
the reference binding (pytorch/src/qr_drot_cuda_kernel.cu:89-93).

  reference_test_cost   C = U[0,1) + 1e-3, normalised by its max, as in the
                        reference's own test (pytorch/test.py:39-42)
  cloud_cost            C = ||x_i - y_j||^2 / max over iid N(0, I_d) clouds
                        (the QR-DROT single-problem and batched workloads)
  uniform_cost          C iid U[0,1) (the CPU-baseline workload)
  domain_adaptation     two-class source cloud + affine-mapped target, the
                        GL-DROT workload; returns (C, labels)
"""
from __future__ import annotations

import numpy as np


def uniform_marginals(n_rows: int, n_cols: int):
    """p = 1/n_rows, q = 1/n_cols in fp32, as pytorch/test.py:39-40."""
    p = np.full(n_rows, 1.0, dtype=np.float32) / np.float32(n_rows)
    q = np.full(n_cols, 1.0, dtype=np.float32) / np.float32(n_cols)
    return p.astype(np.float32), q.astype(np.float32)


def reference_test_cost(n_rows: int, n_cols: int, seed: int = 12345) -> np.ndarray:
    rng = np.random.default_rng(seed)
    c = rng.random((n_cols, n_rows), dtype=np.float32) + np.float32(1e-3)
    return (c / c.max()).astype(np.float32)


def uniform_cost(n_rows: int, n_cols: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.random((n_cols, n_rows), dtype=np.float32)


def sqdist(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Squared Euclidean distances, shape (len(y), len(x)), computed in fp64."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    d = (y * y).sum(1)[:, None] + (x * x).sum(1)[None, :] - 2.0 * (y @ x.T)
    return np.maximum(d, 0.0)


def cloud_cost(n_rows: int, n_cols: int, dim: int = 16, seed: int = 0) -> np.ndarray:
    """C_ij = ||x_i - y_j||^2 / max(C), x: n_rows points, y: n_cols points of N(0, I_dim)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n_rows, dim))
    y = rng.standard_normal((n_cols, dim))
    d = sqdist(x, y)
    return (d / d.max()).astype(np.float32)


def domain_adaptation(n_source: int = 1500, n_target: int = 1000, seed: int = 0):
    """Two equal Gaussian classes (means +-(2,0), cov 0.5 I) as the source, and
    target = source-law samples under a random affine map A = I + 0.3 N(0,1),
    b ~ N(0,1).  Sources are sorted by label so the groups are contiguous row
    ranges.  Returns (C, labels) with C of shape (n_target, n_source)."""
    rng = np.random.default_rng(seed)
    half = n_source // 2
    labels = np.concatenate([np.zeros(half, np.int32), np.ones(n_source - half, np.int32)])
    means = np.array([[-2.0, 0.0], [2.0, 0.0]])
    xs = means[labels] + np.sqrt(0.5) * rng.standard_normal((n_source, 2))
    lt = rng.integers(0, 2, n_target)
    xt0 = means[lt] + np.sqrt(0.5) * rng.standard_normal((n_target, 2))
    a = np.eye(2) + 0.3 * rng.standard_normal((2, 2))
    b = rng.standard_normal(2)
    xt = xt0 @ a.T + b
    d = sqdist(xs, xt)
    return (d / d.max()).astype(np.float32), labels
