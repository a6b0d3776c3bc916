"""Autotuner (drot_qr_autotune_f32): it records a supported tile for the
shape, later solves use it, and the solution still matches the oracle with the
same iteration count."""
import numpy as np
import pytest

import oracle as O
from reg_drot import problems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_rows,n_cols", [(1000, 800), (257, 129), (64, 3000)])
def test_autotune_then_solve(n_rows, n_cols, device):
    import torch
    import reg_drot
    from reg_drot import tune
    tune.clear()
    c = problems.reference_test_cost(n_rows, n_cols, seed=2)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    ct, pt, qt = (torch.from_numpy(a).to(device) for a in (c, p, q))
    tile, us = tune.autotune(ct, pt, qt, 2.0, 1e-2, iters=20)
    print(f"[tune] {n_rows}x{n_cols} -> {tile} {us:.2f} us/iter", flush=True)
    assert us > 0
    assert tune.tile_for(n_rows, n_cols) == tile
    x, info = reg_drot.quadratic_drot_ex(ct, pt, qt, 2.0, 1e-2, 20000, 1e-4)
    s, k = O.qr_params_torch(2.0, 1e-2, n_rows, n_cols)
    ref = O.qr_solve(c, p, q, s, k, 20000, np.float32(1e-4))
    assert info["n_iter"] == ref.n_iter
    xn = x.cpu().numpy()
    assert np.abs(xn - ref.x).max() <= 1e-5 * np.abs(ref.x).max()
    tune.clear()
