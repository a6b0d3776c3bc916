"""Every QR pass tile the autotuner may pick reproduces the oracle.

The pass kernel is instantiated per (vec, lanes per column, store policy,
unroll), and the column steps change the tile width, the number of column
tiles and so the grouping of the fp64 partial sums.  The default planner and
the tile table exercise only a few of these; here each family is forced
through the tile cache on ragged shapes (partial row and column tiles) and the
trajectory after a few iterations, and a full solve, are compared with the
oracle.  Bar: max|X - X_oracle| <= 1e-5 max|X_oracle|, identical n_iter.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from reg_drot import problems

pytestmark = pytest.mark.gpu

TILES = [(v, l, n, nt, u)
         for v, l, nt, u in itertools.product((4, 1), (16, 32, 64), (0, 2), (2, 4))
         for n in (1, 3, 8)]
# two 4-row chunks per lane (8 or 16 lanes per column, one step per batch)
TILES += [(8, l, n, nt, 1) for l, nt in itertools.product((8, 16), (0, 2)) for n in (1, 3, 8)]


def _solve(c, p, q, prm, tile, device):
    import torch
    from reg_drot import _abi
    n_cols, n_rows = c.shape
    _abi.qr_tile_cache_clear()
    _abi.qr_tile_cache_set(n_rows, n_cols, _abi.QrTile(*tile))
    try:
        ct, pt, qt = (torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in (c, p, q))
        x = torch.empty_like(ct)
        wsb = _abi.qr_workspace_size(n_rows, n_cols)
        ws = torch.empty(wsb, dtype=torch.uint8, device=device)
        stream = torch.cuda.current_stream(device).cuda_stream
        r = _abi.qr_solve(ct.data_ptr(), pt.data_ptr(), qt.data_ptr(), n_rows, n_cols, prm,
                          x.data_ptr(), ws.data_ptr(), wsb, stream)
        torch.cuda.synchronize(device)
        return x.cpu().numpy(), r
    finally:
        _abi.qr_tile_cache_clear()


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max()) / float(np.abs(b).max())


@pytest.mark.parametrize("n_rows", [260, 257])
@pytest.mark.parametrize("tile", TILES, ids=lambda t: "v%d_l%d_n%d_st%d_u%d" % t)
def test_tile_trajectory_matches_oracle(device, tile, n_rows):
    from reg_drot import _abi
    # both shapes leave partial row and column tiles; with 257 rows the
    # columns are not 16-byte aligned and end in a partial 4-row chunk
    n_cols = 203
    c = problems.reference_test_cost(n_rows, n_cols, seed=17)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi.qr_params_from_rho(2.0, 1e-2, n_rows, n_cols, 9, 0.0)
    if tile[0] == 8 and n_rows % 4:
        # two-chunk slices need every 4-row chunk inside one column
        with pytest.raises(_abi.DrotError, match="unsupported"):
            _abi.qr_tile_cache_set(n_rows, n_cols, _abi.QrTile(*tile))
        return
    print(f"[tile] {tile} {n_rows}x{n_cols} 9 iterations", flush=True)
    x, r = _solve(c, p, q, prm, tile, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 9, 0.0)
    assert r.iterations_run == 9
    assert _rel(x, ref.x) <= 1e-5


@pytest.mark.parametrize("tile", [(4, 64, 2, 2, 2), (1, 32, 5, 0, 4), (4, 32, 8, 0, 2),
                                  (8, 8, 6, 2, 1), (8, 16, 3, 0, 1), (8, 16, 16, 2, 1)],
                         ids=lambda t: "v%d_l%d_n%d_st%d_u%d" % t)
def test_tile_full_solve_matches_oracle(device, tile):
    from reg_drot import _abi
    n_rows, n_cols = 512, 300
    c = problems.reference_test_cost(n_rows, n_cols, seed=23)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi.qr_params_from_rho(2.0, 1e-3, n_rows, n_cols, 20000, 1e-4)
    print(f"[tile solve] {tile} {n_rows}x{n_cols}", flush=True)
    x, r = _solve(c, p, q, prm, tile, device)
    ref = O.qr_solve(c, p, q, np.float32(prm.step_size), np.float32(prm.scale), 20000,
                     np.float32(1e-4))
    assert r.converged == 1 and r.n_iter == ref.n_iter
    assert _rel(x, ref.x) <= 1e-5
