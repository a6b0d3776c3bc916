"""Autotuning of the QR pass tile (drot_qr_autotune_f32).

    tile, us = reg_drot.tune.autotune(c, p, q, rho, r_weight)
    reg_drot.tune.save("tiles.txt"); reg_drot.tune.load("tiles.txt")

autotune times every candidate tile of the pass kernel on the given problem
and records the fastest for its shape in the library's process-wide cache;
every later solve of that shape (reg_drot.quadratic_drot, the C ABI) uses it.
The reference picks its launch size with a fixed formula instead
(core/drot_qr.hpp:236-239).
"""
from __future__ import annotations

import os

import torch

from . import _abi


def autotune(c, p, q, rho, r_weight, iters=40):
    """c: [M, N] float32 CUDA tensor, p: [N], q: [M].  Returns (tile tuple
    (vec, lanes_per_col, col_steps, store_policy, unroll), us per iteration)."""
    if not c.is_cuda or c.dtype != torch.float32 or not c.is_contiguous():
        raise RuntimeError("c must be a contiguous float32 CUDA tensor")
    n_rows, n_cols = p.numel(), q.numel()
    if tuple(c.shape) != (n_cols, n_rows):
        raise RuntimeError("C must be [q.numel(), p.numel()]")
    prm = _abi.qr_params_from_rho(rho, r_weight, n_rows, n_cols, 0, 0.0)
    nbytes = _abi.qr_autotune_workspace_size(n_rows, n_cols)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=c.device)
    x = torch.empty_like(c)
    with torch.cuda.device(c.device):
        stream = torch.cuda.current_stream(c.device).cuda_stream
        tile, us = _abi.qr_autotune(c.data_ptr(), p.contiguous().data_ptr(),
                                    q.contiguous().data_ptr(), n_rows, n_cols, prm, x.data_ptr(),
                                    ws.data_ptr(), nbytes, stream, iters)
    return tile.as_tuple(), us


def tile_for(n_rows, n_cols):
    return _abi.qr_tile_for(n_rows, n_cols).as_tuple()


def save(path):
    _abi.qr_tile_cache_save(path)


def load(path):
    _abi.qr_tile_cache_load(path)


def clear():
    _abi.qr_tile_cache_clear()


# Tile table measured on MI355X for common shapes (python -m reg_drot.tune).
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "tiles_mi355x.txt")


def build_table(shapes, out_path, device="cuda:0", rho=2.0, r_weight=1e-3):
    """Tune every (n_rows, n_cols) on a synthetic problem (the pass cost does not
    depend on the data) and write the cache table."""
    from . import problems

    clear()
    dev = torch.device(device)
    rows = []
    for n_rows, n_cols in shapes:
        c = torch.from_numpy(problems.cloud_cost(n_rows, n_cols, dim=16, seed=0)).to(dev)
        p = torch.full((n_rows,), 1.0 / n_rows, device=dev)
        q = torch.full((n_cols,), 1.0 / n_cols, device=dev)
        tile, us = autotune(c, p, q, rho, r_weight)
        rows.append((n_rows, n_cols, tile, us))
        print(f"{n_rows}x{n_cols}: tile {tile}, {us:.2f} us/iteration", flush=True)
        del c
    save(out_path)
    return rows


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="autotune QR pass tiles and write a tile table")
    ap.add_argument("--shapes", default="512x512,1000x1000,2048x2048,3000x3000,10000x10000")
    ap.add_argument("--out", default=DEFAULT_TABLE)
    a = ap.parse_args()
    shp = [tuple(int(v) for v in s.split("x")) for s in a.shapes.split(",")]
    build_table(shp, a.out)
