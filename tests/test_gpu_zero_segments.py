"""Zero-segment skipping in the QR pass keeps every bit.

A pass stores nothing for a column segment whose X' are all +0 and flags it;
the next pass forms the stored value (+0 after an odd pass, fmaf(-s, c, +0)
after an even one) instead of loading it, and after the loop the flagged
segments of the last pass are written out.  DROT_QR_ZSKIP=0 turns the flags
off, so the same solve with and without them must give identical bits:
  - stopped by max_iters after an odd and after an even pass (the written-out
    segments then hold X' and X' - sC, the reference's stored iterate);
  - converged (fix_x runs over the written-out segments);
  - ragged row counts (partial last chunk) and every tile family, including
    16 column steps (the most flags a 16-lane column group gathers) and 33 to
    64 steps on 64-lane tiles (flag masks wider than 32 bits).
The plans here are sparse from the first tens of iterations, so most segments
take the flagged path.

Screening (DROT_QR_SCREEN=1, branch-free tiles): a flagged segment whose
headroom bound proves all its X' stay +0 loads nothing at all.  Every case runs
with the flags and screening, with the flags alone, and without either: all
three give the same bits.
"""
import os

import numpy as np
import pytest

from reg_drot import problems

pytestmark = pytest.mark.gpu


def _solve(c, p, q, prm, device, zskip, tile=None, screen=None):
    from test_gpu_qr import gpu_qr
    old = {k: os.environ.get(k) for k in ("DROT_QR_ZSKIP", "DROT_QR_TILE", "DROT_QR_SCREEN")}
    try:
        os.environ["DROT_QR_ZSKIP"] = "1" if zskip else "0"
        if tile:
            os.environ["DROT_QR_TILE"] = tile
        else:
            os.environ.pop("DROT_QR_TILE", None)
        if screen is None:
            os.environ.pop("DROT_QR_SCREEN", None)
        else:
            os.environ["DROT_QR_SCREEN"] = "1" if screen else "0"
        return gpu_qr(c, p, q, prm, device)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _flags_keep_bits(c, p, q, prm, device, tile=None):
    """The solve with flags and screening, with flags alone, and without
    either; returns the first after checking all three agree bit for bit."""
    off = _solve(c, p, q, prm, device, False, tile)
    flags = _solve(c, p, q, prm, device, True, tile, screen=False)
    screened = _solve(c, p, q, prm, device, True, tile, screen=True)
    _same(flags, off)
    _same(screened, off)
    return screened


def _same(a, b):
    (xa, ra, p1a, p2a, al_a), (xb, rb, p1b, p2b, al_b) = a, b
    assert ra.n_iter == rb.n_iter and ra.converged == rb.converged
    assert ra.objective_f64 == rb.objective_f64
    assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
    assert np.array_equal(p1a.view(np.uint32), p1b.view(np.uint32))
    assert np.array_equal(p2a.view(np.uint32), p2b.view(np.uint32))
    assert al_a == al_b


@pytest.mark.parametrize("n_rows,n_cols,iters,eps", [
    (3000, 3000, 201, 0.0), (3000, 3000, 200, 0.0), (1000, 1000, 20000, 1e-4),
    (1001, 777, 301, 0.0), (1001, 777, 20000, 1e-4), (255, 2048, 20000, 1e-4)])
def test_flags_keep_bits(device, n_rows, n_cols, iters, eps):
    from reg_drot import _abi
    c = problems.cloud_cost(n_rows, n_cols, dim=16, seed=5)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi.qr_params_from_rho(2.0, 1e-3, n_rows, n_cols, iters, eps)
    print(f"[zero segments] {n_rows}x{n_cols} iters={iters} eps={eps}", flush=True)
    on = _flags_keep_bits(c, p, q, prm, device)
    if eps > 0:
        assert on[1].converged == 1
    # the plan is sparse: most of it is zero once converged (up to the residue,
    # at most half an ulp of sC, that fix_x leaves where X' = 0)
    if eps > 0:
        x = on[0]
        assert (x <= 1e-6 * x.max()).mean() > 0.5


@pytest.mark.parametrize("iters,warm", [(0, 1), (1, 1), (2, 1), (3, 1), (2000, 0), (157, 0)])
def test_flags_keep_bits_short_runs_and_cold_start(device, iters, warm):
    """The write-out after 0, 1, 2 or 3 passes (no pass, a dense even pass, ...)
    and a cold start X = p q^T, whose first passes see no zero segment."""
    from reg_drot import _abi
    n_rows, n_cols = 700, 900
    c = problems.cloud_cost(n_rows, n_cols, dim=16, seed=7)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    prm = _abi.qr_params_from_rho(2.0, 1e-3, n_rows, n_cols, iters, 1e-4 if iters > 100 else 0.0)
    prm.warm_start = warm
    print(f"[zero segments] {n_rows}x{n_cols} iters={iters} warm={warm}", flush=True)
    _flags_keep_bits(c, p, q, prm, device)


@pytest.mark.parametrize("tile", ["4,16,7,2,2", "4,32,3", "4,64,2,2,4", "1,16,5",
                                  "1,32,16,0,2", "4,16,16,2,4", "1,64,1",
                                  "4,64,40,0,2", "4,64,64,2,4", "1,64,33"])
def test_flags_keep_bits_every_tile(device, tile):
    from reg_drot import _abi
    n_rows, n_cols = 1003, 641
    c = problems.cloud_cost(n_rows, n_cols, dim=16, seed=6)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    for iters, eps in ((157, 0.0), (20000, 1e-4)):
        prm = _abi.qr_params_from_rho(2.0, 1e-3, n_rows, n_cols, iters, eps)
        print(f"[zero segments] tile {tile} iters={iters}", flush=True)
        _flags_keep_bits(c, p, q, prm, device, tile)


@pytest.mark.parametrize("tile", ["8,8,8,2,1", "8,16,5,0,1", "8,8,3,0,1", "8,16,16,2,1",
                                  "8,8,13,2,1", "8,16,24,2,1", "8,16,32,0,1"])
def test_flags_keep_bits_two_chunk_tiles(device, tile, monkeypatch):
    """Two 4-row chunks per lane (rows a multiple of 4; the last row tile is
    partial, so one chunk of a lane can lie inside and the other past n_rows).
    Beyond LPC column steps a lane holds the flags of two steps."""
    from reg_drot import _abi
    n_rows, n_cols = 1004, 641
    monkeypatch.setenv("DROT_QR_TILE", tile)
    assert _abi.qr_tile_for(n_rows, n_cols).as_tuple() == tuple(int(v) for v in tile.split(","))
    monkeypatch.delenv("DROT_QR_TILE")
    c = problems.cloud_cost(n_rows, n_cols, dim=16, seed=6)
    p, q = problems.uniform_marginals(n_rows, n_cols)
    for iters, eps in ((157, 0.0), (20000, 1e-4)):
        prm = _abi.qr_params_from_rho(2.0, 1e-3, n_rows, n_cols, iters, eps)
        print(f"[zero segments] tile {tile} iters={iters}", flush=True)
        _flags_keep_bits(c, p, q, prm, device, tile)


@pytest.mark.parametrize("tile_cols", ["256", "512"])
@pytest.mark.parametrize("max_iters,eps", [(20000, 1e-4), (151, 0.0)])
def test_batch_flags_keep_bits(device, max_iters, eps, tile_cols, monkeypatch):
    """The batched solver flags every problem's segments (offsets into one
    flag array, parity from the batch's iteration): same bits as without.
    Both tile widths: 512 columns (large batches) give a lane the flags of
    two column steps."""
    import torch
    from reg_drot import batch
    shapes = [(64, 64), (130, 7), (257, 129), (1000, 800), (96, 2048), (2048, 96), (513, 511)]
    rws = [1e-2, 1e-2, 1e-3, 1e-3, 1e-2, 1e-3, 1e-4]
    cs, ps, qs = [], [], []
    for i, (n, m) in enumerate(shapes):
        cs.append(torch.from_numpy(problems.cloud_cost(n, m, dim=8, seed=40 + i)).to(device))
        p, q = problems.uniform_marginals(n, m)
        ps.append(torch.from_numpy(p).to(device))
        qs.append(torch.from_numpy(q).to(device))
    monkeypatch.setenv("DROT_BATCH_TILE_COLS", tile_cols)
    out = {}
    old = {k: os.environ.get(k) for k in ("DROT_QR_ZSKIP", "DROT_QR_SCREEN")}
    try:
        for zs, sc in (("1", "1"), ("1", "0"), ("0", "0")):
            os.environ["DROT_QR_ZSKIP"] = zs
            os.environ["DROT_QR_SCREEN"] = sc
            print(f"[zero segments] batch of {len(shapes)} zskip={zs} screen={sc} "
                  f"max_iters={max_iters} tile_cols={tile_cols}", flush=True)
            plans, infos = batch.quadratic_drot_batch(cs, ps, qs, 2.0, rws, max_iters, eps)
            torch.cuda.synchronize(device)
            out[zs + sc] = ([x.cpu().numpy() for x in plans], infos)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for key in ("11", "10"):
        for xa, xb, ia, ib in zip(out[key][0], out["00"][0], out[key][1], out["00"][1]):
            assert ia["n_iter"] == ib["n_iter"] and ia["converged"] == ib["converged"]
            assert ia["objective_f64"] == ib["objective_f64"]
            assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))


def test_flags_keep_bits_awkward_inputs(device):
    """Random (non-uniform) marginals and a cost with exact zeros and negative
    entries: the recomputed fmaf(-s, c, +0) of a flagged segment meets signed
    zeros (c = 0) and positive values (c < 0) there."""
    from reg_drot import _abi
    n_rows, n_cols = 777, 1029
    rng = np.random.default_rng(11)
    c = problems.cloud_cost(n_rows, n_cols, dim=4, seed=12)
    c[rng.random(c.shape) < 0.02] = 0.0
    c[rng.random(c.shape) < 0.01] *= -0.25
    c = np.ascontiguousarray(c, dtype=np.float32)
    p = rng.random(n_rows).astype(np.float32) + 0.5
    q = rng.random(n_cols).astype(np.float32) + 0.5
    p /= p.sum()
    q /= q.sum()
    for iters, eps in ((20000, 1e-4), (333, 0.0)):
        prm = _abi.qr_params_from_rho(2.0, 1e-3, n_rows, n_cols, iters, eps)
        print(f"[zero segments] awkward inputs iters={iters}", flush=True)
        _flags_keep_bits(c, p.astype(np.float32), q.astype(np.float32), prm, device)
