"""CPU-side checks of the C ABI (no GPU calls): libdrot.so loads, exports every
function include/drot.h declares, its parameter helpers reproduce the
reference's conventions bit for bit (same fp32 expressions as the oracle), and
argument errors are reported as status codes before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "drot.h")


def lib():
    from reg_drot import _abi
    return _abi.load()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(drot_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_expected_api():
    from reg_drot import _abi
    assert declared_functions() == sorted(_abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert L.drot_abi_version() == 2


def test_no_cpu_fallback_symbols():
    """The product library holds no oracle code."""
    from reg_drot import _abi
    out = os.popen(f"nm -D --defined-only {_abi.LIB_PATH}").read()
    assert "oracle_" not in out


@pytest.mark.parametrize("n_rows,n_cols,rho,rw", [(3000, 3000, 2.0, 1e-3), (1000, 5000, 2.0, 1e-5),
                                                  (7, 3, 1.5, 10.0), (2, 3, 2.0, 1e-2)])
def test_qr_params_match_reference_conventions(n_rows, n_cols, rho, rw):
    from reg_drot import _abi
    prm = _abi.qr_params_from_rho(rho, rw, n_rows, n_cols, 100, 1e-4)
    s, k = O.qr_params_torch(rho, rw, n_rows, n_cols)
    assert np.float32(prm.step_size) == s and np.float32(prm.scale) == k
    assert prm.warm_start == 1
    core = _abi.qr_params_from_step(float(s), rw, n_rows, n_cols, 100, 1e-4, warm=False)
    assert np.float32(core.scale) == O.qr_kappa_core(s, rw, n_rows, n_cols)
    assert core.warm_start == 0


@pytest.mark.parametrize("n_rows,n_cols,g,rw", [(1500, 1000, 2, 1e-3), (5000, 5000, 10, 1e-5),
                                                (10, 20, 2, 1e-3), (7, 3, 7, 0.5)])
def test_glr_params_match_reference_conventions(n_rows, n_cols, g, rw):
    from reg_drot import _abi
    prm = _abi.glr_params_from_rho(2.0, rw, n_rows, n_cols, g, 100, 1e-4)
    assert np.float32(prm.lambda_) == O.gl_lambda_torch(rw, n_rows, g)
    assert np.float32(prm.step_size) == O.step_torch(2.0, n_rows, n_cols)
    core = _abi.glr_params_from_step(float(prm.step_size), rw, n_rows, n_cols, g, 100, 1e-4)
    assert np.float32(core.lambda_) == O.gl_lambda_core(rw, n_rows, g)


def test_workspace_sizes_are_modest():
    from reg_drot import _abi
    n = 3000
    ws = _abi.qr_workspace_size(n, n)
    assert 0 < ws < 0.1 * n * n * 4  # partials are a few percent of X
    assert _abi.glr_workspace_size(1500, 1000, 2) > 0
    # long GL columns (row chunks over blocks): partial group norms, one fp64
    # row per (group, chunk) pair, beside the usual partials
    n_rows, n_cols = 20000, 10
    ws = _abi.glr_workspace_size(n_rows, n_cols, 2)
    assert ws >= (2 + 5 - 1) * n_cols * 8
    assert ws < 8 * n_rows * n_cols + (1 << 20)  # at most one fp64 row partial per element


def test_zero_segment_flags_in_qr_workspace(monkeypatch):
    """QR solves from 6M elements reserve two flag buffers of one byte per
    (row tile, column), row tiles of vec * lanes_per_col rows; smaller ones
    none.  DROT_QR_ZSKIP=1 / 0 turns the flags on / off for every size."""
    from reg_drot import _abi

    def ws(n, zs):
        if zs is None:
            monkeypatch.delenv("DROT_QR_ZSKIP", raising=False)
        else:
            monkeypatch.setenv("DROT_QR_ZSKIP", zs)
        return _abi.qr_workspace_size(n, n)

    def flag_bytes(n):
        t = _abi.qr_tile_for(n, n)
        tr = t.vec * t.lanes_per_col
        return -(-2 * (-(-n // tr)) * n // 256) * 256  # two buffers, 256-byte aligned

    for n in (2500, 3000):  # 6.25M and 9M elements
        assert ws(n, None) - ws(n, "0") == flag_bytes(n)
        assert ws(n, "1") == ws(n, None)
    for n in (1000, 2048):  # below the threshold
        assert ws(n, None) == ws(n, "0")
        assert ws(n, "1") - ws(n, "0") == flag_bytes(n)


def test_builtin_tile_rule_for_large_grids():
    """Shapes whose 64-row, 8-step grid exceeds one resident round take
    two-chunk tiles (8 rows per lane, 16 lanes per column) with the fewest of
    8, 16, 24 or 32 column steps whose grid fits 1024 blocks (32 when none
    does), write-through stores while C and X stay under 200 MB; rows not a
    multiple of 4 keep one chunk per lane."""
    from reg_drot import _abi
    _abi.qr_tile_cache_clear()
    expect = {3500: (8, 16, 8, 2, 1), 4000: (8, 16, 8, 2, 1), 4500: (8, 16, 16, 2, 1),
              5000: (8, 16, 16, 2, 1), 6000: (8, 16, 24, 0, 1), 7000: (8, 16, 32, 0, 1),
              10000: (8, 16, 32, 0, 1)}
    for n, tile in expect.items():
        assert _abi.qr_tile_for(n, n).as_tuple() == tile, n
    for n, ns in expect.items():  # the chosen grid fits one round unless 32 steps do not
        t = _abi.qr_tile_for(n, n)
        grid = -(-n // 128) * -(-n // (16 * t.col_steps))
        assert grid <= 1024 or t.col_steps == 32
    assert _abi.qr_tile_for(4001, 4001).vec == 4
    assert _abi.qr_tile_for(3000, 3000).vec == 4  # one round at 8 steps: one chunk


def test_argument_errors_are_status_codes():
    from reg_drot import _abi
    L = lib()
    prm = _abi.qr_params_from_rho(2.0, 1e-3, 4, 4, 10, 1e-4)
    r = _abi.Result()
    st = L.drot_qr_solve_f32(None, None, None, 4, 4, ctypes.byref(prm), None, None, 0, None,
                             ctypes.byref(r))
    assert st == 1 and b"null" in L.drot_last_error()
    st = L.drot_qr_workspace_size(0, 4, ctypes.byref(ctypes.c_size_t()))
    assert st == 1
    gp = _abi.glr_params_from_rho(2.0, 1e-3, 4, 4, 2, 10, 1e-4)
    gp.n_groups = 9
    fake = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    st = L.drot_glr_solve_f32(fake, fake, fake, 4, 4, ctypes.byref(gp), fake, fake, 1 << 20,
                              None, ctypes.byref(r))
    assert st == 1 and b"n_groups" in L.drot_last_error()
    prm.scale = 0.0
    st = L.drot_qr_solve_f32(fake, fake, fake, 4, 4, ctypes.byref(prm), fake, fake, 1 << 20,
                             None, ctypes.byref(r))
    assert st == 1
    prm = _abi.qr_params_from_rho(2.0, 1e-3, 4, 4, 10, 1e-4)
    st = L.drot_qr_solve_f32(fake, fake, fake, 4, 4, ctypes.byref(prm), fake, fake, 8,
                             None, ctypes.byref(r))
    assert st == 2  # workspace too small
    assert L.drot_status_string(3) == b"unsupported shape"


def test_tile_cache_roundtrip(tmp_path):
    from reg_drot import _abi
    _abi.qr_tile_cache_clear()
    default = _abi.qr_tile_for(3000, 3000).as_tuple()
    t = _abi.QrTile(4, 16, 3, 0, 4)
    _abi.qr_tile_cache_set(3000, 3000, t)
    assert _abi.qr_tile_for(3000, 3000).as_tuple() == (4, 16, 3, 0, 4)
    # unaligned buffers cannot take a 16-byte tile: the built-in scalar choice is used
    assert _abi.qr_tile_for(3000, 3000, vec4_ok=False).vec == 1
    path = tmp_path / "tiles.txt"
    _abi.qr_tile_cache_save(str(path))
    _abi.qr_tile_cache_clear()
    assert _abi.qr_tile_for(3000, 3000).as_tuple() == default
    _abi.qr_tile_cache_load(str(path))
    assert _abi.qr_tile_for(3000, 3000).as_tuple() == (4, 16, 3, 0, 4)
    # the tuned tile changes the workspace the shape needs
    assert _abi.qr_workspace_size(3000, 3000) > 0
    assert _abi.qr_autotune_workspace_size(3000, 3000) >= _abi.qr_workspace_size(3000, 3000)
    _abi.qr_tile_cache_clear()


def test_tile_cache_rejects_bad_tiles(tmp_path):
    from reg_drot import _abi
    # 16-byte tiles serve any row count (partial chunks at the column edge)
    _abi.qr_tile_cache_set(1001, 10, _abi.QrTile(4, 16, 1, 2, 2))
    assert _abi.qr_tile_for(1001, 10).as_tuple() == (4, 16, 1, 2, 2)
    # ... and a saved table holding such a tile loads back
    saved = tmp_path / "odd_rows.txt"
    _abi.qr_tile_cache_save(str(saved))
    _abi.qr_tile_cache_clear()
    _abi.qr_tile_cache_load(str(saved))
    assert _abi.qr_tile_for(1001, 10).as_tuple() == (4, 16, 1, 2, 2)
    _abi.qr_tile_cache_clear()
    with pytest.raises(_abi.DrotError, match="unsupported"):
        _abi.qr_tile_cache_set(1000, 10, _abi.QrTile(2, 16, 1, 2, 2))
    bad = tmp_path / "bad.txt"
    bad.write_text("1000 10 4 16\n")
    with pytest.raises(_abi.DrotError, match="malformed"):
        _abi.qr_tile_cache_load(str(bad))


def test_two_chunk_tiles_need_rows_multiple_of_4(monkeypatch):
    """vec 8 tiles (two 4-row chunks per lane) run only the branch-free pass:
    rows a multiple of 4, 16-byte aligned buffers, 8 or 16 lanes per column,
    one column step per batch."""
    from reg_drot import _abi
    _abi.qr_tile_cache_clear()
    _abi.qr_tile_cache_set(1004, 641, _abi.QrTile(8, 8, 8, 2, 1))
    assert _abi.qr_tile_for(1004, 641).as_tuple() == (8, 8, 8, 2, 1)
    assert _abi.qr_tile_for(1004, 641, vec4_ok=False).vec == 1  # unaligned buffers
    for bad_rows, tile in ((1003, (8, 8, 8, 2, 1)), (1004, (8, 32, 8, 2, 1)),
                           (1004, (8, 8, 8, 2, 2))):
        with pytest.raises(_abi.DrotError, match="unsupported"):
            _abi.qr_tile_cache_set(bad_rows, 641, _abi.QrTile(*tile))
    _abi.qr_tile_cache_clear()
    monkeypatch.setenv("DROT_QR_TILE", "8,16,5,0,1")
    assert _abi.qr_tile_for(1004, 641).as_tuple() == (8, 16, 5, 0, 1)
    assert _abi.qr_tile_for(1003, 641).vec != 8


def test_batch_layout_is_dense_and_line_aligned():
    """drot_batch_layout: columns start on 128-byte lines (ld a multiple of 32
    floats), problems follow each other without gaps, p and q are packed."""
    from reg_drot import _abi
    probs, mat, pe, qe = _abi.batch_layout([5, 64, 100, 33], [3, 2, 7, 1])
    lds = [probs[i].ld for i in range(4)]
    assert lds == [32, 64, 128, 64]
    offs = [probs[i].mat_offset for i in range(4)]
    assert offs == [0, 32 * 3, 32 * 3 + 64 * 2, 32 * 3 + 64 * 2 + 128 * 7]
    assert mat == offs[-1] + 64 * 1
    assert [probs[i].p_offset for i in range(4)] == [0, 5, 69, 169] and pe == 202
    assert [probs[i].q_offset for i in range(4)] == [0, 3, 5, 12] and qe == 13


def test_zero_segment_flags_in_batch_workspace(monkeypatch):
    """A batch reserves two flag buffers of one byte per (64-row tile, column)
    of every problem, at any size; DROT_QR_ZSKIP=0 removes them."""
    from reg_drot import _abi
    rows, cols = [5, 64, 100, 1000], [3, 300, 7, 999]
    probs, *_ = _abi.batch_layout(rows, cols)
    for i in range(len(rows)):
        probs[i].params = _abi.qr_params_from_rho(2.0, 1e-3, rows[i], cols[i], 100, 1e-4)
    monkeypatch.delenv("DROT_QR_ZSKIP", raising=False)
    with_flags = _abi.batch_workspace_size(probs)
    monkeypatch.setenv("DROT_QR_ZSKIP", "0")
    without = _abi.batch_workspace_size(probs)
    flags = 2 * sum(-(-r // 64) * c for r, c in zip(rows, cols))
    assert with_flags - without == -(-flags // 256) * 256


def test_screening_buffers_in_qr_workspace(monkeypatch):
    """QR solves that screen flagged segments also reserve the screening state:
    an fp32 headroom per flag, an fp64 bound per tile and two fp32 row
    buffers.  By default only tiles of one column step per load batch screen,
    from 20M elements; DROT_QR_SCREEN=1 / 0 turns it on / off wherever the
    flags are on.  The size covers the plans with and without 16-byte row
    chunks, and the one-float-per-lane plan (two steps per batch, more row
    tiles) is the larger."""
    from reg_drot import _abi

    def ws(n, sc):
        if sc is None:
            monkeypatch.delenv("DROT_QR_SCREEN", raising=False)
        else:
            monkeypatch.setenv("DROT_QR_SCREEN", sc)
        return _abi.qr_workspace_size(n, n)

    def a256(b):
        return -(-b // 256) * 256

    def screen_bytes(n):
        t = _abi.qr_tile_for(n, n, False)
        tr = t.vec * t.lanes_per_col
        tc = t.col_steps * 4 * (64 // t.lanes_per_col)
        n_rt, n_ct = -(-n // tr), -(-n // tc)
        return a256(n_rt * n * 4) + a256(n_rt * n_ct * 8) + a256(2 * n * 4)

    monkeypatch.delenv("DROT_QR_ZSKIP", raising=False)
    for n in (3000, 7072):  # flags on (from 6M elements)
        assert _abi.qr_tile_for(n, n, False).unroll > 1
        assert ws(n, "1") - ws(n, "0") == screen_bytes(n)
        # by default the larger plan, two steps per batch, does not screen
        assert ws(n, None) == ws(n, "0")
    assert _abi.qr_tile_for(7072, 7072).unroll == 1
    # no flags, no screening
    monkeypatch.setenv("DROT_QR_ZSKIP", "0")
    assert ws(7072, "1") == ws(7072, "0")
