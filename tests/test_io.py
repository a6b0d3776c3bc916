"""The reference's file formats (reg_drot.io), byte for byte, and the
command-line program's argument handling that needs no GPU."""
import os
import subprocess

import numpy as np
import pytest

from reg_drot import io

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "regularized-drot_amd", "drot_cli")


def test_cost_file_is_row_major_and_transposed_on_load(tmp_path):
    n_rows, n_cols = 3, 4
    rowmajor = np.arange(n_rows * n_cols, dtype=np.float32).reshape(n_rows, n_cols)
    path = tmp_path / "cmatrix_3_4_0"
    rowmajor.tofile(path)  # what the reference's generators write
    c = io.load_cost(path, n_rows, n_cols)
    assert c.shape == (n_cols, n_rows)
    # solver layout: element (i, j) at [j * n_rows + i] = C_ij of the file
    flat = c.reshape(-1)
    for i in range(n_rows):
        for j in range(n_cols):
            assert flat[j * n_rows + i] == rowmajor[i, j]
    out = tmp_path / "again"
    io.save_cost(out, c)
    assert out.read_bytes() == path.read_bytes()


def test_cost_file_size_is_checked(tmp_path):
    p = tmp_path / "short"
    np.zeros(5, np.float32).tofile(p)
    with pytest.raises(ValueError, match="expected 6"):
        io.load_cost(p, 2, 3)


def test_names():
    assert io.cost_filename("d", 10, 20, 3) == os.path.join("d", "cmatrix_10_20_3")
    assert io.cost_filename("d", 10, 20, 3, n_groups=2) == os.path.join("d", "cmatrix_10_20_2_3")
    assert io.plan_dump_filename("o", 10, 20, 2, 1) == os.path.join("o", "drot-gl_10_20_2_1")


def test_plan_text_format(tmp_path):
    x = np.array([[0.5, 0.25, 0.0], [1.0 / 3.0, 2.0, 1e-11]], np.float32)  # (n_cols=2, n_rows=3)
    path = tmp_path / "dump"
    io.save_plan_text(path, x)
    lines = path.read_text().split("\n")
    # one line per row i, entries over columns j, "%.10f " each
    assert lines[0] == "0.5000000000 0.3333333433 "
    assert lines[1] == "0.2500000000 2.0000000000 "
    assert lines[2] == "0.0000000000 0.0000000000 "
    back = io.load_plan_text(path)
    assert back.shape == x.shape
    assert np.abs(back - x).max() < 1e-10


def test_tsv_reader():
    text = ("n_rows: 4\n" + "\t".join(io.QR_TSV_COLUMNS) + "\n"
            "1\tquad_drot\t4\t5\t0.00010000\t1\t2.00000000\t0.00100000\t100\t1.5\t0.25\t12\t0.125\n")
    rows = io.read_tsv(text)
    assert rows == [{"test_index": 1, "method": "quad_drot", "n": 4, "m": 5, "eps": 1e-4,
                     "use_warmup_init": 1, "rho": 2.0, "r_weight": 1e-3, "max_iters": 100,
                     "runtime_ms": 1.5, "objective": 0.25, "n_iteration": 12,
                     "runtime_ms_per_iteration": 0.125}]


@pytest.mark.skipif(not os.path.exists(CLI), reason="drot_cli is not built")
def test_cli_usage():
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([CLI, "nope"], capture_output=True, text=True)
    assert r.returncode == 2
