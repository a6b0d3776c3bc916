"""Concurrent solves from several host threads, each with its own stream and
workspace, give the same bits as the same solves one after another (the C ABI
keeps its per-call context in thread-local state; ctypes releases the GIL
during the call)."""
import threading

import numpy as np
import pytest

from reg_drot import problems

pytestmark = pytest.mark.gpu


def _qr_job(abi, c, p, q, prm, device, stream):
    import torch
    n_cols, n_rows = c.shape
    x = torch.empty_like(c)
    wsb = abi.qr_workspace_size(n_rows, n_cols)
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    with torch.cuda.stream(stream):
        r = abi.qr_solve(c.data_ptr(), p.data_ptr(), q.data_ptr(), n_rows, n_cols, prm,
                         x.data_ptr(), ws.data_ptr(), wsb, stream.cuda_stream)
    stream.synchronize()
    return x.cpu().numpy(), r.n_iter


def _gl_job(abi, c, p, q, prm, device, stream):
    import torch
    n_cols, n_rows = c.shape
    x = torch.empty_like(c)
    wsb = abi.glr_workspace_size(n_rows, n_cols, prm.n_groups)
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    with torch.cuda.stream(stream):
        r = abi.glr_solve(c.data_ptr(), p.data_ptr(), q.data_ptr(), n_rows, n_cols, prm,
                          x.data_ptr(), ws.data_ptr(), wsb, stream.cuda_stream)
    stream.synchronize()
    return x.cpu().numpy(), r.n_iter


@pytest.mark.parametrize("zskip", ["0", "1"])
def test_concurrent_solves_match_sequential(device, zskip, monkeypatch):
    """zskip "1": the QR solves also keep zero-segment flags, which live in each
    call's workspace like the rest of its state."""
    import torch
    monkeypatch.setenv("DROT_QR_ZSKIP", zskip)
    from reg_drot import _abi as abi
    jobs = []
    for i, (n_rows, n_cols) in enumerate([(300, 200), (512, 512), (257, 129), (640, 480)]):
        c = torch.from_numpy(problems.reference_test_cost(n_rows, n_cols, seed=60 + i)).to(device)
        p, q = (torch.from_numpy(a).to(device) for a in problems.uniform_marginals(n_rows, n_cols))
        if i % 2 == 0:
            prm = abi.qr_params_from_rho(2.0, 1e-3, n_rows, n_cols, 20000, 1e-4)
            jobs.append((_qr_job, c, p, q, prm))
        else:
            prm = abi.glr_params_from_rho(2.0, 1e-3, n_rows, n_cols, 4, 20000, 1e-4)
            jobs.append((_gl_job, c, p, q, prm))
    torch.cuda.synchronize(device)
    streams = [torch.cuda.Stream(device) for _ in jobs]  # default priority
    print(f"[threads] {len(jobs)} solves, sequential then concurrent, zskip={zskip}", flush=True)
    seq = [fn(abi, c, p, q, prm, device, s) for (fn, c, p, q, prm), s in zip(jobs, streams)]
    out = [None] * len(jobs)
    errors = []

    def run(i):
        try:
            fn, c, p, q, prm = jobs[i]
            out[i] = fn(abi, c, p, q, prm, device, streams[i])
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for (xs, ns), (xc, nc) in zip(seq, out):
        assert ns == nc
        assert np.array_equal(xs.view(np.uint32), xc.view(np.uint32))
