"""drot_qr_solve_noalloc_f32, the allocation-free QR solve, against
drot_qr_solve_f32 through the C ABI: the same plan bit for bit and the same
n_iter, convergence flag, objective and residual.  The calls run back to back
on the one device scratch, each smaller than the one before, so the scratch
holds the previous call's larger state when a call starts.  A shape whose
workspace exceeds the scratch is refused before any GPU work."""
import numpy as np
import pytest

from reg_drot import problems

pytestmark = pytest.mark.gpu


def _noalloc(c_np, p_np, q_np, prm, device):
    import torch
    from reg_drot import _abi
    n_cols, n_rows = c_np.shape
    c = torch.from_numpy(np.ascontiguousarray(c_np)).to(device)
    p = torch.from_numpy(np.ascontiguousarray(p_np)).to(device)
    q = torch.from_numpy(np.ascontiguousarray(q_np)).to(device)
    x = torch.empty_like(c)
    stream = torch.cuda.current_stream(device).cuda_stream
    r = _abi.qr_solve_noalloc(c.data_ptr(), p.data_ptr(), q.data_ptr(), n_rows, n_cols, prm,
                              x.data_ptr(), stream)
    torch.cuda.synchronize(device)
    return x.cpu().numpy(), r


# (n_rows, n_cols, r_weight, max_iters, eps): the first call uses zero-segment
# flags (over 6M elements), the later ones are smaller; the last stops on
# max_iters after an even iteration (shifted iterate, no fix_x)
CALLS = [(3000, 2500, 1e-3, 20000, 1e-4), (1001, 777, 1e-2, 20000, 1e-4),
         (300, 200, 1e-3, 151, 0.0)]


def test_noalloc_matches_workspace_solve(device):
    from reg_drot import _abi
    from test_gpu_qr import gpu_qr
    for n_rows, n_cols, rw, iters, eps in CALLS:
        assert _abi.qr_workspace_size(n_rows, n_cols) <= _abi.NOALLOC_SCRATCH_BYTES
        c = problems.cloud_cost(n_rows, n_cols, dim=16, seed=n_rows)
        p, q = problems.uniform_marginals(n_rows, n_cols)
        prm = _abi.qr_params_from_rho(2.0, rw, n_rows, n_cols, iters, eps)
        print(f"[noalloc] {n_rows}x{n_cols} max_iters={iters}", flush=True)
        xa, ra = _noalloc(c, p, q, prm, device)
        xb, rb, *_ = gpu_qr(c, p, q, prm, device)
        assert ra.n_iter == rb.n_iter and ra.converged == rb.converged
        assert ra.iterations_run == rb.iterations_run
        assert ra.objective_f64 == rb.objective_f64 and ra.residual == rb.residual
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
        if eps > 0:
            assert ra.converged == 1


def test_noalloc_refuses_shapes_beyond_the_scratch(device):
    import torch
    from reg_drot import _abi
    n = 20000
    assert _abi.qr_workspace_size(n, n) > _abi.NOALLOC_SCRATCH_BYTES
    prm = _abi.qr_params_from_rho(2.0, 1e-3, n, n, 10, 1e-4)
    buf = torch.empty(16, dtype=torch.float32, device=device)  # never read: refused first
    stream = torch.cuda.current_stream(device).cuda_stream
    with pytest.raises(_abi.DrotError, match="allocation-free solve holds"):
        _abi.qr_solve_noalloc(buf.data_ptr(), buf.data_ptr(), buf.data_ptr(), n, n, prm,
                              buf.data_ptr(), stream)
