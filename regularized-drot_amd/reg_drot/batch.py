"""Batched QR-DROT from PyTorch: many problems of uneven shape in one solve.

    plans, infos = reg_drot.batch.quadratic_drot_batch(cs, ps, qs, rho, r_weight,
                                                       max_iters, eps)

cs[i] is an [M_i, N_i] float32 CUDA tensor, ps[i] / qs[i] its marginals
(the reference binding's layout, pytorch/src/qr_drot_cuda_kernel.cu:89-93);
r_weight may be one float or one per problem.  The problems are packed into
the library's batch layout (leading dimension rounded up to 32 floats), solved by
drot_qr_batch_solve_f32, and unpacked; each plan is that of
reg_drot.quadratic_drot on that problem alone up to the grouping of the fp64
partial sums by the tile shape (same n_iter, plans within 1e-6).  PackedBatch exposes the packed
buffers for callers that keep problems resident between solves.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _abi


@dataclass
class PackedBatch:
    probs: object          # ctypes array of _abi.BatchProblem
    c: torch.Tensor        # packed costs (float32, device)
    p: torch.Tensor
    q: torch.Tensor
    x: torch.Tensor        # packed plans
    ws: torch.Tensor       # workspace
    shapes: list           # [(n_rows, n_cols)]

    def solve(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.c.device)
        return _abi.batch_solve(self.probs, self.c.data_ptr(), self.p.data_ptr(),
                                self.q.data_ptr(), self.x.data_ptr(), self.ws.data_ptr(),
                                self.ws.numel(), s.cuda_stream)

    def block(self, t: torch.Tensor, i: int) -> torch.Tensor:
        """The [n_cols, n_rows] view of problem i in a packed matrix buffer t."""
        pr = self.probs[i]
        n, m = pr.n_rows, pr.n_cols
        return t[pr.mat_offset: pr.mat_offset + pr.ld * m].view(m, pr.ld)[:, :n]

    def plan(self, i: int) -> torch.Tensor:
        return self.block(self.x, i)


def _check_marginals(ps, qs):
    for i, (p, q) in enumerate(zip(ps, qs)):
        if not (p.is_cuda and q.is_cuda):
            raise RuntimeError(f"problem {i}: tensors must be CUDA tensors")
        if p.dtype != torch.float32 or q.dtype != torch.float32:
            raise RuntimeError(f"problem {i}: tensors must be float32")


def layout(ps, qs, rho, r_weight, max_iters, eps) -> PackedBatch:
    """Packed buffers for problems with marginals ps[i], qs[i]; the cost buffer
    is zero-filled (write it, e.g. with reg_drot.cost.sqdist_packed)."""
    if len(ps) != len(qs) or len(ps) == 0:
        raise RuntimeError("ps and qs must be non-empty lists of the same length")
    _check_marginals(ps, qs)
    dev = ps[0].device
    n_rows = [p.numel() for p in ps]
    n_cols = [q.numel() for q in qs]
    probs, mat_elems, p_elems, q_elems = _abi.batch_layout(n_rows, n_cols)
    rws = list(r_weight) if isinstance(r_weight, (list, tuple)) else [r_weight] * len(ps)
    for i in range(len(ps)):
        probs[i].params = _abi.qr_params_from_rho(rho, float(rws[i]), n_rows[i], n_cols[i],
                                                  max_iters, eps)
    cp = torch.zeros(mat_elems, dtype=torch.float32, device=dev)
    pp = torch.empty(p_elems, dtype=torch.float32, device=dev)
    qp = torch.empty(q_elems, dtype=torch.float32, device=dev)
    for i, (p, q) in enumerate(zip(ps, qs)):
        pr = probs[i]
        pp[pr.p_offset: pr.p_offset + pr.n_rows].copy_(p)
        qp[pr.q_offset: pr.q_offset + pr.n_cols].copy_(q)
    x = torch.empty_like(cp)
    ws = torch.empty(_abi.batch_workspace_size(probs), dtype=torch.uint8, device=dev)
    return PackedBatch(probs, cp, pp, qp, x, ws, list(zip(n_rows, n_cols)))


def pack(cs, ps, qs, rho, r_weight, max_iters, eps) -> PackedBatch:
    if not (len(cs) == len(ps) == len(qs)) or len(cs) == 0:
        raise RuntimeError("cs, ps and qs must be non-empty lists of the same length")
    for i, (c, p, q) in enumerate(zip(cs, ps, qs)):
        if not (c.is_cuda and p.is_cuda and q.is_cuda):
            raise RuntimeError(f"problem {i}: tensors must be CUDA tensors")
        if c.dtype != torch.float32:
            raise RuntimeError(f"problem {i}: tensors must be float32")
        if c.dim() != 2 or c.shape[0] != q.numel() or c.shape[1] != p.numel():
            raise RuntimeError(f"problem {i}: C must be [q.numel(), p.numel()]")
    pb = layout(ps, qs, rho, r_weight, max_iters, eps)
    for i, c in enumerate(cs):
        pb.block(pb.c, i).copy_(c)
    return pb


def quadratic_drot_batch(cs, ps, qs, rho, r_weight, max_iters, eps):
    """Solve every problem; returns (plans, infos)."""
    pb = pack(cs, ps, qs, rho, r_weight, max_iters, eps)
    res = pb.solve()
    plans = [pb.plan(i).contiguous() for i in range(len(cs))]
    return plans, [r.as_dict() for r in res]
