"""Cost matrices from point clouds on the GPU, differentiable in the points.

    c = reg_drot.cost.sqdist(x, y)                 # [len(y), len(x)], / max
    cp = reg_drot.cost.sqdist_packed(xs, ys, pb)   # straight into a PackedBatch layout

C[j, i] = ||x_i - y_j||^2 / max C (normalize=True), the cost the reference
builds on the host with POT before a solve (benchmark/generate_ot.py:22-23,
benchmark/generate_da.py:40-41).  x (source points, the p side) is [N, dim],
y (target points, the q side) is [M, dim], dim <= 64, float32 CUDA tensors.
The forward and the backward to x and y are the kernels of drot_sqdist_f32 /
drot_sqdist_backward_f32 (libdrot.so).
"""
from __future__ import annotations

import torch

from . import _abi


def _check_points(x, y):
    for name, t in (("x", x), ("y", y)):
        if not t.is_cuda or t.dtype != torch.float32 or t.dim() != 2:
            raise RuntimeError(f"{name} must be a 2-D float32 CUDA tensor")
    if x.shape[1] != y.shape[1]:
        raise RuntimeError("x and y must have the same point dimension")


class _SqdistFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src, tgt, probs, numel, normalize):
        src = src.contiguous()
        tgt = tgt.contiguous()
        dim = src.shape[1]
        dev = src.device
        n = len(probs)
        c = torch.empty(numel, dtype=torch.float32, device=dev)
        scale = torch.ones(n, dtype=torch.float32, device=dev)
        argmax = torch.zeros(2 * n, dtype=torch.int32, device=dev)
        nbytes = _abi.sqdist_workspace_size(probs, dim)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        _abi.sqdist(probs, dim, src.data_ptr(), tgt.data_ptr(), normalize, c.data_ptr(),
                    scale.data_ptr(), argmax.data_ptr(), ws.data_ptr(), nbytes, stream)
        ctx.save_for_backward(src, tgt, scale, argmax)
        ctx.probs = probs
        ctx.normalize = normalize
        ctx.mark_non_differentiable(scale)
        return c, scale

    @staticmethod
    def backward(ctx, grad_c, _grad_scale):
        src, tgt, scale, argmax = ctx.saved_tensors
        dim = src.shape[1]
        gs = torch.empty_like(src)
        gt = torch.empty_like(tgt)
        nbytes = _abi.sqdist_workspace_size(ctx.probs, dim)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=src.device)
        stream = torch.cuda.current_stream(src.device).cuda_stream
        g = grad_c.contiguous().to(torch.float32)
        _abi.sqdist_backward(ctx.probs, dim, src.data_ptr(), tgt.data_ptr(), g.data_ptr(),
                             ctx.normalize, scale.data_ptr(), argmax.data_ptr(), gs.data_ptr(),
                             gt.data_ptr(), ws.data_ptr(), nbytes, stream)
        return gs, gt, None, None, None


def sqdist(x, y, normalize=True):
    """[len(y), len(x)] squared distances (divided by their max when normalize)."""
    _check_points(x, y)
    n, m = x.shape[0], y.shape[0]
    probs = _abi.cost_problems([(n, m)])
    c, _ = _SqdistFunction.apply(x, y, probs, n * m, bool(normalize))
    return c.view(m, n)


def sqdist_packed(xs, ys, pb, normalize=True):
    """Costs of problems (xs[i], ys[i]) packed in the layout of a
    reg_drot.batch.PackedBatch (leading dimensions and offsets of pb.probs);
    returns a flat tensor of pb.c's size, differentiable in every xs[i], ys[i]."""
    if len(xs) != len(ys) or len(xs) != len(pb.probs):
        raise RuntimeError("xs, ys and the packed batch must hold the same number of problems")
    for x, y, pr in zip(xs, ys, pb.probs):
        _check_points(x, y)
        if x.shape[0] != pr.n_rows or y.shape[0] != pr.n_cols:
            raise RuntimeError("point counts must match the packed problems")
    probs = _abi.cost_problems([(pr.n_rows, pr.n_cols) for pr in pb.probs],
                               lds=[pr.ld for pr in pb.probs],
                               mat_offsets=[pr.mat_offset for pr in pb.probs])
    src = torch.cat(list(xs))
    tgt = torch.cat(list(ys))
    c, _ = _SqdistFunction.apply(src, tgt, probs, pb.c.numel(), bool(normalize))
    return c
