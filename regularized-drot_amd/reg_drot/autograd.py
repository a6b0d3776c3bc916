"""Differentiable QR-DROT: torch.autograd.Function wrappers of the HIP solver
and its implicit-differentiation backward.

    x = reg_drot.quadratic_drot_diff(c, p, q, rho, r_weight, max_iters, eps)
    x, F = reg_drot.quadratic_drot_diff(..., return_objective=True)
    plans, Fs = reg_drot.quadratic_drot_batch_diff(cs, ps, qs, rho, r_weight, max_iters, eps)

The forward is reg_drot.quadratic_drot (pytorch/src/drot_cuda.cpp:26-40 in the
reference, whose output carries no gradient: qr_drot_cuda_kernel.cu:99).  F is
the regularized objective <C,X> + (reg/2)||X||^2 with reg = (scale-1)/step_size
(fp64).  The backward returns dL/dC only (p, q, rho and r_weight get None):
    dL/dC = S o (w 1^T + 1 w'^T - dL/dX) / reg + (dL/dF) X
with S the plan's support and [w; w'] the solution of the bipartite-Laplacian
system of the KKT conditions on S (drot_qr_backward_f32).  The derivative is
that of the exact optimum on the support the forward found, so the forward
should be converged well below the accuracy the gradient needs.
"""
from __future__ import annotations

import torch

from . import _abi, _C, batch as _batch

def _objective(c, x, reg):
    return (torch.sum(c * x, dtype=torch.float64)
            + 0.5 * reg * torch.sum(x * x, dtype=torch.float64))


class QuadraticDROTFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, p, q, rho, r_weight, max_iters, eps, support_tol, max_cg_iters, cg_tol):
        x, info = _C.quadratic_drot_ex(c, p, q, rho, r_weight, max_iters, eps)
        reg = (float(info["scale"]) - 1.0) / float(info["step_size"])
        obj = _objective(c, x, reg)
        ctx.save_for_backward(c, p, q, x)
        ctx.args = (float(rho), float(r_weight), float(support_tol), int(max_cg_iters),
                    float(cg_tol))
        ctx.info = info
        return x, obj

    @staticmethod
    def backward(ctx, grad_x, grad_obj):
        c, p, q, x = ctx.saved_tensors
        if not ctx.needs_input_grad[0]:
            return (None,) * 10
        gx = torch.zeros_like(x) if grad_x is None else grad_x.contiguous().to(torch.float32)
        gc, info = _C.quadratic_drot_backward(c, p, q, x, gx, grad_obj, *ctx.args)
        ctx.backward_info = info
        return (gc,) + (None,) * 9


def quadratic_drot_diff(c, p, q, rho, r_weight, max_iters, eps, *, return_objective=False,
                        support_tol=0.0, max_cg_iters=0, cg_tol=0.0):
    """reg_drot.quadratic_drot with a gradient for c (see module docstring)."""
    x, obj = QuadraticDROTFunction.apply(c, p, q, rho, r_weight, max_iters, eps, support_tol,
                                         max_cg_iters, cg_tol)
    return (x, obj) if return_objective else x


# ------------------------------------------------------------------ batches

class _BatchFunction(torch.autograd.Function):
    """Packed costs -> (packed plans, objectives).  pb holds the packed layout;
    pb.c is overwritten with the packed costs of this call."""

    @staticmethod
    def forward(ctx, cp, pb, bp):
        pb.c.copy_(cp)
        res = pb.solve()
        xp = pb.x.clone()
        objs = []
        for i, pr in enumerate(pb.probs):
            reg = (float(pr.params.scale) - 1.0) / float(pr.params.step_size)
            objs.append(_objective(pb.block(pb.c, i), pb.block(xp, i), reg))
        ctx.save_for_backward(cp, xp)
        ctx.pb = pb
        ctx.bp = bp
        ctx.results = res
        return xp, torch.stack(objs)

    @staticmethod
    def backward(ctx, grad_xp, grad_objs):
        cp, xp = ctx.saved_tensors
        pb = ctx.pb
        gx = torch.zeros_like(xp) if grad_xp is None else grad_xp.contiguous().to(torch.float32)
        go = None if grad_objs is None else grad_objs.to(torch.float32).contiguous()
        grad_cp = torch.empty_like(cp)
        nbytes = _abi.batch_backward_workspace_size(pb.probs, ctx.bp)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=cp.device)
        stream = torch.cuda.current_stream(cp.device).cuda_stream
        ctx.backward_infos = _abi.batch_backward(
            pb.probs, cp.data_ptr(), xp.data_ptr(), gx.data_ptr(),
            go.data_ptr() if go is not None else 0, ctx.bp, grad_cp.data_ptr(), ws.data_ptr(),
            nbytes, stream)
        return grad_cp, None, None


class _Unpack(torch.autograd.Function):
    """Packed buffer -> the per-problem [n_cols, n_rows] views.  Slicing the
    packed tensor directly would give every view's gradient the full packed
    size (one fill and one add of the whole buffer per problem); here the
    backward fills the packed gradient once and copies each problem's block."""

    @staticmethod
    def forward(ctx, xp, pb):
        ctx.set_materialize_grads(False)
        ctx.pb = pb
        ctx.meta = (xp.shape, xp.dtype, xp.device)
        return tuple(pb.block(xp, i) for i in range(len(pb.shapes)))

    @staticmethod
    def backward(ctx, *grads):
        shape, dtype, device = ctx.meta
        g = torch.zeros(shape, dtype=dtype, device=device)
        for i, gi in enumerate(grads):
            if gi is not None:
                ctx.pb.block(g, i).copy_(gi)
        return g, None


def unpack(xp, pb):
    """Per-problem plans (views of the packed xp) whose gradients flow back to
    xp in one pass; use it instead of pb.block when a loss reads many plans."""
    return list(_Unpack.apply(xp, pb))


def pack_costs(cs, pb) -> torch.Tensor:
    """Differentiable packing of cs into pb's layout (padding rows are zero)."""
    parts = []
    off = 0
    for c, pr in zip(cs, pb.probs):
        if pr.mat_offset != off:
            raise RuntimeError("pack_costs needs the dense layout of drot_batch_layout")
        parts.append(torch.nn.functional.pad(c, (0, pr.ld - pr.n_rows)).reshape(-1))
        off += pr.ld * pr.n_cols
    return torch.cat(parts)


def batch_solve_diff(cp, pb, *, support_tol=0.0, max_cg_iters=0, cg_tol=0.0, nnz_capacity=0):
    """Differentiable solve of the packed costs cp (a flat tensor in pb's layout,
    e.g. from reg_drot.cost.sqdist_packed): returns (packed plans, objectives)."""
    bp = _abi.backward_params(support_tol, max_cg_iters, cg_tol, nnz_capacity)
    return _BatchFunction.apply(cp, pb, bp)


def quadratic_drot_batch_diff(cs, ps, qs, rho, r_weight, max_iters, eps, *, support_tol=0.0,
                              max_cg_iters=0, cg_tol=0.0, nnz_capacity=0, packed=None):
    """Differentiable batched solve: returns (plans, objectives[n] fp64).

    Gradients flow to every cs[i] through one batched backward launch sequence
    (drot_qr_batch_backward_f32).  `packed` may pass a PackedBatch from an
    earlier call with the same shapes and parameters, to reuse its buffers."""
    pb = packed if packed is not None else _batch.layout(ps, qs, rho, r_weight, max_iters, eps)
    cp = pack_costs(cs, pb)
    xp, objs = batch_solve_diff(cp, pb, support_tol=support_tol, max_cg_iters=max_cg_iters,
                                cg_tol=cg_tol, nnz_capacity=nnz_capacity)
    return unpack(xp, pb), objs
