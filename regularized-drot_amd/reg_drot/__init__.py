"""reg_drot -- MI355X-native drop-in for the reference's PyTorch extension.

Same module name and call signatures as the reference's binding
(pytorch/src/drot_cuda.cpp:59-62, documented in pytorch/README.md:32-72):

    x = reg_drot.quadratic_drot(c, p, q, rho, r_weight, max_iters, eps)
    x = reg_drot.group_lasso_drot(c, p, q, n_class, rho, r_weight, max_iters, eps)

c is a contiguous float32 GPU tensor of shape [M, N] = [len(q), len(p)]; p and
q are float32 GPU tensors.  Groups are row ranges of p
([round(i*N/n_class), round((i+1)*N/n_class))).  The returned plan has c's
shape.  Errors raise RuntimeError as the reference's TORCH_CHECKs do.

The solve runs in libdrot.so (hand-written gfx950 HIP kernels behind the C ABI
of include/drot.h) on the caller's current HIP stream.  There is no CPU
fallback: importing this module fails if the native extension is not built.

Extras that the reference computes but discards:
    x, info = reg_drot.quadratic_drot_ex(...)    # info: n_iter, objective, residual
    x, info = reg_drot.group_lasso_drot_ex(...)

Additions with no reference counterpart (the reference has no backward and no
batched path): reg_drot.quadratic_drot_diff / quadratic_drot_batch_diff
(autograd through the implicit-differentiation backward, reg_drot.autograd)
and reg_drot.batch (many problems per solve).
"""
from __future__ import annotations

import torch  # noqa: F401  (loads torch's HIP runtime before libdrot.so)

try:
    from . import _C
except ImportError as exc:  # pragma: no cover - exercised only on a broken install
    raise ImportError(
        "reg_drot native extension is not built; run "
        "`python regularized-drot_amd/build.py` (or __graft_entry__.build())"
    ) from exc

quadratic_drot = _C.quadratic_drot
group_lasso_drot = _C.group_lasso_drot
quadratic_drot_ex = _C.quadratic_drot_ex
group_lasso_drot_ex = _C.group_lasso_drot_ex

from .autograd import quadratic_drot_batch_diff, quadratic_drot_diff  # noqa: E402

__all__ = ["quadratic_drot", "group_lasso_drot", "quadratic_drot_ex", "group_lasso_drot_ex",
           "quadratic_drot_diff", "quadratic_drot_batch_diff"]
