"""ctypes view of the C ABI in include/drot.h (libdrot.so).

This is the boundary a non-PyTorch host (or a test) binds; the PyTorch module
``reg_drot`` uses the same functions from C++.  Device buffers are passed as
integer addresses (e.g. ``tensor.data_ptr()``) and the stream as a hipStream_t
address (``torch.cuda.current_stream().cuda_stream``).
"""
from __future__ import annotations

import ctypes
import os

try:  # load torch's HIP runtime first so libdrot.so binds to the same one
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libdrot.so")

# Every function include/drot.h declares.
EXPORTS = (
    "drot_qr_params_from_rho", "drot_qr_params_from_step",
    "drot_glr_params_from_rho", "drot_glr_params_from_step",
    "drot_qr_workspace_size", "drot_glr_workspace_size",
    "drot_qr_solve_f32", "drot_glr_solve_f32", "drot_export_duals_f32",
    "drot_qr_wrapper_f32", "drot_glr_wrapper_f32", "drot_free",
    "drot_status_string", "drot_last_error", "drot_abi_version", "drot_set_timing",
    "drot_batch_layout", "drot_qr_batch_workspace_size", "drot_qr_batch_solve_f32",
    "drot_qr_backward_workspace_size", "drot_qr_backward_f32",
    "drot_qr_batch_backward_workspace_size", "drot_qr_batch_backward_f32",
    "drot_sqdist_workspace_size", "drot_sqdist_f32", "drot_sqdist_backward_f32",
    "drot_qr_tile_for", "drot_qr_autotune_workspace_size", "drot_qr_autotune_f32",
    "drot_qr_tile_cache_set", "drot_qr_tile_cache_save", "drot_qr_tile_cache_load",
    "drot_qr_tile_cache_clear", "drot_qr_solve_noalloc_f32",
)

# DROT_NOALLOC_SCRATCH_BYTES of include/drot.h: the device scratch of the
# allocation-free solve
NOALLOC_SCRATCH_BYTES = 32 << 20

DROT_OK = 0
ABI_VERSION = 2  # DROT_ABI_VERSION of the include/drot.h these prototypes follow


class QrParams(ctypes.Structure):
    _fields_ = [("step_size", ctypes.c_float), ("scale", ctypes.c_float),
                ("r_weight", ctypes.c_float), ("max_iters", ctypes.c_int),
                ("eps", ctypes.c_float), ("warm_start", ctypes.c_int)]


class GlrParams(ctypes.Structure):
    _fields_ = [("step_size", ctypes.c_float), ("lambda_", ctypes.c_float),
                ("n_groups", ctypes.c_int), ("max_iters", ctypes.c_int),
                ("eps", ctypes.c_float), ("warm_start", ctypes.c_int)]


class BatchProblem(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int), ("n_cols", ctypes.c_int), ("ld", ctypes.c_int),
                ("mat_offset", ctypes.c_longlong), ("p_offset", ctypes.c_longlong),
                ("q_offset", ctypes.c_longlong), ("params", QrParams)]


class Result(ctypes.Structure):
    _fields_ = [("n_iter", ctypes.c_int), ("converged", ctypes.c_int),
                ("iterations_run", ctypes.c_int), ("objective", ctypes.c_float),
                ("objective_f64", ctypes.c_double), ("residual", ctypes.c_double),
                ("run_ms", ctypes.c_float), ("prep_ms", ctypes.c_float),
                ("pass_ms", ctypes.c_double * 2), ("pass_count", ctypes.c_int * 2),
                ("pass_ms_events", ctypes.c_double * 2),
                ("pass_count_events", ctypes.c_int * 2)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["pass_ms"] = list(self.pass_ms)
        d["pass_count"] = list(self.pass_count)
        d["pass_ms_events"] = list(self.pass_ms_events)
        d["pass_count_events"] = list(self.pass_count_events)
        return d


class BackwardParams(ctypes.Structure):
    _fields_ = [("support_tol", ctypes.c_float), ("max_cg_iters", ctypes.c_int),
                ("cg_tol", ctypes.c_double), ("nnz_capacity", ctypes.c_longlong)]


class BackwardInfo(ctypes.Structure):
    _fields_ = [("nnz", ctypes.c_longlong), ("cg_iters", ctypes.c_int),
                ("converged", ctypes.c_int), ("cg_residual", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CostProblem(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int), ("n_cols", ctypes.c_int), ("ld", ctypes.c_int),
                ("mat_offset", ctypes.c_longlong), ("src_offset", ctypes.c_longlong),
                ("tgt_offset", ctypes.c_longlong)]


class QrTile(ctypes.Structure):
    _fields_ = [("vec", ctypes.c_int), ("lanes_per_col", ctypes.c_int),
                ("col_steps", ctypes.c_int), ("store_policy", ctypes.c_int),
                ("unroll", ctypes.c_int)]

    def as_tuple(self):
        return (self.vec, self.lanes_per_col, self.col_steps, self.store_policy, self.unroll)


class DrotError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load libdrot.so and declare the prototypes.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DrotError(f"{path} is not built; run regularized-drot_amd/build.py")
    lib = ctypes.CDLL(path)
    c_int, c_float, c_size, vp = ctypes.c_int, ctypes.c_float, ctypes.c_size_t, ctypes.c_void_p
    P = ctypes.POINTER
    lib.drot_qr_params_from_rho.argtypes = [c_float, c_float, c_int, c_int, c_int, c_float,
                                            P(QrParams)]
    lib.drot_qr_params_from_step.argtypes = [c_float, c_float, c_int, c_int, c_int, c_float,
                                             c_int, P(QrParams)]
    lib.drot_glr_params_from_rho.argtypes = [c_float, c_float, c_int, c_int, c_int, c_int,
                                             c_float, P(GlrParams)]
    lib.drot_glr_params_from_step.argtypes = [c_float, c_float, c_int, c_int, c_int, c_int,
                                              c_float, c_int, P(GlrParams)]
    lib.drot_qr_workspace_size.argtypes = [c_int, c_int, P(c_size)]
    lib.drot_glr_workspace_size.argtypes = [c_int, c_int, c_int, P(c_size)]
    lib.drot_qr_solve_f32.argtypes = [vp, vp, vp, c_int, c_int, P(QrParams), vp, vp, c_size,
                                      vp, P(Result)]
    lib.drot_glr_solve_f32.argtypes = [vp, vp, vp, c_int, c_int, P(GlrParams), vp, vp, c_size,
                                       vp, P(Result)]
    lib.drot_qr_solve_noalloc_f32.argtypes = [vp, vp, vp, c_int, c_int, P(QrParams), vp, vp,
                                              P(Result)]
    lib.drot_export_duals_f32.argtypes = [c_int, c_int, c_int, c_int, vp, c_size, vp, vp,
                                          P(c_float), vp]
    fp = P(c_float)
    lib.drot_qr_wrapper_f32.argtypes = [fp, fp, fp, c_int, c_int, c_float, c_float, c_int,
                                        c_float, fp, fp, P(c_int), fp, c_int, c_int]
    lib.drot_qr_wrapper_f32.restype = fp
    lib.drot_glr_wrapper_f32.argtypes = [fp, fp, fp, c_int, c_int, c_int, c_float, c_float,
                                         c_int, c_float, fp, fp, P(c_int), fp, c_int, c_int]
    lib.drot_glr_wrapper_f32.restype = fp
    lib.drot_free.argtypes = [vp]
    lib.drot_free.restype = None
    lib.drot_status_string.argtypes = [c_int]
    lib.drot_status_string.restype = ctypes.c_char_p
    lib.drot_last_error.argtypes = []
    lib.drot_last_error.restype = ctypes.c_char_p
    lib.drot_abi_version.argtypes = []
    lib.drot_abi_version.restype = c_int
    if lib.drot_abi_version() != ABI_VERSION:
        raise DrotError(f"{path} implements ABI {lib.drot_abi_version()}, expected {ABI_VERSION}; "
                        "rebuild it")
    lib.drot_set_timing.argtypes = [c_int]
    lib.drot_set_timing.restype = None
    ll = ctypes.c_longlong
    lib.drot_batch_layout.argtypes = [c_int, P(c_int), P(c_int), P(BatchProblem), P(ll), P(ll),
                                      P(ll)]
    lib.drot_qr_batch_workspace_size.argtypes = [c_int, P(BatchProblem), P(c_size)]
    lib.drot_qr_batch_solve_f32.argtypes = [c_int, P(BatchProblem), vp, vp, vp, vp, vp, c_size,
                                            vp, P(Result)]
    BP, BI = P(BackwardParams), P(BackwardInfo)
    lib.drot_qr_backward_workspace_size.argtypes = [c_int, c_int, P(QrParams), BP, P(c_size)]
    lib.drot_qr_backward_f32.argtypes = [vp, vp, vp, vp, c_int, c_int, P(QrParams), BP, vp, vp,
                                         c_size, vp, BI]
    lib.drot_qr_batch_backward_workspace_size.argtypes = [c_int, P(BatchProblem), BP, P(c_size)]
    lib.drot_qr_batch_backward_f32.argtypes = [c_int, P(BatchProblem), vp, vp, vp, vp, BP, vp,
                                               vp, c_size, vp, BI]
    CP = P(CostProblem)
    lib.drot_sqdist_workspace_size.argtypes = [c_int, CP, c_int, P(c_size)]
    lib.drot_sqdist_f32.argtypes = [c_int, CP, c_int, vp, vp, c_int, vp, vp, vp, vp, c_size, vp]
    lib.drot_sqdist_backward_f32.argtypes = [c_int, CP, c_int, vp, vp, vp, c_int, vp, vp, vp, vp,
                                             vp, c_size, vp]
    TP = P(QrTile)
    lib.drot_qr_tile_for.argtypes = [c_int, c_int, c_int, TP]
    lib.drot_qr_autotune_workspace_size.argtypes = [c_int, c_int, P(c_size)]
    lib.drot_qr_autotune_f32.argtypes = [vp, vp, vp, c_int, c_int, P(QrParams), vp, vp, c_size,
                                         vp, c_int, TP, P(ctypes.c_double)]
    lib.drot_qr_tile_cache_set.argtypes = [c_int, c_int, TP]
    lib.drot_qr_tile_cache_save.argtypes = [ctypes.c_char_p]
    lib.drot_qr_tile_cache_load.argtypes = [ctypes.c_char_p]
    lib.drot_qr_tile_cache_clear.argtypes = []
    lib.drot_qr_tile_cache_clear.restype = None
    for name in ("drot_qr_tile_for", "drot_qr_autotune_workspace_size", "drot_qr_autotune_f32",
                 "drot_qr_tile_cache_set", "drot_qr_tile_cache_save", "drot_qr_tile_cache_load",
                 "drot_sqdist_workspace_size", "drot_sqdist_f32", "drot_sqdist_backward_f32",
                 "drot_qr_backward_workspace_size", "drot_qr_backward_f32",
                 "drot_qr_batch_backward_workspace_size", "drot_qr_batch_backward_f32",
                 "drot_qr_params_from_rho", "drot_qr_params_from_step",
                 "drot_glr_params_from_rho", "drot_glr_params_from_step",
                 "drot_qr_workspace_size", "drot_glr_workspace_size",
                 "drot_qr_solve_f32", "drot_glr_solve_f32", "drot_export_duals_f32",
                 "drot_qr_solve_noalloc_f32",
                 "drot_batch_layout", "drot_qr_batch_workspace_size", "drot_qr_batch_solve_f32"):
        getattr(lib, name).restype = c_int
    _lib = lib
    return lib


def check(status: int, what: str = "drot call"):
    if status != DROT_OK:
        lib = load()
        raise DrotError(f"{what}: {lib.drot_status_string(status).decode()}: "
                        f"{lib.drot_last_error().decode()}")


def qr_params_from_rho(rho, r_weight, n_rows, n_cols, max_iters, eps) -> QrParams:
    prm = QrParams()
    check(load().drot_qr_params_from_rho(rho, r_weight, n_rows, n_cols, max_iters, eps,
                                         ctypes.byref(prm)), "drot_qr_params_from_rho")
    return prm


def qr_params_from_step(step, r_weight, n_rows, n_cols, max_iters, eps, warm=True) -> QrParams:
    prm = QrParams()
    check(load().drot_qr_params_from_step(step, r_weight, n_rows, n_cols, max_iters, eps,
                                          int(bool(warm)), ctypes.byref(prm)),
          "drot_qr_params_from_step")
    return prm


def glr_params_from_rho(rho, r_weight, n_rows, n_cols, n_groups, max_iters, eps) -> GlrParams:
    prm = GlrParams()
    check(load().drot_glr_params_from_rho(rho, r_weight, n_rows, n_cols, n_groups, max_iters,
                                          eps, ctypes.byref(prm)), "drot_glr_params_from_rho")
    return prm


def glr_params_from_step(step, r_weight, n_rows, n_cols, n_groups, max_iters, eps,
                         warm=True) -> GlrParams:
    prm = GlrParams()
    check(load().drot_glr_params_from_step(step, r_weight, n_rows, n_cols, n_groups,
                                           max_iters, eps, int(bool(warm)), ctypes.byref(prm)),
          "drot_glr_params_from_step")
    return prm


def qr_workspace_size(n_rows, n_cols) -> int:
    b = ctypes.c_size_t()
    check(load().drot_qr_workspace_size(n_rows, n_cols, ctypes.byref(b)), "workspace size")
    return b.value


def glr_workspace_size(n_rows, n_cols, n_groups) -> int:
    b = ctypes.c_size_t()
    check(load().drot_glr_workspace_size(n_rows, n_cols, n_groups, ctypes.byref(b)),
          "workspace size")
    return b.value


def qr_solve(c, p, q, n_rows, n_cols, prm: QrParams, x, ws, ws_bytes, stream) -> Result:
    """All buffer arguments are device addresses (ints)."""
    r = Result()
    check(load().drot_qr_solve_f32(c, p, q, n_rows, n_cols, ctypes.byref(prm), x, ws,
                                   ws_bytes, stream, ctypes.byref(r)), "drot_qr_solve_f32")
    return r


def qr_solve_noalloc(c, p, q, n_rows, n_cols, prm: QrParams, x, stream) -> Result:
    """drot_qr_solve_noalloc_f32: no workspace argument, nothing allocated."""
    r = Result()
    check(load().drot_qr_solve_noalloc_f32(c, p, q, n_rows, n_cols, ctypes.byref(prm), x,
                                           stream, ctypes.byref(r)),
          "drot_qr_solve_noalloc_f32")
    return r


def glr_solve(c, p, q, n_rows, n_cols, prm: GlrParams, x, ws, ws_bytes, stream) -> Result:
    r = Result()
    check(load().drot_glr_solve_f32(c, p, q, n_rows, n_cols, ctypes.byref(prm), x, ws,
                                    ws_bytes, stream, ctypes.byref(r)), "drot_glr_solve_f32")
    return r


def set_timing(enable: bool) -> None:
    load().drot_set_timing(int(bool(enable)))


def export_duals(kind, n_rows, n_cols, n_groups, ws, ws_bytes, phi1, phi2, stream):
    alpha = ctypes.c_float()
    check(load().drot_export_duals_f32(kind, n_rows, n_cols, n_groups, ws, ws_bytes, phi1,
                                       phi2, ctypes.byref(alpha), stream),
          "drot_export_duals_f32")
    return alpha.value


def batch_layout(n_rows, n_cols):
    """Packed layout of a batch: (problems array, mat_elems, p_elems, q_elems)."""
    n = len(n_rows)
    probs = (BatchProblem * n)()
    nr = (ctypes.c_int * n)(*[int(v) for v in n_rows])
    nc = (ctypes.c_int * n)(*[int(v) for v in n_cols])
    m, pe, qe = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
    check(load().drot_batch_layout(n, nr, nc, probs, ctypes.byref(m), ctypes.byref(pe),
                                   ctypes.byref(qe)), "drot_batch_layout")
    return probs, m.value, pe.value, qe.value


def batch_workspace_size(probs) -> int:
    b = ctypes.c_size_t()
    check(load().drot_qr_batch_workspace_size(len(probs), probs, ctypes.byref(b)),
          "drot_qr_batch_workspace_size")
    return b.value


def batch_solve(probs, c, p, q, x, ws, ws_bytes, stream):
    """Device addresses for c/p/q/x/ws; returns a list of Result."""
    n = len(probs)
    res = (Result * n)()
    check(load().drot_qr_batch_solve_f32(n, probs, c, p, q, x, ws, ws_bytes, stream, res),
          "drot_qr_batch_solve_f32")
    return list(res)


def backward_params(support_tol=0.0, max_cg_iters=0, cg_tol=0.0, nnz_capacity=0) -> BackwardParams:
    return BackwardParams(float(support_tol), int(max_cg_iters), float(cg_tol), int(nnz_capacity))


def qr_backward_workspace_size(n_rows, n_cols, prm: QrParams, bp: BackwardParams = None) -> int:
    b = ctypes.c_size_t()
    check(load().drot_qr_backward_workspace_size(n_rows, n_cols, ctypes.byref(prm),
                                                 ctypes.byref(bp) if bp else None,
                                                 ctypes.byref(b)),
          "drot_qr_backward_workspace_size")
    return b.value


def qr_backward(c, x, grad_x, grad_obj, n_rows, n_cols, prm: QrParams, bp, grad_c, ws, ws_bytes,
                stream) -> BackwardInfo:
    """Device addresses for c/x/grad_x/grad_obj (0 = none)/grad_c/ws."""
    info = BackwardInfo()
    check(load().drot_qr_backward_f32(c, x, grad_x, grad_obj or None, n_rows, n_cols,
                                      ctypes.byref(prm), ctypes.byref(bp) if bp else None,
                                      grad_c, ws, ws_bytes, stream, ctypes.byref(info)),
          "drot_qr_backward_f32")
    return info


def batch_backward_workspace_size(probs, bp: BackwardParams = None) -> int:
    b = ctypes.c_size_t()
    check(load().drot_qr_batch_backward_workspace_size(len(probs), probs,
                                                       ctypes.byref(bp) if bp else None,
                                                       ctypes.byref(b)),
          "drot_qr_batch_backward_workspace_size")
    return b.value


def batch_backward(probs, c, x, grad_x, grad_obj, bp, grad_c, ws, ws_bytes, stream):
    """Device addresses; grad_obj: device float array of len(probs) or 0.  Returns
    a list of BackwardInfo."""
    n = len(probs)
    infos = (BackwardInfo * n)()
    check(load().drot_qr_batch_backward_f32(n, probs, c, x, grad_x, grad_obj or None,
                                            ctypes.byref(bp) if bp else None, grad_c, ws,
                                            ws_bytes, stream, infos),
          "drot_qr_batch_backward_f32")
    return list(infos)


def cost_problems(shapes, lds=None, mat_offsets=None):
    """ctypes array of CostProblem for [(n_rows, n_cols)]: points of problem i
    follow those of problem i-1 in the packed src / tgt buffers; C blocks are
    dense (ld = n_rows) unless lds / mat_offsets are given."""
    n = len(shapes)
    probs = (CostProblem * n)()
    so = to = mo = 0
    for i, (nr, nc) in enumerate(shapes):
        ld = int(lds[i]) if lds is not None else int(nr)
        off = int(mat_offsets[i]) if mat_offsets is not None else mo
        probs[i] = CostProblem(int(nr), int(nc), ld, off, so, to)
        so += nr
        to += nc
        mo = off + ld * nc
    return probs


def sqdist_workspace_size(probs, dim) -> int:
    b = ctypes.c_size_t()
    check(load().drot_sqdist_workspace_size(len(probs), probs, int(dim), ctypes.byref(b)),
          "drot_sqdist_workspace_size")
    return b.value


def sqdist(probs, dim, src, tgt, normalize, c, scale, argmax, ws, ws_bytes, stream):
    check(load().drot_sqdist_f32(len(probs), probs, int(dim), src, tgt, int(bool(normalize)), c,
                                 scale or None, argmax or None, ws, ws_bytes, stream),
          "drot_sqdist_f32")


def sqdist_backward(probs, dim, src, tgt, grad_c, normalize, scale, argmax, grad_src, grad_tgt,
                    ws, ws_bytes, stream):
    check(load().drot_sqdist_backward_f32(len(probs), probs, int(dim), src, tgt, grad_c,
                                          int(bool(normalize)), scale or None, argmax or None,
                                          grad_src, grad_tgt, ws, ws_bytes, stream),
          "drot_sqdist_backward_f32")


def qr_tile_for(n_rows, n_cols, vec4_ok=True) -> QrTile:
    t = QrTile()
    check(load().drot_qr_tile_for(n_rows, n_cols, int(bool(vec4_ok)), ctypes.byref(t)),
          "drot_qr_tile_for")
    return t


def qr_tile_cache_set(n_rows, n_cols, tile: QrTile):
    check(load().drot_qr_tile_cache_set(n_rows, n_cols, ctypes.byref(tile)),
          "drot_qr_tile_cache_set")


def qr_tile_cache_save(path):
    check(load().drot_qr_tile_cache_save(os.fsencode(path)), "drot_qr_tile_cache_save")


def qr_tile_cache_load(path):
    check(load().drot_qr_tile_cache_load(os.fsencode(path)), "drot_qr_tile_cache_load")


def qr_tile_cache_clear():
    load().drot_qr_tile_cache_clear()


def qr_autotune_workspace_size(n_rows, n_cols) -> int:
    b = ctypes.c_size_t()
    check(load().drot_qr_autotune_workspace_size(n_rows, n_cols, ctypes.byref(b)),
          "drot_qr_autotune_workspace_size")
    return b.value


def qr_autotune(c, p, q, n_rows, n_cols, prm: QrParams, x, ws, ws_bytes, stream, iters=0):
    """Device addresses; returns (QrTile, us_per_iter)."""
    t = QrTile()
    us = ctypes.c_double()
    check(load().drot_qr_autotune_f32(c, p, q, n_rows, n_cols, ctypes.byref(prm), x, ws, ws_bytes,
                                      stream, int(iters), ctypes.byref(t), ctypes.byref(us)),
          "drot_qr_autotune_f32")
    return t, us.value
