"""ctypes front end of the CPU oracle (liboracle.so, built from drot_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py``'s ``cpu_baseline`` leg.  The product path
(``reg_drot`` / ``libdrot.so``) never imports it.

Arrays follow the reference's column-major layout (``core/kernel_qr.hpp:3``):
a plan or cost matrix is a numpy array of shape ``(n_cols, n_rows)`` in C order,
i.e. the PyTorch ``[M, N]`` tensor of ``pytorch/src/qr_drot_cuda_kernel.cu:89-93``,
with ``n_rows = len(p)`` contiguous.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    i, f, d = ctypes.c_int, ctypes.c_float, ctypes.c_double
    for sfx, rp, rt in (("_f32", _f32p, f), ("_f64", _f64p, d)):
        fn = getattr(lib, "oracle_warm_start" + sfx)
        fn.argtypes = [i, i, rt, rp, rp, rp, rp, rp, ctypes.POINTER(rt)]
        fn.restype = None
        fn = getattr(lib, "oracle_cold_start" + sfx)
        fn.argtypes = [i, i, rp, rp, rp, rp, rp, rp, rp, ctypes.POINTER(rt)]
        fn.restype = None
        fn = getattr(lib, "oracle_qr_step" + sfx)
        fn.argtypes = [i, i, i, rp, rp, rp, rp, rp, rp, rp, rp,
                       ctypes.POINTER(rt), rt, rt, _f64p, i]
        fn.restype = i
        fn = getattr(lib, "oracle_fix_x" + sfx)
        fn.argtypes = [i, i, rp, rp, rt]
        fn.restype = None
        fn = getattr(lib, "oracle_gl_step" + sfx)
        fn.argtypes = [i, i, i, _i32p, rp, rp, rp, rp, rp, rp, rp, rp,
                       ctypes.POINTER(rt), rt, rt, _f64p, i]
        fn.restype = i
        fn = getattr(lib, "oracle_qr_solve" + sfx)
        fn.argtypes = [i, i, rp, rp, rp, rt, rt, i, rt, i, rp, rp, rp, rp, rp,
                       ctypes.POINTER(rt), ctypes.POINTER(i), _f64p, i]
        fn.restype = i
        fn = getattr(lib, "oracle_gl_solve" + sfx)
        fn.argtypes = [i, i, i, _i32p, rp, rp, rp, rt, rt, i, rt, i, rp, rp, rp,
                       rp, rp, ctypes.POINTER(rt), ctypes.POINTER(i), _f64p, i]
        fn.restype = i
    for sfx, rp, rt in (("_f32", _f32p, f), ("_f64", _f64p, d)):
        fn = getattr(lib, "npdrot_qr_solve" + sfx)
        fn.argtypes = [i, i, rp, rp, rp, rt, rt, i, d, rp, ctypes.POINTER(d), i]
        fn.restype = i
    lib.oracle_qr_params_torch.argtypes = [f, f, i, i, ctypes.POINTER(f),
                                           ctypes.POINTER(f)]
    lib.oracle_qr_params_torch.restype = None
    lib.oracle_qr_kappa_core.argtypes = [f, f, i, i]
    lib.oracle_qr_kappa_core.restype = f
    lib.oracle_gl_lambda_torch.argtypes = [f, i, i]
    lib.oracle_gl_lambda_torch.restype = f
    lib.oracle_gl_lambda_core.argtypes = [f, i, i]
    lib.oracle_gl_lambda_core.restype = f
    lib.oracle_step_torch.argtypes = [f, i, i]
    lib.oracle_step_torch.restype = f
    lib.oracle_group_bounds.argtypes = [i, i, _i32p]
    lib.oracle_group_bounds.restype = None
    _lib = lib
    return lib


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------- parameters

def qr_params_torch(rho: float, r_weight: float, n_rows: int, n_cols: int):
    """(step, kappa) as computed by pytorch/src/qr_drot_cuda_kernel.cu:38-40."""
    lib = _load()
    s, k = ctypes.c_float(), ctypes.c_float()
    lib.oracle_qr_params_torch(rho, r_weight, n_rows, n_cols, ctypes.byref(s),
                               ctypes.byref(k))
    return np.float32(s.value), np.float32(k.value)


def qr_kappa_core(step: float, r_weight: float, n_rows: int, n_cols: int):
    """kappa as computed by core/drot_qr.hpp:262."""
    return np.float32(_load().oracle_qr_kappa_core(step, r_weight, n_rows, n_cols))


def step_torch(rho: float, n_rows: int, n_cols: int):
    return np.float32(_load().oracle_step_torch(rho, n_rows, n_cols))


def gl_lambda_torch(r_weight: float, n_rows: int, n_groups: int):
    """lambda = r_weight*sqrt(n_rows//G), pytorch/src/glr_drot_cuda_kernel.cu:47."""
    return np.float32(_load().oracle_gl_lambda_torch(r_weight, n_rows, n_groups))


def gl_lambda_core(r_weight: float, n_rows: int, n_groups: int):
    """lambda = r_weight/sqrt(n_rows//G), core/drot_glr.hpp:385."""
    return np.float32(_load().oracle_gl_lambda_core(r_weight, n_rows, n_groups))


def group_bounds(n_rows: int, n_groups: int) -> np.ndarray:
    """Row bounds of the groups, core/kernel_glr.hpp:79-80 / :185-186."""
    b = np.zeros(n_groups + 1, dtype=np.int32)
    _load().oracle_group_bounds(n_rows, n_groups, b)
    return b


# ---------------------------------------------------------------- state

@dataclass
class State:
    """Full DR state (X, phi1, phi2, a, b, alpha) in the oracle's precision."""
    x: np.ndarray      # (n_cols, n_rows), C order == column-major n_rows x n_cols
    phi1: np.ndarray   # (n_rows,)
    phi2: np.ndarray   # (n_cols,)
    a: np.ndarray
    b: np.ndarray
    alpha: float

    @property
    def dtype(self):
        return self.x.dtype

    def copy(self) -> "State":
        return State(self.x.copy(), self.phi1.copy(), self.phi2.copy(),
                     self.a.copy(), self.b.copy(), self.alpha)


def _sfx(dtype) -> str:
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def _ct(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def new_state(n_rows: int, n_cols: int, dtype=np.float32) -> State:
    z = lambda n: np.zeros(n, dtype=dtype)  # noqa: E731
    return State(np.zeros((n_cols, n_rows), dtype=dtype), z(n_rows), z(n_cols),
                 z(n_rows), z(n_cols), 0.0)


def warm_start(n_rows, n_cols, step, dtype=np.float32) -> State:
    """Closed-form warm start (pytorch/src/qr_drot_cuda_kernel.cu:45-71)."""
    st = new_state(n_rows, n_cols, dtype)
    al = _ct(dtype)()
    getattr(_load(), "oracle_warm_start" + _sfx(dtype))(
        n_rows, n_cols, step, st.x, st.phi1, st.phi2, st.a, st.b, ctypes.byref(al))
    st.alpha = al.value
    return st


def cold_start(p, q, dtype=np.float32) -> State:
    """X = p q^T start (core/drot_qr.hpp:308-320)."""
    n_rows, n_cols = len(p), len(q)
    st = new_state(n_rows, n_cols, dtype)
    al = _ct(dtype)()
    getattr(_load(), "oracle_cold_start" + _sfx(dtype))(
        n_rows, n_cols, np.ascontiguousarray(p, dtype), np.ascontiguousarray(q, dtype),
        st.x, st.phi1, st.phi2, st.a, st.b, ctypes.byref(al))
    st.alpha = al.value
    return st


def qr_step(st: State, k: int, c, p, q, step, kappa, nthreads=None):
    """One QR iteration in place; returns aux = [obj, sumX, sum rs^2, sum cs^2]."""
    dt = st.dtype
    aux = np.zeros(4, dtype=np.float64)
    al = _ct(dt)(st.alpha)
    n_cols, n_rows = st.x.shape
    rc = getattr(_load(), "oracle_qr_step" + _sfx(dt))(
        k, n_rows, n_cols, st.x, np.ascontiguousarray(c, dt),
        np.ascontiguousarray(p, dt), np.ascontiguousarray(q, dt),
        st.phi1, st.phi2, st.a, st.b, ctypes.byref(al), step, kappa, aux,
        nthreads or default_threads())
    if rc != 0:
        raise MemoryError("oracle_qr_step failed")
    st.alpha = al.value
    return aux


def gl_step(st: State, bounds, c, p, q, step, lam, nthreads=None):
    dt = st.dtype
    aux = np.zeros(4, dtype=np.float64)
    al = _ct(dt)(st.alpha)
    n_cols, n_rows = st.x.shape
    bounds = np.ascontiguousarray(bounds, np.int32)
    rc = getattr(_load(), "oracle_gl_step" + _sfx(dt))(
        n_rows, n_cols, len(bounds) - 1, bounds, st.x, np.ascontiguousarray(c, dt),
        np.ascontiguousarray(p, dt), np.ascontiguousarray(q, dt),
        st.phi1, st.phi2, st.a, st.b, ctypes.byref(al), step, lam, aux,
        nthreads or default_threads())
    if rc != 0:
        raise MemoryError("oracle_gl_step failed")
    st.alpha = al.value
    return aux


def fix_x(st: State, c, step):
    n_cols, n_rows = st.x.shape
    getattr(_load(), "oracle_fix_x" + _sfx(st.dtype))(
        n_rows, n_cols, st.x, np.ascontiguousarray(c, st.dtype), step)


@dataclass
class SolveResult:
    x: np.ndarray
    n_iter: int
    residual: float
    objective: float
    converged: bool
    state: State


def qr_solve(c, p, q, step, kappa, max_iters, eps, warm=True, dtype=np.float32,
             nthreads=None) -> SolveResult:
    """Full QR-DROT solve (core/drot_qr.hpp:159-234)."""
    c = np.ascontiguousarray(c, dtype)
    n_cols, n_rows = c.shape
    st = new_state(n_rows, n_cols, dtype)
    al = _ct(dtype)()
    nit = ctypes.c_int()
    out = np.zeros(3, dtype=np.float64)
    rc = getattr(_load(), "oracle_qr_solve" + _sfx(dtype))(
        n_rows, n_cols, c, np.ascontiguousarray(p, dtype), np.ascontiguousarray(q, dtype),
        step, kappa, int(max_iters), eps, int(bool(warm)), st.x, st.phi1, st.phi2,
        st.a, st.b, ctypes.byref(al), ctypes.byref(nit), out,
        nthreads or default_threads())
    if rc != 0:
        raise MemoryError("oracle_qr_solve failed")
    st.alpha = al.value
    return SolveResult(st.x, nit.value, out[0], out[1], bool(out[2]), st)


def gl_solve(c, p, q, n_groups, step, lam, max_iters, eps, warm=True,
             dtype=np.float32, nthreads=None, bounds=None) -> SolveResult:
    """Full GL-DROT solve (core/drot_glr.hpp:242-355)."""
    c = np.ascontiguousarray(c, dtype)
    n_cols, n_rows = c.shape
    if bounds is None:
        bounds = group_bounds(n_rows, n_groups)
    bounds = np.ascontiguousarray(bounds, np.int32)
    st = new_state(n_rows, n_cols, dtype)
    al = _ct(dtype)()
    nit = ctypes.c_int()
    out = np.zeros(3, dtype=np.float64)
    rc = getattr(_load(), "oracle_gl_solve" + _sfx(dtype))(
        n_rows, n_cols, len(bounds) - 1, bounds, c, np.ascontiguousarray(p, dtype),
        np.ascontiguousarray(q, dtype), step, lam, int(max_iters), eps, int(bool(warm)),
        st.x, st.phi1, st.phi2, st.a, st.b, ctypes.byref(al), ctypes.byref(nit),
        out, nthreads or default_threads())
    if rc != 0:
        raise MemoryError("oracle_gl_solve failed")
    st.alpha = al.value
    return SolveResult(st.x, nit.value, out[0], out[1], bool(out[2]), st)


# ---------------------------------------------------------------- npdrot, OpenMP

def npdrot_solve(y0, c, p, q, step, reg, max_iters, eps, dtype=np.float64, nthreads=None):
    """The reference's CPU solver (npdrot regdrot + quadratic prox) restated in
    C with OpenMP, pass for pass (npdrot_omp.c).  y0, c: (n_cols, n_rows)
    arrays in the solver layout (= npdrot's (m, n) F-order arrays); p: (n_rows,),
    q: (n_cols,).  Returns (X, iterations, last primal residual): X is the
    final prox output, as npdrot returns it."""
    c = np.ascontiguousarray(c, dtype)
    n_cols, n_rows = c.shape
    x = np.array(y0, dtype=dtype, order="C", copy=True)
    r = ctypes.c_double()
    k = getattr(_load(), "npdrot_qr_solve" + _sfx(dtype))(
        n_rows, n_cols, c, np.ascontiguousarray(p, dtype), np.ascontiguousarray(q, dtype),
        step, reg, int(max_iters), float(eps), x, ctypes.byref(r),
        nthreads or default_threads())
    if k < 0:
        raise MemoryError("npdrot_qr_solve failed")
    return x, k, r.value
