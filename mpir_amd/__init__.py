"""MI355X-native solve path of yex33/MPIR: ordering, QDLDL LDL^T and GMRES-IR.

Host-side mirror of the reference's interfaces over the C ABI of
libmpir_hip.so (include/mpir_hip.h):

* ``QDLDL_etree``, ``QDLDL_factor``, ``QDLDL_solve`` -- same arguments and
  return codes as the templates in include/MPIR/qdldl.hpp (:48, :172, :417),
  with NumPy arrays in place of raw pointers; the index dtype (int32/int64)
  selects INT and the value dtype (float64/float32) selects PREC.
* ``GmresLDLIR`` -- ``Compute`` / ``Solve`` / ``SetMaxIRIterations`` /
  ``SetMaxGmresIterations`` / ``SetTolerance`` of GmresLDLIR<UF,UW,UR>
  (include/MPIR/gmres_ir.hpp:90-315), LDL^T form.

Everything computes on the GPU; nothing here falls back to the CPU, and the
import fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import GmresStats, last_error, load

F64, F32, F16, BF16 = 0, 1, 2, 3
ORDER_NATURAL, ORDER_AMD, ORDER_ND = 0, 1, 2

_PREC_OF = {np.dtype(np.float64): F64, np.dtype(np.float32): F32}
_FACTOR_PREC_OF = {**_PREC_OF, np.dtype(np.float16): F16}
SCALE_NONE, SCALE_DIAG, SCALE_RUIZ = 0, 1, 2
SCHED_LEVELS, SCHED_DAG = 0, 1


class MpirError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where} failed ({code}): {last_error()}")
        self.code = code


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int, where: str) -> int:
    if rc is not None and rc <= -100:
        raise MpirError(rc, where)
    return rc


def _req(a, dtype, name: str) -> np.ndarray:
    arr = np.asarray(a)
    if arr.dtype != np.dtype(dtype) or not arr.flags.c_contiguous:
        raise TypeError(f"{name} must be a contiguous {np.dtype(dtype).name} array")
    return arr


def device_ok() -> bool:
    return bool(load().mpir_device_ok())


# ---------------------------------------------------------------------- QDLDL
def QDLDL_etree(n, Ap, Ai, work, Lnz, etree) -> int:
    """QDLDL_etree<INT> (qdldl.hpp:48-105): fills work/Lnz/etree, returns sum(Lnz), -1 or -2."""
    it = np.asarray(Ap).dtype
    if it == np.int32:
        fn = load().mpir_qdldl_etree_i32
    elif it == np.int64:
        fn = load().mpir_qdldl_etree_i64
    else:
        raise TypeError("index arrays must be int32 or int64")
    arrs = [_req(x, it, nm) for x, nm in ((Ap, "Ap"), (Ai, "Ai"), (work, "work"), (Lnz, "Lnz"), (etree, "etree"))]
    return _check(int(fn(int(n), *map(_ptr, arrs))), "QDLDL_etree")


def QDLDL_factor(n, Ap, Ai, Ax, Lp, Li, Lx, D, Dinv, Lnz, etree, bwork=None, iwork=None, fwork=None) -> int:
    """QDLDL_factor<INT,PREC> (qdldl.hpp:172-345) on the GPU; returns #D>0 or -1 on a zero pivot."""
    it = np.asarray(Ap).dtype
    vt = np.asarray(Ax).dtype
    if vt not in _PREC_OF:
        raise TypeError("Ax must be float64 or float32")
    suffix = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}[vt]
    isuf = {np.dtype(np.int32): "i32", np.dtype(np.int64): "i64"}.get(it)
    if isuf is None:
        raise TypeError("index arrays must be int32 or int64")
    fn = getattr(load(), f"mpir_qdldl_factor_{suffix}_{isuf}")
    ints = [_req(x, it, nm) for x, nm in ((Ap, "Ap"), (Ai, "Ai"), (Lp, "Lp"), (Li, "Li"), (Lnz, "Lnz"), (etree, "etree"))]
    vals = [_req(x, vt, nm) for x, nm in ((Ax, "Ax"), (Lx, "Lx"), (D, "D"), (Dinv, "Dinv"))]
    Ap_, Ai_, Lp_, Li_, Lnz_, et_ = ints
    Ax_, Lx_, D_, Di_ = vals
    rc = fn(int(n), _ptr(Ap_), _ptr(Ai_), _ptr(Ax_), _ptr(Lp_), _ptr(Li_), _ptr(Lx_), _ptr(D_), _ptr(Di_),
            _ptr(Lnz_), _ptr(et_), None, None, None)
    return _check(int(rc), "QDLDL_factor")


def QDLDL_solve(n, Lp, Li, Lx, Dinv, x) -> None:
    """QDLDL_solve<INT,FACTPREC,XPREC> (qdldl.hpp:417-432) on the GPU, x overwritten."""
    it = np.asarray(Lp).dtype
    ft = np.asarray(Lx).dtype
    xt = np.asarray(x).dtype
    key = ({np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}.get(ft),
           {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}.get(xt),
           {np.dtype(np.int32): "i32", np.dtype(np.int64): "i64"}.get(it))
    name = f"mpir_qdldl_solve_{key[0]}_{key[1]}_{key[2]}"
    fn = getattr(load(), name, None)
    if None in key or fn is None:
        raise TypeError(f"unsupported precision/index combination {key}")
    Lp_, Li_ = _req(Lp, it, "Lp"), _req(Li, it, "Li")
    Lx_, Di_ = _req(Lx, ft, "Lx"), _req(Dinv, ft, "Dinv")
    x_ = _req(x, xt, "x")
    _check(int(fn(int(n), _ptr(Lp_), _ptr(Li_), _ptr(Lx_), _ptr(Di_), _ptr(x_))), "QDLDL_solve")


# ---------------------------------------------------------------------- LUIR
def LUIR_QDLDL(Ap, Ai, Ax, b, FACTPREC=np.float32, maxiter: int = 10):
    """LUIR_QDLDL<FACTPREC, double, double, double, INT> (include/qdldl-luir/
    luir.hpp:33-241) on the GPU: QDLDL factor of the matrix as given (upper
    CSC), classic iterative refinement with a double residual. Returns
    (x, stats) with stats the fields of the reference's Stats (stats.hpp)."""
    it = np.dtype(np.asarray(Ap).dtype)
    it = np.dtype(np.int64) if it == np.int64 else np.dtype(np.int32)
    ft = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}.get(np.dtype(FACTPREC))
    if ft is None:
        raise TypeError("FACTPREC must be float64 or float32")
    Ap_, Ai_ = _req(Ap, it, "Ap"), _req(Ai, it, "Ai")
    Ax_ = np.ascontiguousarray(Ax, dtype=np.float64)
    b_ = np.ascontiguousarray(b, dtype=np.float64)
    n = len(Ap_) - 1
    x = np.zeros(n)
    st = _lib.LuirStats()
    fn = getattr(load(), f"mpir_luir_qdldl_{ft}_{'i64' if it == np.int64 else 'i32'}")
    rc = int(fn(n, _ptr(Ap_), _ptr(Ai_), _ptr(Ax_), _ptr(b_), _ptr(x), C.byref(st), int(maxiter)))
    if rc < 0:
        raise MpirError(rc, "LUIR_QDLDL")
    return x, st.as_dict()


# ---------------------------------------------------------------------- GMRES-IR
class GmresLDLIR:
    """GmresLDLIR<UF,UW,UR> (gmres_ir.hpp:90-315), LDL^T preconditioner, on the GPU.

    UF: np.float64 or np.float32; UW and UR: np.float64. Inputs follow the
    reference: the upper triangle of K in CSC (Ap, Ai, Ax), rows ascending.
    """

    def __init__(self, UF=np.float64, UW=np.float64, UR=np.float64, ordering: int = ORDER_AMD,
                 scaling: int = SCALE_DIAG):
        # UF may also be np.float16 or "bf16": a factor computed with FP32
        # arithmetic and stored (and applied) in the 16-bit format
        uf = BF16 if isinstance(UF, str) and UF.lower() == "bf16" else _FACTOR_PREC_OF.get(np.dtype(UF))
        if uf is None:
            raise TypeError("UF must be float64, float32, float16 or 'bf16'")
        if np.dtype(UW) != np.float64 or np.dtype(UR) != np.float64:
            raise TypeError("UW and UR must be float64")
        h = C.c_void_p()
        _check(load().mpir_gmres_ldl_create(uf, F64, F64, C.byref(h)), "GmresLDLIR()")
        self._h = h
        self._n = 0
        self.UF = UF if uf == BF16 else np.dtype(UF)
        _check(load().mpir_gmres_ldl_set_ordering(self._h, ordering), "set_ordering")
        _check(load().mpir_gmres_ldl_set_scaling(self._h, scaling), "set_scaling")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().mpir_gmres_ldl_destroy(h)
            self._h = None

    # reference setters (gmres_ir.hpp:291-303)
    def SetMaxIRIterations(self, n: int) -> None:
        _check(load().mpir_gmres_ldl_set_max_ir(self._h, int(n)), "SetMaxIRIterations")

    def SetMaxGmresIterations(self, n: int) -> None:
        _check(load().mpir_gmres_ldl_set_max_gmres(self._h, int(n)), "SetMaxGmresIterations")

    def SetTolerance(self, tol: float) -> None:
        _check(load().mpir_gmres_ldl_set_tol(self._h, float(tol)), "SetTolerance")

    def SetOrdering(self, ordering: int) -> None:
        _check(load().mpir_gmres_ldl_set_ordering(self._h, int(ordering)), "SetOrdering")

    def SetScaling(self, mode: int) -> None:
        """SCALE_NONE, SCALE_DIAG (the reference's, default) or SCALE_RUIZ."""
        _check(load().mpir_gmres_ldl_set_scaling(self._h, int(mode)), "SetScaling")

    def SetSolveSchedule(self, schedule: int) -> None:
        """SCHED_DAG (default) or SCHED_LEVELS for the triangular solves of M^{-1}."""
        _check(load().mpir_gmres_ldl_set_schedule(self._h, int(schedule)), "SetSolveSchedule")

    # phases
    def Analyze(self, Ap, Ai) -> None:
        Ap = np.ascontiguousarray(Ap, dtype=np.int32)
        Ai = np.ascontiguousarray(Ai, dtype=np.int32)
        n = len(Ap) - 1
        _check(load().mpir_gmres_ldl_analyze(self._h, n, _ptr(Ap), _ptr(Ai)), "Analyze")
        self._n = n

    def Factor(self, Ax) -> None:
        Ax = np.ascontiguousarray(Ax, dtype=np.float64)
        _check(load().mpir_gmres_ldl_factor(self._h, _ptr(Ax)), "Factor")

    def Compute(self, Ap, Ai, Ax, ILU_k: int = 0) -> None:
        """Compute(Ap, Ai, Ax, ILU_k): diagonal scaling + ordering + LDL^T (ILU_k is ignored)."""
        Ap = np.ascontiguousarray(Ap, dtype=np.int32)
        Ai = np.ascontiguousarray(Ai, dtype=np.int32)
        Ax = np.ascontiguousarray(Ax, dtype=np.float64)
        n = len(Ap) - 1
        _check(load().mpir_gmres_ldl_compute(self._h, n, _ptr(Ap), _ptr(Ai), _ptr(Ax)), "Compute")
        self._n = n

    def Solve(self, b) -> np.ndarray:
        """Solve(b) -> x; b of shape (n,) or (n, nrhs)."""
        b = np.asarray(b, dtype=np.float64)
        n = self._n
        if b.shape[0] != n:
            raise ValueError("b has the wrong length")  # PrecondGmres returns {} here
        if b.ndim == 1:
            bb = np.ascontiguousarray(b)
            x = np.empty(n)
            _check(load().mpir_gmres_ldl_solve(self._h, _ptr(bb), _ptr(x), 1, n), "Solve")
            return x
        bb = np.ascontiguousarray(b.T)
        xx = np.empty_like(bb)
        _check(load().mpir_gmres_ldl_solve(self._h, _ptr(bb), _ptr(xx), bb.shape[0], n), "Solve")
        return xx.T.copy()

    def stats(self) -> dict:
        st = GmresStats()
        _check(load().mpir_gmres_ldl_stats(self._h, C.byref(st)), "stats")
        return st.as_dict()

    def log(self) -> dict:
        """The last solve's log in the schema of the reference's VERBOSE build."""
        import json
        lib = load()
        n = int(lib.mpir_gmres_ldl_log(self._h, None, 0))
        _check(n, "log")
        buf = C.create_string_buffer(n + 1)
        _check(int(lib.mpir_gmres_ldl_log(self._h, buf, n + 1)), "log")
        return json.loads(buf.value.decode())

    def perm(self) -> np.ndarray:
        p = np.empty(self._n, dtype=np.int32)
        _check(load().mpir_gmres_ldl_perm(self._h, _ptr(p)), "perm")
        return p

    def export_factor(self):
        """(Lp, Li, Lx, D, Dinv) of the factor of P K_s P^T, QDLDL CSC layout."""
        s = self.stats()
        n, lnz = s["n"], s["nnz_l"]
        Lp = np.empty(n + 1, dtype=np.int64)
        Li = np.empty(max(lnz, 1), dtype=np.int32)
        Lx = np.empty(max(lnz, 1))
        D = np.empty(n)
        Dinv = np.empty(n)
        _check(load().mpir_gmres_ldl_export_factor(self._h, _ptr(Lp), _ptr(Li), _ptr(Lx), _ptr(D), _ptr(Dinv)),
               "export_factor")
        return Lp, Li[:lnz], Lx[:lnz], D, Dinv

    def time_kernels(self, what: str, reps: int = 10) -> float:
        """Average ms of one kernel group ('spmv', 'precond', 'factor') via HIP events."""
        code = {"spmv": 0, "precond": 1, "factor": 2}[what]
        ms = C.c_double()
        _check(load().mpir_gmres_ldl_time(self._h, code, int(reps), C.byref(ms)), "time_kernels")
        return ms.value

    def factor_dev(self, ax_ptr: int) -> None:
        _check(load().mpir_gmres_ldl_factor_dev(self._h, C.c_void_p(ax_ptr)), "factor_dev")

    def solve_dev(self, b_ptr: int, x_ptr: int, nrhs: int = 1) -> None:
        _check(load().mpir_gmres_ldl_solve_dev(self._h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), nrhs, self._n),
               "solve_dev")

    @property
    def handle(self):
        return self._h


# ---------------------------------------------------------------------- utilities
@dataclass
class KKT:
    n: int
    n1: int
    Ap: np.ndarray
    Ai: np.ndarray
    Ax: np.ndarray
    x_true: np.ndarray
    b: np.ndarray


def gen_kkt(dim=2, k=32, h_shift=1e-2, m=256, a_mode=0, cnt_lo=3, cnt_hi=3, d_value=1e-2, seed=0) -> KKT:
    """Deterministic quasi-definite KKT system (see include/mpir_hip.h)."""
    lib = load()
    n = C.c_int32()
    nnz = C.c_int64()
    n1 = C.c_int32()
    args = (dim, k, float(h_shift), int(m), a_mode, cnt_lo, cnt_hi, float(d_value), int(seed))
    _check(lib.mpir_gen_kkt(*args, C.byref(n), C.byref(nnz), C.byref(n1), None, None, None, None, None), "gen_kkt")
    Ap = np.empty(n.value + 1, dtype=np.int32)
    Ai = np.empty(nnz.value, dtype=np.int32)
    Ax = np.empty(nnz.value)
    xt = np.empty(n.value)
    b = np.empty(n.value)
    _check(lib.mpir_gen_kkt(*args, None, None, None, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(xt), _ptr(b)), "gen_kkt")
    return KKT(n.value, n1.value, Ap, Ai, Ax, xt, b)


def kkt_config(cfg: int, seed: int = 0, d_value: float | None = None, k: int | None = None) -> KKT:
    """The KKT workloads by name: 0 = 2D 32x32 + 256 constraints (n=1280);
    1 = 3D 64^3 + 65536 four-entry constraint rows (n=327680). ``k`` shrinks
    the grid of workload 1 for tests."""
    if cfg == 0:
        return gen_kkt(2, 32 if k is None else k, 1e-2, 256 if k is None else (k * k) // 4, 0, 3, 3,
                       1e-2 if d_value is None else d_value, seed)
    if cfg in (1, 2):
        kk = 64 if k is None else k
        return gen_kkt(3, kk, 1e-3, kk ** 3 // 4, 1, 4, 4, 1e-4 if d_value is None else d_value, seed)
    raise ValueError("unknown configuration")


def read_mtx(path):
    """A Matrix Market file read as examples/example.cpp:87-138 reads it:
    (n, Ap, Ai, Ax, symmetric) of the transposed triplets in CSC, i.e. the
    upper triangle of a symmetric matrix stored by its lower triangle."""
    lib = load()
    n, nnz, sym = C.c_int32(), C.c_int64(), C.c_int32()
    p = str(path).encode()
    _check(lib.mpir_read_mtx_upper(p, C.byref(n), C.byref(nnz), C.byref(sym), None, None, None), "read_mtx")
    Ap = np.empty(n.value + 1, np.int32)
    Ai = np.empty(nnz.value, np.int32)
    Ax = np.empty(nnz.value)
    _check(lib.mpir_read_mtx_upper(p, C.byref(n), C.byref(nnz), C.byref(sym), _ptr(Ap), _ptr(Ai), _ptr(Ax)),
           "read_mtx")
    return n.value, Ap, Ai, Ax, bool(sym.value)


def amd_order(Ap, Ai) -> np.ndarray:
    Ap = np.ascontiguousarray(Ap, dtype=np.int32)
    Ai = np.ascontiguousarray(Ai, dtype=np.int32)
    n = len(Ap) - 1
    perm = np.empty(n, dtype=np.int32)
    _check(load().mpir_amd_order(n, _ptr(Ap), _ptr(Ai), _ptr(perm)), "amd_order")
    return perm


def spmv_upper(Ap, Ai, Ax, x) -> np.ndarray:
    """y = K x from the upper CSC, FP64 on the GPU with the reference's rounding."""
    Ap = np.ascontiguousarray(Ap, dtype=np.int32)
    Ai = np.ascontiguousarray(Ai, dtype=np.int32)
    Ax = np.ascontiguousarray(Ax, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(len(Ap) - 1)
    _check(load().mpir_spmv_upper_f64(len(Ap) - 1, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(x), _ptr(y)), "spmv_upper")
    return y


def residual_upper(Ap, Ai, Ax, x, b) -> np.ndarray:
    Ap = np.ascontiguousarray(Ap, dtype=np.int32)
    Ai = np.ascontiguousarray(Ai, dtype=np.int32)
    Ax = np.ascontiguousarray(Ax, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    r = np.empty(len(Ap) - 1)
    _check(load().mpir_residual_upper_f64(len(Ap) - 1, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(x), _ptr(b), _ptr(r)),
           "residual_upper")
    return r


__all__ = [
    "QDLDL_etree", "QDLDL_factor", "QDLDL_solve", "LUIR_QDLDL", "GmresLDLIR", "MpirError", "KKT", "gen_kkt", "kkt_config",
    "amd_order", "read_mtx", "spmv_upper", "residual_upper", "device_ok", "F64", "F32", "F16", "BF16", "ORDER_NATURAL",
    "ORDER_AMD", "ORDER_ND", "SCALE_NONE", "SCALE_DIAG", "SCALE_RUIZ", "SCHED_LEVELS", "SCHED_DAG",
]
