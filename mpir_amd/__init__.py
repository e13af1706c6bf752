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
PREC_LDL, PREC_FGPILU = 0, 1  # preconditioner of GMRES-IR


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


def _value_suffix(vt, prec):
    """Entry-point suffix of a value array: float64 / float32 / float16, or
    bfloat16 bit patterns in uint16 arrays with prec='bf16'."""
    if prec is not None:
        if str(prec).lower() != "bf16" or vt != np.uint16:
            raise TypeError("prec='bf16' takes uint16 arrays of bfloat16 bit patterns")
        return "bf16"
    suf = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32", np.dtype(np.float16): "f16"}.get(vt)
    if suf is None:
        raise TypeError("values must be float64, float32, float16 (or uint16 with prec='bf16')")
    return suf


def QDLDL_factor(n, Ap, Ai, Ax, Lp, Li, Lx, D, Dinv, Lnz, etree, bwork=None, iwork=None, fwork=None,
                 prec=None) -> int:
    """QDLDL_factor<INT,PREC> (qdldl.hpp:172-345) on the GPU; returns #D>0 or -1 on a zero pivot.
    PREC is the dtype of Ax/Lx/D/Dinv: float64, float32 or float16; bfloat16 as
    uint16 bit patterns with prec='bf16' (16-bit factors: FP32 arithmetic,
    values rounded to the format as they are stored)."""
    it = np.asarray(Ap).dtype
    vt = np.asarray(Ax).dtype
    suffix = _value_suffix(vt, prec)
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


def QDLDL_solve(n, Lp, Li, Lx, Dinv, x, prec=None) -> None:
    """QDLDL_solve<INT,FACTPREC,XPREC> (qdldl.hpp:417-432) on the GPU, x overwritten.
    FACTPREC as for QDLDL_factor (prec='bf16' for bfloat16 bit patterns)."""
    it = np.asarray(Lp).dtype
    ft = np.asarray(Lx).dtype
    xt = np.asarray(x).dtype
    key = (_value_suffix(ft, prec),
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


def qdldl_solve_plan_builds() -> int:
    """Plans this thread's QDLDL_solve engines have built: a call with the
    same factor pattern as the previous one reuses its plan."""
    out = C.c_int64(0)
    _check(int(load().mpir_qdldl_solve_plan_builds(C.byref(out))), "qdldl_solve_plan_builds")
    return int(out.value)


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
                 scaling: int = SCALE_DIAG, preconditioner: int | None = None, MAX_SWEEPS: int = 10):
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
        self._prec = (PREC_LDL, 0, int(MAX_SWEEPS))
        if preconditioner is not None:
            self.SetPreconditioner(preconditioner, 0, MAX_SWEEPS)

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

    def SetPreconditioner(self, kind: int, ILU_k: int = 0, MAX_SWEEPS: int = 10) -> None:
        """PREC_LDL (default): L D L^T of the reordered matrix; PREC_FGPILU: the
        reference's shipped Compute, FGPILU of the scaled matrix on the
        ILU(ILU_k) pattern with at most MAX_SWEEPS sweeps (gmres_ir.hpp:323-531)."""
        _check(load().mpir_gmres_ldl_set_preconditioner(self._h, int(kind), int(ILU_k), int(MAX_SWEEPS)),
               "SetPreconditioner")
        self._prec = (int(kind), int(ILU_k), int(MAX_SWEEPS))

    def SetSolveSchedule(self, schedule: int) -> None:
        """SCHED_DAG (default) or SCHED_LEVELS for the triangular solves of M^{-1}."""
        _check(load().mpir_gmres_ldl_set_schedule(self._h, int(schedule)), "SetSolveSchedule")

    def SetAmalgamation(self, level: int) -> None:
        """Supernode amalgamation, before Analyze: 0 none, 1 .. 3 (default;
        fewer, larger fronts and a shallower tree for more stored zeros)."""
        _check(load().mpir_gmres_ldl_set_amalgamation(self._h, int(level)), "SetAmalgamation")

    def SetPanelSplit(self, min_tasks: int) -> None:
        """Panel-step launches of at least min_tasks tasks (default 512; 0
        never) eliminate each front's diagonal block once, in a launch of
        their own; same factor bit for bit."""
        _check(load().mpir_gmres_ldl_set_panel_split(self._h, int(min_tasks)), "SetPanelSplit")

    def SetFactorSchedule(self, schedule: int) -> None:
        """SCHED_LEVELS (default: one launch per level and panel step) or
        SCHED_DAG (the levels above the small fronts in one dependency-driven
        launch) for the numeric factorization; same factor bit for bit."""
        _check(load().mpir_gmres_ldl_set_factor_schedule(self._h, int(schedule)), "SetFactorSchedule")

    def factor_schedule_info(self) -> dict:
        """schedule, first level of the dependency-driven launch, levels, tasks of that launch."""
        sc, lv, nl, nt = C.c_int(), C.c_int32(), C.c_int32(), C.c_int64()
        _check(load().mpir_gmres_ldl_factor_schedule_info(self._h, C.byref(sc), C.byref(lv), C.byref(nl), C.byref(nt)),
               "factor_schedule_info")
        return {"schedule": sc.value, "dag_level": lv.value, "nlevels": nl.value, "dag_tasks": nt.value}

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
        """Compute(Ap, Ai, Ax, ILU_k): diagonal scaling, then the L D L^T of the
        reordered matrix (ILU_k unused) or, with PREC_FGPILU, FGPILU on the
        ILU(ILU_k) pattern."""
        kind, k, sweeps = self._prec
        if kind == PREC_FGPILU and int(ILU_k) != k:
            self.SetPreconditioner(kind, ILU_k, sweeps)
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

    def export_ilu(self):
        """(Up, Ui, Ux, Lx, UDinv) of the FGPILU factors: U by columns, L by rows on
        the same indices (unit diagonal stored), 1 / U_jj."""
        nnz = C.c_int64()
        _check(load().mpir_gmres_ldl_ilu_nnz(self._h, C.byref(nnz)), "ilu_nnz")
        n, m = self._n, int(nnz.value)
        Up = np.empty(n + 1, dtype=np.int32)
        Ui = np.empty(max(m, 1), dtype=np.int32)
        Ux = np.empty(max(m, 1))
        Lx = np.empty(max(m, 1))
        UDinv = np.empty(n)
        _check(load().mpir_gmres_ldl_export_ilu(self._h, _ptr(Up), _ptr(Ui), _ptr(Ux), _ptr(Lx), _ptr(UDinv)),
               "export_ilu")
        return Up, Ui[:m], Ux[:m], Lx[:m], UDinv

    def trace_factor(self) -> np.ndarray:
        """One factorization with the timeline of its dependency-driven launch:
        an (ntask, 4) uint64 array (kind << 28 | supernode, drawn, waits
        ended, counted << 16 | workgroup; device clock at 100 MHz)."""
        nt = C.c_int64()
        _check(load().mpir_gmres_ldl_trace_factor(self._h, None, 0, C.byref(nt)), "trace_factor")
        out = np.zeros((max(nt.value, 1), 4), dtype=np.uint64)
        _check(load().mpir_gmres_ldl_trace_factor(self._h, _ptr(out), out.size, C.byref(nt)), "trace_factor")
        return out[:nt.value]

    def trace_precond(self):
        """One M^{-1} application with the timeline of its tasks: (an (ntask, 4)
        uint64 array: kind | first supernode << 4, drawn, wait ended,
        done << 16 | workgroup; device clock at 100 MHz), forward task count."""
        nt, nf = C.c_int64(), C.c_int64()
        _check(load().mpir_gmres_ldl_trace_precond(self._h, None, 0, C.byref(nt), C.byref(nf)), "trace_precond")
        out = np.zeros((max(nt.value, 1), 4), dtype=np.uint64)
        _check(load().mpir_gmres_ldl_trace_precond(self._h, _ptr(out), out.size, C.byref(nt), C.byref(nf)),
               "trace_precond")
        return out[:nt.value], nf.value

    def supernode_info(self) -> dict:
        """Per supernode of the analyzed plan, in the indices the traces record:
        'level' (height above the leaves in the supernodal tree), 'ncols',
        'nrows' (rows below the diagonal block); int32 arrays."""
        ns = C.c_int32()
        _check(load().mpir_gmres_ldl_supernode_info(self._h, None, None, None, 0, C.byref(ns)), "supernode_info")
        out = {k: np.zeros(max(ns.value, 1), dtype=np.int32) for k in ("level", "ncols", "nrows")}
        _check(load().mpir_gmres_ldl_supernode_info(self._h, _ptr(out["level"]), _ptr(out["ncols"]),
                                                    _ptr(out["nrows"]), ns.value, C.byref(ns)), "supernode_info")
        return {k: v[:ns.value] for k, v in out.items()}

    def time_kernels(self, what: str, reps: int = 10) -> float:
        """Average ms of one kernel group ('spmv', 'precond', 'factor') via HIP events."""
        code = {"spmv": 0, "precond": 1, "factor": 2, "precond_multi": 3}[what]
        ms = C.c_double()
        _check(load().mpir_gmres_ldl_time(self._h, code, int(reps), C.byref(ms)), "time_kernels")
        return ms.value

    def SetTuning(self, upd_group: int = 4, super_tiles: int = 8, narrow_rows: int = 1024) -> None:
        """A tuning without timing it (see mpir_gmres_ldl_set_tuning)."""
        _check(load().mpir_gmres_ldl_set_tuning(self._h, int(upd_group), int(super_tiles), int(narrow_rows)),
               "SetTuning")

    def Autotune(self, reps: int = 3) -> dict:
        """Times the factor's update grouping / super-block edge and the
        solves' narrow-front limit / schedule on this factor, keeps the
        fastest and refactors with it (see mpir_gmres_ldl_autotune)."""
        choice = np.zeros(5, dtype=np.int32)
        ms = np.zeros(4)
        _check(load().mpir_gmres_ldl_autotune(self._h, int(reps), _ptr(choice), _ptr(ms)), "Autotune")
        return {"upd_group": int(choice[0]), "super_tiles": int(choice[1]), "narrow_rows": int(choice[2]),
                "schedule": int(choice[3]), "candidates": int(choice[4]), "factor_ms": float(ms[0]),
                "precond_ms": float(ms[1]), "factor_ms_before": float(ms[2]), "precond_ms_before": float(ms[3])}

    def factor_dev(self, ax_ptr: int) -> None:
        _check(load().mpir_gmres_ldl_factor_dev(self._h, C.c_void_p(ax_ptr)), "factor_dev")

    def solve_dev(self, b_ptr: int, x_ptr: int, nrhs: int = 1) -> None:
        _check(load().mpir_gmres_ldl_solve_dev(self._h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), nrhs, self._n),
               "solve_dev")

    @property
    def handle(self):
        return self._h


# ---------------------------------------------------------------------- batches
def _ptr_array(arrs) -> C.Array:
    return (C.c_void_p * len(arrs))(*[a.ctypes.data if a is not None else None for a in arrs])


class BatchLDLIR:
    """Many independent systems, each solved as GmresLDLIR<UF,double,double>
    (gmres_ir.hpp:534-741) solves one, on one GPU: the systems are joined
    into one block-diagonal system (one factorization plan over the forest
    of their elimination trees; lock-step GMRES-IR with per-system scalars).
    A system's result is bit-identical alone and in any batch."""

    def __init__(self, UF=np.float32, ordering: int = ORDER_ND, scaling: int = SCALE_DIAG, threads: int = 8):
        self._s = GmresLDLIR(UF, ordering=ordering, scaling=scaling)
        # the level schedule of the triangular solves: on the forest of a
        # batch it is faster than the dependency-driven one, and it never
        # waits across workgroups (SetSolveSchedule changes it)
        self._s.SetSolveSchedule(SCHED_LEVELS)
        self.threads = int(threads)
        self._n: list[int] = []
        self._keep = []

    # the reference's setters apply to every system
    def SetMaxIRIterations(self, n: int) -> None:
        self._s.SetMaxIRIterations(n)

    def SetMaxGmresIterations(self, n: int) -> None:
        self._s.SetMaxGmresIterations(n)

    def SetTolerance(self, tol: float) -> None:
        self._s.SetTolerance(tol)

    def SetSolveSchedule(self, schedule: int) -> None:
        """SCHED_LEVELS (the default for batches) or SCHED_DAG."""
        self._s.SetSolveSchedule(schedule)

    def SetFactorSchedule(self, schedule: int) -> None:
        """SCHED_LEVELS (default) or SCHED_DAG for the batch's factorization;
        same factor bit for bit."""
        self._s.SetFactorSchedule(schedule)

    def SetPanelSplit(self, min_tasks: int) -> None:
        """As GmresLDLIR.SetPanelSplit, for the batch's factorization."""
        self._s.SetPanelSplit(min_tasks)

    def Analyze(self, patterns) -> None:
        """patterns: list of (Ap, Ai), upper CSC of each system."""
        pats = [(np.ascontiguousarray(a, dtype=np.int32), np.ascontiguousarray(i, dtype=np.int32)) for a, i in patterns]
        n = np.array([len(a) - 1 for a, _ in pats], dtype=np.int32)
        _check(load().mpir_batch_analyze(self._s.handle, len(pats), _ptr(n), _ptr_array([a for a, _ in pats]),
                                         _ptr_array([i for _, i in pats]), self.threads), "batch Analyze")
        self._n = [int(v) for v in n]

    def Factor(self, values) -> None:
        vals = [np.ascontiguousarray(v, dtype=np.float64) for v in values]
        if len(vals) != len(self._n):
            raise ValueError("one value array per system")
        _check(load().mpir_batch_factor(self._s.handle, _ptr_array(vals)), "batch Factor")

    def Compute(self, systems) -> None:
        """systems: list of (Ap, Ai, Ax)."""
        self.Analyze([(a, i) for a, i, _ in systems])
        self.Factor([x for _, _, x in systems])

    def Solve(self, rhs) -> list:
        """rhs: per system an array (n,) or (n, nrhs) (or None for none);
        returns the solutions in the same shapes."""
        if len(rhs) != len(self._n):
            raise ValueError("one right-hand-side array per system")
        bs, xs, nr, shapes = [], [], [], []
        for b, n in zip(rhs, self._n):
            if b is None:
                bs.append(None)
                xs.append(None)
                nr.append(0)
                shapes.append(None)
                continue
            b = np.asarray(b, dtype=np.float64)
            if b.shape[0] != n:
                raise ValueError("right-hand side of the wrong length")
            shapes.append(b.shape)
            bb = np.ascontiguousarray(b.reshape(n, -1).T)
            bs.append(bb)
            xs.append(np.empty_like(bb))
            nr.append(bb.shape[0])
        nrhs = np.array(nr, dtype=np.int32)
        _check(load().mpir_batch_solve(self._s.handle, _ptr(nrhs), _ptr_array(bs), _ptr_array(xs)), "batch Solve")
        out = []
        for x, shp in zip(xs, shapes):
            out.append(None if x is None else x.T.reshape(shp).copy())
        return out

    def factor_dev(self, ptrs) -> None:
        arr = (C.c_void_p * len(ptrs))(*ptrs)
        _check(load().mpir_batch_factor_dev(self._s.handle, arr), "batch factor_dev")

    def solve_dev(self, nrhs, b_ptrs, x_ptrs) -> None:
        nr = np.ascontiguousarray(nrhs, dtype=np.int32)
        _check(load().mpir_batch_solve_dev(self._s.handle, _ptr(nr), (C.c_void_p * len(b_ptrs))(*b_ptrs),
                                           (C.c_void_p * len(x_ptrs))(*x_ptrs)), "batch solve_dev")

    def stats(self) -> dict:
        """positive_pivots (per system); ir_iterations, gmres_iterations,
        stop_reason, final_rel_residual, done_ms of shape (systems, max_rhs)."""
        S = len(self._n)
        lib = load()
        mr = C.c_int32()
        _check(lib.mpir_batch_stats(self._s.handle, C.byref(mr), None, None, None, None, None, None), "batch stats")
        R = max(mr.value, 1)
        pos = np.zeros(S, np.int32)
        ir, gm, stop = (np.zeros(S * R, np.int32) for _ in range(3))
        rel, done = np.zeros(S * R), np.zeros(S * R)
        _check(lib.mpir_batch_stats(self._s.handle, C.byref(mr), _ptr(pos), _ptr(ir), _ptr(gm), _ptr(stop), _ptr(rel),
                                    _ptr(done)), "batch stats")
        sh = (S, R)
        counts = np.zeros(3, np.int32)
        _check(lib.mpir_batch_graph(self._s.handle, -1, _ptr(counts)), "batch graph")
        return {"max_rhs": mr.value, "positive_pivots": pos, "ir_iterations": ir.reshape(sh),
                "gmres_iterations": gm.reshape(sh), "stop_reason": stop.reshape(sh),
                "final_rel_residual": rel.reshape(sh), "done_ms": done.reshape(sh), "graph_nodes": int(counts[0]),
                "arnoldi_steps": int(counts[1]), "precond_sweeps": int(counts[2]), "totals": self._s.stats()}

    def SetGraph(self, enable: bool) -> None:
        """Replay repeated solves as one captured HIP graph (default) or launch
        every kernel of every solve (see mpir_batch_graph)."""
        _check(load().mpir_batch_graph(self._s.handle, 1 if enable else 0, None), "SetGraph")

    def time_kernels(self, what: str, reps: int = 5) -> float:
        return self._s.time_kernels(what, reps)

    def Autotune(self, reps: int = 3) -> dict:
        """The level-tile autotuner on the joined batch (GmresLDLIR.Autotune)."""
        return self._s.Autotune(reps)

    def SetTuning(self, upd_group: int = 4, super_tiles: int = 8, narrow_rows: int = 1024) -> None:
        self._s.SetTuning(upd_group, super_tiles, narrow_rows)

    @property
    def sizes(self) -> list:
        return list(self._n)


class Server:
    """A stream of solve requests with continuous batching on one GPU (see
    include/mpir_hip.h, mpir_server_*): requests are admitted in arrival
    order into batches by device-memory footprint, `lanes` batches run
    concurrently, and each pattern's symbolic phase is computed once."""

    def __init__(self, UF=np.float32, lanes: int = 2, budget_bytes: int = 64 << 30, max_batch: int = 64,
                 tol: float = 1e-10, ordering: int = ORDER_ND, keep_x: bool = False, threads: int = 8):
        uf = _PREC_OF.get(np.dtype(UF))
        if uf is None:
            raise TypeError("UF must be float64 or float32")
        h = C.c_void_p()
        _check(load().mpir_server_create(uf, int(lanes), int(budget_bytes), int(max_batch), float(tol), int(ordering),
                                         int(bool(keep_x)), int(threads), C.byref(h)), "Server()")
        self._h = h
        self._n: list[int] = []
        self._nnz: list[int] = []
        self._last = None
        self._req: dict = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().mpir_server_destroy(h)
            self._h = None

    def add_pattern(self, Ap, Ai, base_Ax, n1: int) -> int:
        Ap = np.ascontiguousarray(Ap, dtype=np.int32)
        Ai = np.ascontiguousarray(Ai, dtype=np.int32)
        base = np.ascontiguousarray(base_Ax, dtype=np.float64)
        pid = _check(load().mpir_server_add_pattern(self._h, len(Ap) - 1, _ptr(Ap), _ptr(Ai), _ptr(base), int(n1)),
                     "add_pattern")
        self._n.append(len(Ap) - 1)
        self._nnz.append(int(Ap[-1]))
        return int(pid)

    def add_patterns(self, kkts, threads: int = 8) -> int:
        """Registers many KKT systems' patterns (base values = their values),
        analysing them on `threads` host threads; returns the first id."""
        Ap = [np.ascontiguousarray(P.Ap, dtype=np.int32) for P in kkts]
        Ai = [np.ascontiguousarray(P.Ai, dtype=np.int32) for P in kkts]
        base = [np.ascontiguousarray(P.Ax, dtype=np.float64) for P in kkts]
        n = np.array([len(a) - 1 for a in Ap], dtype=np.int32)
        n1 = np.array([P.n1 for P in kkts], dtype=np.int32)
        first = _check(load().mpir_server_add_patterns(self._h, len(kkts), _ptr(n), _ptr_array(Ap), _ptr_array(Ai),
                                                       _ptr_array(base), _ptr(n1), int(threads)), "add_patterns")
        self._n.extend(int(v) for v in n)
        self._nnz.extend(int(a[-1]) for a in Ap)
        return int(first)

    def run(self, pattern, seed, nrhs, t_arrive_ms) -> dict:
        """Runs the request stream to completion; returns per-request arrays."""
        pattern = np.ascontiguousarray(pattern, dtype=np.int32)
        seed = np.ascontiguousarray(seed, dtype=np.uint64)
        nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
        ta = np.ascontiguousarray(t_arrive_ms, dtype=np.float64)
        m = len(pattern)
        out = {k: np.zeros(m) for k in ("t_admit", "t_done", "rel_residual_max")}
        out.update({k: np.zeros(m, np.int32) for k in ("batch", "lane", "status", "ir_max", "gmres_max")})
        _check(load().mpir_server_run(self._h, m, _ptr(pattern), _ptr(seed), _ptr(nrhs), _ptr(ta), _ptr(out["t_admit"]),
                                      _ptr(out["t_done"]), _ptr(out["batch"]), _ptr(out["lane"]), _ptr(out["status"]),
                                      _ptr(out["ir_max"]), _ptr(out["gmres_max"]), _ptr(out["rel_residual_max"])),
               "Server.run")
        out["t_arrive"] = ta
        out["latency_ms"] = out["t_done"] - ta
        self._last = (pattern, nrhs)
        return out

    def submit(self, pattern: int, Ax, B) -> int:
        """A request with the caller's values Ax (the pattern's upper-CSC
        order) and right-hand sides B (n, or n x nrhs); returns its id."""
        if not 0 <= pattern < len(self._n):
            raise ValueError(f"unknown pattern {pattern}")
        n = self._n[pattern]
        Ax = np.ascontiguousarray(Ax, dtype=np.float64)
        if Ax.shape != (self._nnz[pattern],):
            raise ValueError("Ax must hold the pattern's nnz values")
        B = np.asarray(B, dtype=np.float64)
        B2 = B.reshape(n, -1)
        nrhs = B2.shape[1]
        bcol = np.ascontiguousarray(B2.T)  # column-major n x nrhs
        rid = C.c_int64(-1)
        _check(load().mpir_server_submit(self._h, int(pattern), _ptr(Ax), _ptr(bcol), int(nrhs), C.byref(rid)),
               "Server.submit")
        self._req[int(rid.value)] = (n, nrhs, B.ndim == 1)
        return int(rid.value)

    def poll(self, rid: int, timeout_ms: float = 0.0, release: bool = True):
        """(done, X, info): X has the shape of the request's B once done.
        timeout_ms: 0 looks only, < 0 waits without limit. A failed request
        raises MpirError."""
        n, nrhs, vec = self._req[rid]
        x = np.empty(n * nrhs)
        info = _lib.ServerRequest()
        rc = int(load().mpir_server_poll(self._h, int(rid), float(timeout_ms), _ptr(x), C.byref(info),
                                         int(bool(release))))
        if rc == 0:
            return False, None, None
        if release:
            del self._req[rid]
        if rc < 0:
            raise MpirError(rc, "Server.poll")
        X = x.reshape(nrhs, n).T.copy()
        return True, (X[:, 0] if vec else X), info.as_dict()

    def summary(self) -> dict:
        c = np.zeros(6, np.int64)
        w = np.zeros(5)
        _check(load().mpir_server_summary(self._h, _ptr(c), _ptr(w)), "summary")
        keys = ["completed", "failed", "batches", "max_backlog", "peak_inflight_bytes", "max_concurrent"]
        d = {k: int(v) for k, v in zip(keys, c)}
        for k, v in zip(["wall_ms", "prepare_ms", "plan_ms", "factor_ms", "solve_ms"], w):
            d[k] = float(v)
        j, u, g = C.c_double(), C.c_double(), C.c_int64()
        _check(load().mpir_server_plan_phases(self._h, C.byref(j), C.byref(u), C.byref(g)), "summary")
        d["plan_join_ms"], d["plan_upload_ms"], d["device_regrowths"] = j.value, u.value, g.value
        return d

    def solution(self, i: int) -> np.ndarray:
        pattern, nrhs = self._last
        n, r = self._n[pattern[i]], int(nrhs[i])
        x = np.empty(n * r)
        _check(load().mpir_server_solution(self._h, int(i), _ptr(x)), "solution")
        return x.reshape(r, n).T.copy()

    def request_data(self, pattern: int, seed: int, nrhs: int):
        """(Ax, B, X_true) of a request, on the host."""
        n = self._n[pattern]
        Ax = np.empty(self._nnz[pattern])
        B = np.empty(n * nrhs)
        X = np.empty(n * nrhs)
        _check(load().mpir_server_request_data(self._h, int(pattern), C.c_uint64(int(seed)), int(nrhs), _ptr(Ax),
                                               _ptr(B), _ptr(X)), "request_data")
        return Ax, B.reshape(nrhs, n).T.copy(), X.reshape(nrhs, n).T.copy()


def serving_patterns(count: int = 64, k_lo: int = 16, k_hi: int = 48, seed: int = 0, d_value: float = 1e-3,
                     h_shift: float = 1e-3) -> list:
    """The serving pool: pattern p is the 3D k_p^3 Laplacian, k_p = k_lo +
    round((k_hi - k_lo) p / (count - 1)), plus m = k_p^3 / 4 constraint rows
    of 4 entries (a uniform cell and its +x, +y, +z neighbours), D = d_value I;
    the base values of each pattern."""
    out = []
    for p in range(count):
        k = k_lo + int(round((k_hi - k_lo) * p / max(1, count - 1)))
        out.append(gen_kkt(3, k, h_shift, k ** 3 // 4, 1, 4, 4, d_value, seed * 7919 + p))
    return out


def poisson_stream(rate_per_s: float, seconds: float, npatterns: int, seed: int = 0, nrhs_lo: int = 1,
                   nrhs_hi: int = 4):
    """Arrival times (ms) of a Poisson process, a uniform pattern and a
    uniform right-hand-side count per request, and a seed for its data."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t, times = 0.0, []
    while True:
        t += rng.exponential(1000.0 / rate_per_s)
        if t >= seconds * 1000.0:
            break
        times.append(t)
    m = len(times)
    pattern = rng.integers(0, npatterns, m).astype(np.int32)
    nrhs = rng.integers(nrhs_lo, nrhs_hi + 1, m).astype(np.int32)
    seeds = rng.integers(1, 2 ** 62, m, dtype=np.int64).astype(np.uint64)
    return np.asarray(times), pattern, nrhs, seeds


def pattern_cache_info() -> dict:
    e, h, m = C.c_int64(), C.c_int64(), C.c_int64()
    _check(load().mpir_pattern_cache_info(C.byref(e), C.byref(h), C.byref(m)), "pattern_cache_info")
    return {"entries": e.value, "hits": h.value, "misses": m.value}


def batch_plan_info(patterns, ordering: int = ORDER_ND) -> dict:
    """Host-only symbolic phase of a batch (no device needed)."""
    pats = [(np.ascontiguousarray(a, dtype=np.int32), np.ascontiguousarray(i, dtype=np.int32)) for a, i in patterns]
    S = len(pats)
    n = np.array([len(a) - 1 for a, _ in pats], dtype=np.int32)
    info = np.zeros(8 + 2 * S, dtype=np.int64)
    _check(load().mpir_batch_plan_info(S, _ptr(n), _ptr_array([a for a, _ in pats]), _ptr_array([i for _, i in pats]),
                                       ordering, _ptr(info)), "batch_plan_info")
    keys = ["n", "nnz_l", "factor_flops", "nsupernodes", "nlevels", "panel_entries", "panel_tasks", "fwd_tasks"]
    out = {k: int(v) for k, v in zip(keys, info[:8])}
    out["nnz_l_each"] = info[8::2].copy()
    out["bytes_each"] = info[9::2].copy()
    return out


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


@dataclass
class BatchSystem:
    kkt: KKT
    dim: int
    k: int
    B: np.ndarray       # right-hand sides, (n, nrhs)
    X_true: np.ndarray  # (n, nrhs)


def batch_config(nsys: int = 512, nrhs: int = 8, seed: int = 0, n1_lo: float = 2e3, n1_hi: float = 6e4,
                 d_value: float = 1e-3, h_shift: float = 1e-3) -> list:
    """The batched workload: `nsys` independent quasi-definite systems.
    System i draws (numpy PCG64, seed) n1 log-uniform in [n1_lo, n1_hi], a 2D
    or 3D grid by a fair coin with side round(sqrt(n1)) / round(cbrt(n1)),
    m = floor(k^dim / 4) constraint rows of 3-5 entries (a uniform cell and
    its nearest in-grid neighbours, values N(0,1)), D = d_value I, H shift
    h_shift, and `nrhs` right-hand sides b = K x_true with x_true ~ U(-1,1)."""
    import scipy.sparse as sp
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(nsys):
        n1 = float(np.exp(rng.uniform(np.log(n1_lo), np.log(n1_hi))))
        dim = 2 if rng.uniform() < 0.5 else 3
        k = max(3, int(round(n1 ** (1.0 / dim))))
        n1r = k ** dim
        P = gen_kkt(dim, k, h_shift, n1r // 4, 2, 3, 5, d_value, seed * 1000003 + i)
        X = rng.uniform(-1.0, 1.0, size=(P.n, nrhs))
        U = sp.csc_matrix((P.Ax, P.Ai, P.Ap), shape=(P.n, P.n))
        K = (U + sp.triu(U, 1).T).tocsr()
        B = np.asarray(K @ X)
        out.append(BatchSystem(P, dim, k, B, X))
    return out


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
    "PREC_LDL", "PREC_FGPILU", "BatchLDLIR", "BatchSystem", "batch_config", "batch_plan_info", "pattern_cache_info", "Server",
    "serving_patterns", "poisson_stream",
]
