"""ctypes access to the oracle (oracle/liboracle.so) and to the compiled
reference headers (oracle/_ref/libqdldl_ref.so). Test infrastructure only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libqdldl_ref.so"
P = C.c_void_p
i32, i64, f64 = C.c_int32, C.c_int64, C.c_double


def ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def _bind(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


class Oracle:
    def __init__(self):
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        self.lib = C.CDLL(str(ORACLE_SO))
        _bind(self.lib, {
            "or_etree": (i64, [i32, P, P, P, P, P]),
            "or_factor_f64": (i32, [i32, P, P, P, P, P, P, P, P, P, P]),
            "or_factor_f32": (i32, [i32, P, P, P, P, P, P, P, P, P, P]),
            "or_solve_f64_f64": (None, [i32, P, P, P, P, P]),
            "or_solve_f32_f64": (None, [i32, P, P, P, P, P]),
            "or_solve_f32_f32": (None, [i32, P, P, P, P, P]),
            "or_matmul_upper": (None, [i32, P, P, P, P, P]),
            "or_dnrm2": (f64, [i64, P]),
            "or_dot": (f64, [i64, P, P]),
            "or_infnrm": (f64, [i64, P]),
            "or_gmres_ir": (i32, [i32, P, P, P, P, i32, i64, i64, f64, P, P, P]),
            "or_scaling": (None, [i32, P, P, P, i32, P, P]),
            "or_luir_f64": (i32, [i32, P, P, P, P, P, i32, P]),
            "or_luir_f32": (i32, [i32, P, P, P, P, P, i32, P]),
            "or_fgpilu": (i64, [i32, P, P, P, i32, i32, i32, f64, P, P, P, P, P, P]),
            "or_fgpilu_solve": (None, [i32, P, P, P, P, P, P]),
            "or_gmres_ilu_ir": (i32, [i32, P, P, P, i32, i32, i32, i64, i64, f64, P, P, P]),
        })

    def etree(self, Ap, Ai):
        n = len(Ap) - 1
        w, Lnz, et = (np.zeros(n, np.int32) for _ in range(3))
        tot = self.lib.or_etree(n, ptr(Ap), ptr(Ai), ptr(w), ptr(Lnz), ptr(et))
        return int(tot), Lnz, et

    def factor(self, Ap, Ai, Ax, dtype=np.float64):
        n = len(Ap) - 1
        tot, Lnz, et = self.etree(Ap, Ai)
        if tot < 0:
            return tot, None
        Lp = np.zeros(n + 1, np.int32)
        Li = np.zeros(max(tot, 1), np.int32)
        Lx = np.zeros(max(tot, 1), dtype)
        D = np.zeros(n, dtype)
        Dinv = np.zeros(n, dtype)
        ax = np.ascontiguousarray(Ax, dtype=dtype)
        fn = self.lib.or_factor_f64 if dtype == np.float64 else self.lib.or_factor_f32
        ret = fn(n, ptr(Ap), ptr(Ai), ptr(ax), ptr(Lp), ptr(Li), ptr(Lx), ptr(D), ptr(Dinv), ptr(Lnz), ptr(et))
        return int(ret), dict(Lp=Lp, Li=Li[:tot], Lx=Lx[:tot], D=D, Dinv=Dinv, Lnz=Lnz, etree=et)

    def solve(self, Lp, Li, Lx, Dinv, b, xdtype=np.float64):
        n = len(Lp) - 1
        x = np.array(b, dtype=xdtype)
        if Lx.dtype == np.float64:
            self.lib.or_solve_f64_f64(n, ptr(Lp), ptr(Li), ptr(Lx), ptr(Dinv), ptr(x))
        elif xdtype == np.float64:
            self.lib.or_solve_f32_f64(n, ptr(Lp), ptr(Li), ptr(Lx), ptr(Dinv), ptr(x))
        else:
            self.lib.or_solve_f32_f32(n, ptr(Lp), ptr(Li), ptr(Lx), ptr(Dinv), ptr(x))
        return x

    def matmul(self, Ap, Ai, Ax, x):
        y = np.zeros(len(Ap) - 1)
        self.lib.or_matmul_upper(len(Ap) - 1, ptr(Ap), ptr(Ai), ptr(np.ascontiguousarray(Ax, np.float64)),
                                 ptr(np.ascontiguousarray(x, np.float64)), ptr(y))
        return y

    def dnrm2(self, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.lib.or_dnrm2(len(x), ptr(x))

    def dot(self, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        return self.lib.or_dot(len(a), ptr(a), ptr(b))

    def infnrm(self, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.lib.or_infnrm(len(x), ptr(x))

    def luir(self, Ap, Ai, Ax, b, factprec=np.float32, maxiter=10):
        """LUIR_QDLDL restated (luir.hpp:33-241): (ret, x, steps, rel_residual)."""
        n = len(Ap) - 1
        x = np.zeros(n)
        info = np.zeros(3)
        f = self.lib.or_luir_f64 if factprec == np.float64 else self.lib.or_luir_f32
        ret = f(n, ptr(Ap), ptr(Ai), ptr(np.ascontiguousarray(Ax, np.float64)),
                ptr(np.ascontiguousarray(b, np.float64)), ptr(x), int(maxiter), ptr(info))
        return ret, x, int(info[0]), float(info[1])

    def fgpilu(self, Ap, Ai, Ax, uf=np.float64, ilu_k=0, max_sweeps=10, tol=1e-10):
        """FGPILU factors of the matrix as given: (Up, Ui, Ux, Lx, UDinv, info)."""
        n = len(Ap) - 1
        Ap = np.ascontiguousarray(Ap, np.int32)
        Ai = np.ascontiguousarray(Ai, np.int32)
        Ax = np.ascontiguousarray(Ax, np.float64)
        u = 0 if uf == np.float64 else 1
        nnz = int(self.lib.or_fgpilu(n, ptr(Ap), ptr(Ai), ptr(Ax), u, ilu_k, max_sweeps, tol, None, None, None, None,
                                     None, None))
        Up = np.zeros(n + 1, np.int32)
        Ui = np.zeros(max(nnz, 1), np.int32)
        Ux, Lx, Di, info = np.zeros(max(nnz, 1)), np.zeros(max(nnz, 1)), np.zeros(n), np.zeros(3)
        self.lib.or_fgpilu(n, ptr(Ap), ptr(Ai), ptr(Ax), u, ilu_k, max_sweeps, tol, ptr(Up), ptr(Ui), ptr(Ux), ptr(Lx),
                           ptr(Di), ptr(info))
        return Up, Ui[:nnz], Ux[:nnz], Lx[:nnz], Di, {"sweeps": int(info[0]), "nl_res0": info[1], "nl_res": info[2]}

    def fgpilu_solve(self, Up, Ui, Ux, Lx, UDinv, x):
        x = np.ascontiguousarray(x, np.float64).copy()
        self.lib.or_fgpilu_solve(len(Up) - 1, ptr(Up), ptr(Ui), ptr(Ux), ptr(Lx), ptr(UDinv), ptr(x))
        return x

    def gmres_ilu_ir(self, Ap, Ai, Ax, b, uf=np.float64, ilu_k=0, max_sweeps=10, ir_iter=10, gmres_iter=10,
                     tol=1e-10):
        n = len(Ap) - 1
        x = np.zeros(n)
        info = np.zeros(10)
        rc = self.lib.or_gmres_ilu_ir(n, ptr(Ap), ptr(Ai), ptr(np.ascontiguousarray(Ax, np.float64)),
                                      0 if uf == np.float64 else 1, ilu_k, max_sweeps, ir_iter, gmres_iter, tol,
                                      ptr(np.ascontiguousarray(b, np.float64)), ptr(x), ptr(info))
        keys = ["ir_iterations", "gmres_iterations", "stop_reason", "factor_ret", "precond", "r_norm", "factor_s",
                "solve_s", "sweeps", "nl_res"]
        return rc, x, dict(zip(keys, info.tolist()))

    def scaling(self, Ap, Ai, Ax, uf=np.float64):
        """GmresLDLIR::Compute's diagonal scaling: (AD, scaled values)."""
        n = len(Ap) - 1
        code = {np.float64: 0, np.float32: 1, np.float16: 2, "bf16": 3}[uf]
        AD = np.zeros(n)
        Axs = np.zeros(len(Ai))
        self.lib.or_scaling(n, ptr(Ap), ptr(Ai), ptr(np.ascontiguousarray(Ax, np.float64)), code, ptr(AD), ptr(Axs))
        return AD, Axs

    def gmres_ir(self, Ap, Ai, Ax, b, perm=None, uf=np.float64, ir_iter=10, gmres_iter=10, tol=1e-10):
        n = len(Ap) - 1
        x = np.zeros(n)
        info = np.zeros(8)
        pm = None if perm is None else np.ascontiguousarray(perm, np.int32)
        rc = self.lib.or_gmres_ir(n, ptr(Ap), ptr(Ai), ptr(np.ascontiguousarray(Ax, np.float64)), ptr(pm),
                                  0 if uf == np.float64 else 1, ir_iter, gmres_iter, tol,
                                  ptr(np.ascontiguousarray(b, np.float64)), ptr(x), ptr(info))
        keys = ["ir_iterations", "gmres_iterations", "stop_reason", "factor_ret", "precond", "r_norm",
                "factor_s", "solve_s"]
        return rc, x, dict(zip(keys, info.tolist()))


class Reference:
    """The reference's own qdldl.hpp / ops.hpp compiled as shipped."""

    @classmethod
    def load(cls):
        if not REF_SO.exists():
            return None
        return cls()

    def __init__(self):
        self.lib = C.CDLL(str(REF_SO))
        _bind(self.lib, {
            "ref_etree_i32": (i32, [i32, P, P, P, P, P]),
            "ref_factor_f64_i32": (i32, [i32, P, P, P, P, P, P, P, P, P, P]),
            "ref_factor_f32_i32": (i32, [i32, P, P, P, P, P, P, P, P, P, P]),
            "ref_solve_f64_f64": (None, [i32, P, P, P, P, P]),
            "ref_solve_f32_f64": (None, [i32, P, P, P, P, P]),
            "ref_solve_f32_f32": (None, [i32, P, P, P, P, P]),
            "ref_matmul_upper": (None, [i32, P, P, P, P, P]),
            "ref_dot": (f64, [i64, P, P]),
            "ref_dnrm2": (f64, [i64, P]),
            "ref_infnrm": (f64, [i64, P]),
        })

    def etree(self, Ap, Ai):
        n = len(Ap) - 1
        w, Lnz, et = (np.zeros(n, np.int32) for _ in range(3))
        tot = self.lib.ref_etree_i32(n, ptr(Ap), ptr(Ai), ptr(w), ptr(Lnz), ptr(et))
        return int(tot), Lnz, et

    def factor(self, Ap, Ai, Ax, dtype=np.float64):
        n = len(Ap) - 1
        tot, Lnz, et = self.etree(Ap, Ai)
        if tot < 0:
            return tot, None
        Lp = np.zeros(n + 1, np.int32)
        Li = np.zeros(max(tot, 1), np.int32)
        Lx = np.zeros(max(tot, 1), dtype)
        D = np.zeros(n, dtype)
        Dinv = np.zeros(n, dtype)
        ax = np.ascontiguousarray(Ax, dtype=dtype)
        fn = self.lib.ref_factor_f64_i32 if dtype == np.float64 else self.lib.ref_factor_f32_i32
        ret = fn(n, ptr(Ap), ptr(Ai), ptr(ax), ptr(Lp), ptr(Li), ptr(Lx), ptr(D), ptr(Dinv), ptr(Lnz), ptr(et))
        return int(ret), dict(Lp=Lp, Li=Li[:tot], Lx=Lx[:tot], D=D, Dinv=Dinv, Lnz=Lnz, etree=et)

    def solve(self, Lp, Li, Lx, Dinv, b, xdtype=np.float64):
        n = len(Lp) - 1
        x = np.array(b, dtype=xdtype)
        if Lx.dtype == np.float64:
            self.lib.ref_solve_f64_f64(n, ptr(Lp), ptr(Li), ptr(Lx), ptr(Dinv), ptr(x))
        elif xdtype == np.float64:
            self.lib.ref_solve_f32_f64(n, ptr(Lp), ptr(Li), ptr(Lx), ptr(Dinv), ptr(x))
        else:
            self.lib.ref_solve_f32_f32(n, ptr(Lp), ptr(Li), ptr(Lx), ptr(Dinv), ptr(x))
        return x

    def matmul(self, Ap, Ai, Ax, x):
        y = np.zeros(len(Ap) - 1)
        self.lib.ref_matmul_upper(len(Ap) - 1, ptr(Ap), ptr(Ai), ptr(np.ascontiguousarray(Ax, np.float64)),
                                  ptr(np.ascontiguousarray(x, np.float64)), ptr(y))
        return y

    def dot(self, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        return self.lib.ref_dot(len(a), ptr(a), ptr(b))

    def dnrm2(self, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.lib.ref_dnrm2(len(x), ptr(x))

    def infnrm(self, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.lib.ref_infnrm(len(x), ptr(x))


def permute_upper(Ap, Ai, Ax, perm):
    """C = triu(P K P^T) in upper CSC (perm: new -> old), pure NumPy."""
    n = len(Ap) - 1
    pinv = np.empty(n, np.int64)
    pinv[np.asarray(perm)] = np.arange(n)
    cols = np.repeat(np.arange(n), np.diff(Ap))
    a = pinv[np.asarray(Ai)]
    b = pinv[cols]
    col = np.maximum(a, b)
    row = np.minimum(a, b)
    order = np.lexsort((row, col))
    Cp = np.zeros(n + 1, np.int32)
    np.add.at(Cp, col + 1, 1)
    Cp = np.cumsum(Cp).astype(np.int32)
    return Cp, row[order].astype(np.int32), np.asarray(Ax, np.float64)[order]
