"""FGPILU, the reference's shipped preconditioner (GmresLDLIR::Compute,
include/MPIR/gmres_ir.hpp:323-531), on the GPU against the oracle.

Bars:
* the ILU(k) pattern and the factors U, L, 1/U_jj, in FP64 and FP32, are
  bit-identical to the oracle's (the reference's single-threaded sweeps) on
  the same diagonally scaled matrix;
* the sweep count equals the oracle's; the nonlinear residuals agree to
  1e-10 of the initial one (their sums run in another order);
* M^{-1} = U^{-1} L^{-1} on the GPU is bit-identical to the oracle's
  TriangularSolve on the exported factors, one vector or several at once;
* GMRES-IR with it: the oracle's IR iteration count and stop reason, and a
  backward error within 10x of the oracle's.
"""
import numpy as np
import pytest

import mpir_amd

pytestmark = pytest.mark.gpu

TOL = 1e-12
SWEEPS = 5


def scaled_values(P, uf):
    """The reference's Compute scaling (gmres_ir.hpp:347-363): s_i = 1/sqrt|a_ii|
    where |a_ii| > eps(UF), a_ij s_i s_j."""
    eps = np.finfo(np.float64 if uf == np.float64 else np.float32).eps
    ad = np.ones(P.n)
    for j in range(P.n):
        q = P.Ap[j + 1] - 1
        if P.Ai[q] == j and abs(P.Ax[q]) > eps:
            ad[j] = 1.0 / np.sqrt(abs(P.Ax[q]))
    cols = np.repeat(np.arange(P.n), np.diff(P.Ap))
    return P.Ax * (ad[P.Ai] * ad[cols])


def fgpilu_solver(uf, ilu_k, P):
    s = mpir_amd.GmresLDLIR(uf, preconditioner=mpir_amd.PREC_FGPILU, MAX_SWEEPS=SWEEPS)
    s.SetTolerance(TOL)
    s.Compute(P.Ap, P.Ai, P.Ax, ILU_k=ilu_k)
    return s


@pytest.mark.parametrize("cfg,k,uf,ilu_k", [(0, None, np.float32, 0), (0, None, np.float64, 1),
                                            (1, 12, np.float32, 0), (1, 12, np.float64, 2), (1, 16, np.float32, 1)])
def test_fgpilu_factors_bitwise(gpu, oracle, cfg, k, uf, ilu_k):
    P = mpir_amd.kkt_config(cfg, k=k)
    print(f"cfg {cfg} k {k} uf {np.dtype(uf).name} ILU({ilu_k}) n {P.n}", flush=True)
    s = fgpilu_solver(uf, ilu_k, P)
    Up, Ui, Ux, Lx, Di = s.export_ilu()
    oUp, oUi, oUx, oLx, oDi, info = oracle.fgpilu(P.Ap, P.Ai, scaled_values(P, uf), uf=uf, ilu_k=ilu_k,
                                                  max_sweeps=SWEEPS, tol=TOL)
    np.testing.assert_array_equal(Up, oUp)
    np.testing.assert_array_equal(Ui, oUi)
    np.testing.assert_array_equal(Ux, oUx)
    np.testing.assert_array_equal(Lx, oLx)
    np.testing.assert_array_equal(Di, oDi)
    st = s.stats()
    assert st["preconditioner"] == mpir_amd.PREC_FGPILU
    assert st["nnz_l"] == len(oUi)
    assert st["ilu_sweeps"] == info["sweeps"]
    assert abs(st["ilu_nl_residual0"] - info["nl_res0"]) <= 1e-10 * info["nl_res0"]
    assert abs(st["ilu_nl_residual"] - info["nl_res"]) <= 1e-10 * info["nl_res0"]
    assert st["positive_pivots"] == np.count_nonzero(oDi > 0)


@pytest.mark.parametrize("uf", [np.float32, np.float64])
def test_fgpilu_preconditioner_bitwise(gpu, oracle, uf):
    import torch
    P = mpir_amd.kkt_config(1, k=12)
    s = fgpilu_solver(uf, 1, P)
    Up, Ui, Ux, Lx, Di = s.export_ilu()
    lib = mpir_amd._lib.load()
    X = np.random.default_rng(9).uniform(-1, 1, (3, P.n))
    expect = np.stack([oracle.fgpilu_solve(Up, Ui, Ux, Lx, Di, x) for x in X])
    one = torch.tensor(X, device="cuda")
    for r in range(3):
        assert lib.mpir_gmres_ldl_precond_dev(s.handle, mpir_amd.C.c_void_p(one[r].data_ptr())) == 0
    many = torch.tensor(X, device="cuda")
    assert lib.mpir_gmres_ldl_precond_multi_dev(s.handle, mpir_amd.C.c_void_p(many.data_ptr()), 3) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(one.cpu().numpy(), expect)
    np.testing.assert_array_equal(many.cpu().numpy(), expect)


def backward_error(oracle, P, x):
    y = oracle.matmul(P.Ap, P.Ai, P.Ax, x)
    knorm = np.abs(oracle.matmul(P.Ap, P.Ai, np.abs(P.Ax), np.ones(P.n))).max()
    return np.abs(P.b - y).max() / (knorm * np.abs(x).max() + np.abs(P.b).max())


@pytest.mark.parametrize("cfg,k,uf,ilu_k", [(0, None, np.float32, 0), (1, 12, np.float32, 0),
                                            (1, 12, np.float64, 1)])
def test_fgpilu_gmres_ir_matches_oracle(gpu, oracle, cfg, k, uf, ilu_k):
    P = mpir_amd.kkt_config(cfg, k=k)
    print(f"cfg {cfg} k {k} uf {np.dtype(uf).name} ILU({ilu_k})", flush=True)
    s = fgpilu_solver(uf, ilu_k, P)
    x = s.Solve(P.b)
    st = s.stats()
    rc, xo, info = oracle.gmres_ilu_ir(P.Ap, P.Ai, P.Ax, P.b, uf=uf, ilu_k=ilu_k, max_sweeps=SWEEPS, tol=TOL)
    assert rc == 0
    assert st["ir_iterations"] == info["ir_iterations"]
    assert st["stop_reason"] == info["stop_reason"]
    eta, eta_o = backward_error(oracle, P, x), backward_error(oracle, P, xo)
    print(f"eta {eta:.3e} oracle {eta_o:.3e} ir {st['ir_iterations']} gmres {st['gmres_iterations']}", flush=True)
    assert eta <= 10 * eta_o + 1e-16


def test_fgpilu_handle_rules(gpu):
    P = mpir_amd.kkt_config(0)
    s = fgpilu_solver(np.float32, 0, P)
    with pytest.raises(mpir_amd.MpirError):
        s.export_factor()
    # ILU_k changes with Compute's argument; back to L D L^T on the same handle
    s.Compute(P.Ap, P.Ai, P.Ax, ILU_k=2)
    assert s.stats()["nnz_l"] > len(P.Ai)
    s.SetPreconditioner(mpir_amd.PREC_LDL)
    s.Compute(P.Ap, P.Ai, P.Ax)
    st = s.stats()
    assert st["preconditioner"] == mpir_amd.PREC_LDL and st["positive_pivots"] == P.n1
    with pytest.raises(mpir_amd.MpirError):
        s.SetPreconditioner(7)
    with pytest.raises(mpir_amd.MpirError):
        mpir_amd.GmresLDLIR(np.float16, preconditioner=mpir_amd.PREC_FGPILU)
