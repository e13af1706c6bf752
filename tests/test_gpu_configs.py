"""Every workload at its stated size on the GPU.

* 64^3 KKT (n = 327 680), FP32 and FP64 factors: backward error, inertia
  (n1 positive pivots, QDLDL_factor's return for a quasi-definite K) and the
  forward error against x_true (b = K x_true).
* Its precision form: D in {1e-2, 1e-4, 1e-6}, FP16 and BF16 factors with
  diagonal or Ruiz scaling. FP32 and FP64 factors converge at every D; a
  16-bit factor either converges or says that it did not (stop reason 0, a
  reported residual above tolerance, or MPIR_ERR_NUMERIC for a factor
  outside the format's range) -- it never reports a solution it does not
  have. Forward errors are printed per leg.
* The batched form: 512 systems x 8 right-hand sides, every pair checked.

The oracle is not run here: its FP32 factorization of the 64^3 system takes
more than a minute on one core. These tests check what is independent of
the elimination order: K x = b itself.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import mpir_amd

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
TOL = 1e-10


def full_k(Ap, Ai, Ax):
    n = len(Ap) - 1
    U = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    return (U + sp.triu(U, 1).T).tocsr()


def errors(K, x, b, x_true):
    r = b - K @ x
    eta = np.abs(r).max() / (abs(K).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    fwd = np.abs(x - x_true).max() / np.abs(x_true).max()
    return float(eta), float(fwd)


@pytest.fixture(scope="module")
def cube():
    return mpir_amd.kkt_config(1)


@pytest.mark.parametrize("uf", [np.float32, np.float64])
def test_64cubed(gpu, cube, uf):
    P = cube
    s = mpir_amd.GmresLDLIR(uf, ordering=mpir_amd.ORDER_ND)
    s.SetTolerance(TOL)
    s.Compute(P.Ap, P.Ai, P.Ax)
    x = s.Solve(P.b)
    st = s.stats()
    eta, fwd = errors(full_k(P.Ap, P.Ai, P.Ax), x, P.b, P.x_true)
    print(f"64^3 {np.dtype(uf).name}: eta {eta:.2e} fwd {fwd:.2e} ir {st['ir_iterations']} "
          f"gmres {st['gmres_iterations']} stop {st['stop_reason']} factor {st['factor_ms']:.1f} ms", flush=True)
    assert st["n"] == 327680 and st["positive_pivots"] == P.n1
    assert st["stop_reason"] in (1, 2, 3)
    assert eta <= TOL and eta < 1e-14
    assert fwd < 1e-8


@pytest.mark.parametrize("d_value", [1e-2, 1e-4, 1e-6])
def test_precision_form_64cubed(gpu, d_value):
    P = mpir_amd.kkt_config(2, d_value=d_value)
    K = full_k(P.Ap, P.Ai, P.Ax)
    legs = [(np.float64, mpir_amd.SCALE_DIAG), (np.float32, mpir_amd.SCALE_DIAG)]
    legs += [(uf, sc) for uf in (np.float16, "bf16") for sc in (mpir_amd.SCALE_DIAG, mpir_amd.SCALE_RUIZ)]
    for uf, scaling in legs:
        name = f"D={d_value:g} {uf if isinstance(uf, str) else np.dtype(uf).name} " \
               f"{'ruiz' if scaling == mpir_amd.SCALE_RUIZ else 'diag'}"
        s = mpir_amd.GmresLDLIR(uf, ordering=mpir_amd.ORDER_ND, scaling=scaling)
        s.SetTolerance(TOL)
        sixteen = uf in (np.float16, "bf16")
        if sixteen:
            s.SetMaxGmresIterations(50)
            s.SetMaxIRIterations(20)
        try:
            s.Compute(P.Ap, P.Ai, P.Ax)
        except mpir_amd.MpirError as e:
            # a factor entry outside the 16-bit range: reported, not hidden
            print(f"{name}: {e}", flush=True)
            assert sixteen and e.code == -104
            continue
        x = s.Solve(P.b)
        st = s.stats()
        eta, fwd = errors(K, x, P.b, P.x_true)
        print(f"{name}: eta {eta:.2e} fwd {fwd:.2e} ir {st['ir_iterations']} gmres {st['gmres_iterations']} "
              f"stop {st['stop_reason']} reported {st['final_rel_residual']:.2e} pivots+ {st['positive_pivots']} "
              f"(n1 {P.n1})", flush=True)
        if not sixteen:
            assert st["positive_pivots"] == P.n1
            assert eta < 1e-13 and fwd < 1e-6
        elif eta > TOL:
            # not converged: the solver's own residual says so
            assert st["stop_reason"] == 0 or st["final_rel_residual"] > TOL


def test_batched_512x8(gpu):
    systems = mpir_amd.batch_config()
    assert len(systems) == 512 and all(q.B.shape[1] == 8 for q in systems)
    s = mpir_amd.BatchLDLIR(np.float32)
    s.SetTolerance(TOL)
    s.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax) for q in systems])
    xs = s.Solve([q.B for q in systems])
    st = s.stats()
    worst_eta = worst_fwd = 0.0
    for i, q in enumerate(systems):
        P = q.kkt
        assert st["positive_pivots"][i] == P.n1
        K = full_k(P.Ap, P.Ai, P.Ax)
        for r in range(8):
            eta, fwd = errors(K, xs[i][:, r], q.B[:, r], q.X_true[:, r])
            worst_eta, worst_fwd = max(worst_eta, eta), max(worst_fwd, fwd)
            assert st["stop_reason"][i, r] in (1, 2, 3)
    print(f"512 x 8: worst eta {worst_eta:.2e} worst fwd {worst_fwd:.2e} sum n {sum(q.kkt.n for q in systems)}",
          flush=True)
    assert worst_eta < 1e-13 and worst_fwd < 1e-6
