"""The GPU's control layer of GmresLDLIR::Solve / PrecondGmres and Compute's
scaling, held to the same known answers as the oracle (tests/solve_rules.py),
to the oracle's IR count and stop reason, and the failure behaviour the
reference leaves silent: non-finite inputs and factors end with
MPIR_ERR_NUMERIC instead of a quiet NaN or zero solution."""
import numpy as np
import pytest

import mpir_amd
from tests import solve_rules as R

pytestmark = pytest.mark.gpu
ERR_NUMERIC = -104


def backward_error(oracle, Ap, Ai, Ax, x, b):
    r = b - oracle.matmul(Ap, Ai, Ax, x)
    knorm = np.abs(oracle.matmul(Ap, Ai, np.abs(Ax), np.ones(len(b)))).max()
    return np.abs(r).max() / (knorm * np.abs(x).max() + np.abs(b).max())


@pytest.mark.parametrize("tol", sorted(R.IDENTITY_TRACES))
def test_identity_traces(gpu, tol):
    x_is_b, ir, gm, apps, reason = R.IDENTITY_TRACES[tol]
    Ap, Ai, Ax = R.identity_upper(R.IDENTITY_N)
    b = np.full(R.IDENTITY_N, R.IDENTITY_B)
    s = mpir_amd.GmresLDLIR(np.float64)
    s.SetTolerance(tol)
    s.Compute(Ap, Ai, Ax)
    x = s.Solve(b)
    st = s.stats()
    np.testing.assert_array_equal(x, b if x_is_b else np.zeros_like(b))
    got = (st["ir_iterations"], st["gmres_iterations"], st["precond_applications"], st["stop_reason"])
    assert got == (ir, gm, apps, reason), st


@pytest.mark.parametrize("uf", [np.float64, np.float32])
def test_scaling_threshold_through_the_factor(gpu, uf):
    # D of the factor of K_s = AD K AD in the given order: closed forms of the
    # eps(UF) threshold (row 1 couples to row 0: D_1 = a11 s1^2 - (a01 s0 s1)^2)
    Ap, Ai, Ax = R.scaling_matrix()
    ad = np.asarray(R.SCALING_AD[uf])
    s = mpir_amd.GmresLDLIR(uf, ordering=mpir_amd.ORDER_NATURAL)
    s.Compute(Ap, Ai, Ax)
    _, _, _, D, _ = s.export_factor()
    diag = R.SCALING_DIAG * ad * ad
    expect = diag.copy()
    a01 = R.SCALING_A01 * ad[0] * ad[1]
    expect[1] = diag[1] - a01 * a01 / diag[0]
    rtol = 1e-14 if uf == np.float64 else 2e-7
    np.testing.assert_allclose(D, expect, rtol=rtol)


def test_stagnation_stop_is_absolute(gpu):
    P = mpir_amd.kkt_config(0)
    s = mpir_amd.GmresLDLIR(np.float32, ordering=mpir_amd.ORDER_AMD)
    s.SetTolerance(R.STAGNATION_TOL)
    s.Compute(P.Ap, P.Ai, P.Ax)
    s.Solve(P.b)
    st = s.stats()
    assert (st["ir_iterations"], st["stop_reason"]) == (R.STAGNATION_IR, R.STAGNATION_REASON), st


def test_gmres_limit_clamp_persists(gpu, oracle):
    # gmres_iter_ = min(gmres_iter_, n) is stored (gmres_ir.hpp:643): after a
    # 5 x 5 system the handle keeps at most 5 Arnoldi steps per call, also
    # for a larger system computed afterwards
    from tests.test_oracle import random_qd
    K, Ap, Ai, Ax = random_qd(3, 2, 0.6, 4)
    b = K @ np.linspace(-1, 1, 5)
    s = mpir_amd.GmresLDLIR(np.float32)
    s.SetTolerance(0.0)
    s.SetMaxIRIterations(2)
    s.SetMaxGmresIterations(50)
    s.Compute(Ap, Ai, Ax)
    s.Solve(b)
    st = s.stats()
    rc, _, info = oracle.gmres_ir(Ap, Ai, Ax, b, uf=np.float32, tol=0.0, ir_iter=2, gmres_iter=50)
    assert st["gmres_iterations"] == info["gmres_iterations"] == 5 * (st["ir_iterations"] + 1)
    P = mpir_amd.kkt_config(0)
    s.Compute(P.Ap, P.Ai, P.Ax)
    s.Solve(P.b)
    st = s.stats()
    assert st["gmres_iterations"] == 5 * (st["ir_iterations"] + 1), st


@pytest.mark.parametrize("cfg,k,uf,order", [(0, None, np.float64, 1), (1, 12, np.float64, 2), (1, 16, np.float64, 1),
                                            (0, None, np.float32, 1), (1, 12, np.float32, 2), (1, 16, np.float32, 1)])
@pytest.mark.parametrize("tol", [1e-10, 1e-12])
def test_ir_count_and_stop_reason_match_oracle(gpu, oracle, cfg, k, uf, order, tol):
    # the GPU's Arnoldi basis comes from CGS2 (the reference: MGS with a
    # conditional second pass) and its factor from another elimination
    # order, so FP32 trajectories may differ by one refinement step; FP64
    # factors must agree exactly
    P = mpir_amd.kkt_config(cfg, k=k)
    s = mpir_amd.GmresLDLIR(uf, ordering=order)
    s.SetTolerance(tol)
    s.Compute(P.Ap, P.Ai, P.Ax)
    x = s.Solve(P.b)
    st = s.stats()
    rc, xo, info = oracle.gmres_ir(P.Ap, P.Ai, P.Ax, P.b, perm=s.perm(), uf=uf, tol=tol)
    assert rc == 0
    if uf == np.float64:
        assert st["ir_iterations"] == info["ir_iterations"], (st, info)
        if st["stop_reason"] != info["stop_reason"]:
            # a tie at the last IR iteration: the small-update rule (1) holds on
            # one side and not on the other because the two factors round
            # differently; then the GPU's ||d||_inf must sit at the rule's
            # threshold 10 eps ||x_old||_inf (scaled space, x_s = x / AD)
            assert {st["stop_reason"], info["stop_reason"]} == {1, 2}, (st, info)
            AD, _ = oracle.scaling(P.Ap, P.Ai, P.Ax, uf)
            d_last = s.log()["events"][0]["iterations"][-1]["d_norm"]
            ratio = d_last / (10 * np.finfo(np.float64).eps * np.abs(x / AD).max())
            print(f"stop-rule tie: GPU {st['stop_reason']} oracle {info['stop_reason']}, ||d|| / threshold = {ratio:.3f}")
            assert 0.25 <= ratio <= 4.0, ratio
    else:
        assert abs(st["ir_iterations"] - info["ir_iterations"]) <= 1, (st, info)
        assert st["stop_reason"] in (1, 2, 3) and info["stop_reason"] in (1, 2, 3)
        if st["ir_iterations"] == info["ir_iterations"]:
            assert st["stop_reason"] == info["stop_reason"], (st, info)
    assert backward_error(oracle, P.Ap, P.Ai, P.Ax, x, P.b) < 1e-13


@pytest.mark.parametrize("max_ir,reason", [(1, 0), (2, 1)])
def test_iteration_limit_and_small_update_match_oracle(gpu, oracle, max_ir, reason):
    # tol = 0: nothing stagnates and every call runs all 10 Arnoldi steps;
    # after one IR iteration the correction is still above 10 eps ||x||
    # (stop reason 0 at the limit), after the second it is not (reason 1)
    P = mpir_amd.kkt_config(0)
    s = mpir_amd.GmresLDLIR(np.float32, ordering=mpir_amd.ORDER_AMD)
    s.SetTolerance(0.0)
    s.SetMaxIRIterations(max_ir)
    s.Compute(P.Ap, P.Ai, P.Ax)
    s.Solve(P.b)
    st = s.stats()
    rc, _, info = oracle.gmres_ir(P.Ap, P.Ai, P.Ax, P.b, perm=s.perm(), uf=np.float32, tol=0.0, ir_iter=max_ir)
    assert (st["ir_iterations"], st["stop_reason"]) == (max_ir, reason) == (info["ir_iterations"], info["stop_reason"])
    assert st["gmres_iterations"] == info["gmres_iterations"] == 10 * (max_ir + 1)


def _raises_numeric(fn):
    with pytest.raises(mpir_amd.MpirError) as e:
        fn()
    assert e.value.code == ERR_NUMERIC, str(e.value)


@pytest.mark.parametrize("where", [0, 700, 1279])
def test_nan_in_b_is_reported(gpu, where):
    P = mpir_amd.kkt_config(0)
    s = mpir_amd.GmresLDLIR(np.float32)
    s.Compute(P.Ap, P.Ai, P.Ax)
    b = P.b.copy()
    b[where] = np.nan
    _raises_numeric(lambda: s.Solve(b))
    b[where] = np.inf
    _raises_numeric(lambda: s.Solve(b))
    # the handle stays usable
    x = s.Solve(P.b)
    assert np.abs(x - P.x_true).max() < 1e-8


@pytest.mark.parametrize("uf", [np.float64, np.float32])
@pytest.mark.parametrize("diag", [False, True])
def test_nonfinite_matrix_is_reported(gpu, uf, diag):
    P = mpir_amd.kkt_config(0)
    Ax = P.Ax.copy()
    q = P.Ap[300 + 1] - 1 if diag else P.Ap[300]  # the diagonal of column 300, or its first entry
    Ax[q] = np.inf
    s = mpir_amd.GmresLDLIR(uf)
    _raises_numeric(lambda: s.Compute(P.Ap, P.Ai, Ax))
    Ax[q] = np.nan
    _raises_numeric(lambda: s.Compute(P.Ap, P.Ai, Ax))


def test_fp32_factor_overflow_is_reported(gpu):
    # values near the FP32 limit: the scaled matrix is fine, but with the
    # scaling switched off the FP32 factor overflows
    P = mpir_amd.kkt_config(0)
    s = mpir_amd.GmresLDLIR(np.float32, scaling=mpir_amd.SCALE_NONE)
    _raises_numeric(lambda: s.Compute(P.Ap, P.Ai, P.Ax * 1e38))
    s = mpir_amd.GmresLDLIR(np.float32)
    s.Compute(P.Ap, P.Ai, P.Ax * 1e38)
    x = s.Solve(P.b * 1e38)
    assert np.abs(x - P.x_true).max() < 1e-8


def test_huge_right_hand_side_converges(gpu, oracle):
    # Dnrm2 is a scaled sum of squares (ops.hpp:231-255): a right-hand side
    # near 1e200, whose squares overflow, solves like any other
    P = mpir_amd.kkt_config(0)
    s = mpir_amd.GmresLDLIR(np.float32)
    s.SetTolerance(1e-12)
    s.Compute(P.Ap, P.Ai, P.Ax)
    for f in (1e200, 1e-200):
        x = s.Solve(P.b * f)
        np.testing.assert_allclose(x / f, P.x_true, rtol=0, atol=1e-9)
        assert s.stats()["stop_reason"] in (1, 2)


def test_indefinite_h_inertia_matches_oracle(gpu, oracle):
    # H shifted to be indefinite: no longer quasi-definite, the factor (no
    # pivoting, as QDLDL) has positive pivots != n1; the GPU's count equals
    # the oracle's on the same scaled, permuted matrix
    from tests.oracle_lib import permute_upper
    P = mpir_amd.gen_kkt(3, 8, -2.0, 128, 1, 4, 4, 1e-2, 0)
    s = mpir_amd.GmresLDLIR(np.float64, ordering=mpir_amd.ORDER_AMD)
    s.Compute(P.Ap, P.Ai, P.Ax)
    AD, Axs = oracle.scaling(P.Ap, P.Ai, P.Ax, np.float64)
    Cp, Ci, Cx = permute_upper(P.Ap, P.Ai, Axs, s.perm())
    ret, _ = oracle.factor(Cp, Ci, Cx, np.float64)
    assert ret == s.stats()["positive_pivots"] != P.n1
    x = s.Solve(P.b)
    rc, xo, info = oracle.gmres_ir(P.Ap, P.Ai, P.Ax, P.b, perm=s.perm(), uf=np.float64, tol=1e-10)
    assert rc == 0
    assert backward_error(oracle, P.Ap, P.Ai, P.Ax, x, P.b) < 1e-13
    assert (s.stats()["ir_iterations"], s.stats()["stop_reason"]) == (info["ir_iterations"], info["stop_reason"])
