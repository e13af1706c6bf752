"""Tunings of the factor and solve schedules, and the autotuner, on the GPU.

Bars:
* under every tuning the autotuner can choose (and the extremes of the
  ranges), the FP64 factor has the oracle's pattern and inertia exactly and
  its values within 1e-9; the DAG and level schedules agree bit for bit and
  GMRES-IR reaches backward error < 1e-13;
* a tuning set after the factorization re-plans the solves of the factor in
  place, and the solution stays within the tolerance;
* Autotune returns one of its candidates, with factor and M^-1 times no
  worse than those of the tuning in place before the call, and leaves a
  factor whose solves are those of a handle set to the same tuning, bit for
  bit; the batched form solves every system after tuning.
"""
import numpy as np
import pytest

import mpir_amd
from tests.oracle_lib import permute_upper

pytestmark = pytest.mark.gpu


def backward_error(oracle, Ap, Ai, Ax, x, b):
    r = b - oracle.matmul(Ap, Ai, Ax, x)
    knorm = np.abs(oracle.matmul(Ap, Ai, np.abs(Ax), np.ones(len(b)))).max()
    return np.abs(r).max() / (knorm * np.abs(x).max() + np.abs(b).max())


def oracle_factor(oracle, P, perm):
    # the reference's Compute scaling (gmres_ir.hpp:347-363), then QDLDL_factor
    ad = np.ones(P.n)
    for j in range(P.n):
        q = P.Ap[j + 1] - 1
        if P.Ai[q] == j and abs(P.Ax[q]) > np.finfo(np.float64).eps:
            ad[j] = 1.0 / np.sqrt(abs(P.Ax[q]))
    cols = np.repeat(np.arange(P.n), np.diff(P.Ap))
    Cp, Ci, Cx = permute_upper(P.Ap, P.Ai, P.Ax * (ad[P.Ai] * ad[cols]), perm)
    return oracle.factor(Cp, Ci, Cx, np.float64)


@pytest.fixture(scope="module")
def cube24(oracle):
    P = mpir_amd.kkt_config(1, k=24)
    s = mpir_amd.GmresLDLIR(np.float64, ordering=mpir_amd.ORDER_ND)
    s.Analyze(P.Ap, P.Ai)
    ret, F = oracle_factor(oracle, P, s.perm())
    return P, ret, F


# (update grouping, super-block edge, narrow-front rows)
TUNINGS = [(4, 8, 1024), (1, 1, 0), (2, 4, 256), (8, 16, 4096), (3, 2, 64), (64, 64, 1 << 30)]


@pytest.mark.parametrize("tuning", TUNINGS)
def test_every_tuning_matches_oracle(gpu, oracle, cube24, tuning):
    P, ret, F = cube24
    s = mpir_amd.GmresLDLIR(np.float64, ordering=mpir_amd.ORDER_ND)
    s.SetTuning(*tuning)
    s.SetTolerance(1e-12)
    s.Compute(P.Ap, P.Ai, P.Ax)
    Lp, Li, Lx, D, _ = s.export_factor()
    assert ret == P.n1 == s.stats()["positive_pivots"]
    np.testing.assert_array_equal(Lp, F["Lp"])
    np.testing.assert_array_equal(Li, F["Li"])
    np.testing.assert_allclose(Lx, F["Lx"], rtol=1e-9, atol=1e-11 * np.abs(F["Lx"]).max())
    np.testing.assert_allclose(D, F["D"], rtol=1e-10)
    x_dag = s.Solve(P.b)
    s.SetSolveSchedule(mpir_amd.SCHED_LEVELS)
    x_lev = s.Solve(P.b)
    np.testing.assert_array_equal(x_dag, x_lev)
    assert backward_error(oracle, P.Ap, P.Ai, P.Ax, x_dag, P.b) < 1e-13


def test_tuning_after_factorization(gpu, oracle):
    P = mpir_amd.kkt_config(1, k=16)
    s = mpir_amd.GmresLDLIR(np.float32, ordering=mpir_amd.ORDER_ND)
    s.SetTolerance(1e-12)
    s.Compute(P.Ap, P.Ai, P.Ax)
    x0 = s.Solve(P.b)
    for tuning in [(2, 4, 0), (8, 16, 4096), (4, 8, 256)]:
        s.SetTuning(*tuning)  # the solves of the same factor, re-planned
        x = s.Solve(P.b)
        assert backward_error(oracle, P.Ap, P.Ai, P.Ax, x, P.b) < 1e-13
        np.testing.assert_allclose(x, x0, rtol=0, atol=1e-9 * np.abs(x0).max())
    with pytest.raises(mpir_amd.MpirError):
        s.SetTuning(0, 8, 1024)
    with pytest.raises(mpir_amd.MpirError):
        s.SetTuning(4, 65, 1024)


def test_autotune_single_system(gpu, oracle):
    P = mpir_amd.kkt_config(1, k=20)
    s = mpir_amd.GmresLDLIR(np.float32, ordering=mpir_amd.ORDER_ND)
    s.SetTolerance(1e-12)
    s.Compute(P.Ap, P.Ai, P.Ax)
    x0 = s.Solve(P.b)
    r = s.Autotune(reps=2)
    print(r, flush=True)
    assert r["upd_group"] in (2, 4, 8) and r["super_tiles"] in (4, 8, 16)
    assert r["narrow_rows"] in (0, 256, 512, 1024, 2048, 4096)
    assert r["schedule"] in (mpir_amd.SCHED_LEVELS, mpir_amd.SCHED_DAG)
    assert r["candidates"] == 9 + 12
    assert 0 < r["factor_ms"] <= r["factor_ms_before"] and 0 < r["precond_ms"] <= r["precond_ms_before"]
    x1 = s.Solve(P.b)
    assert backward_error(oracle, P.Ap, P.Ai, P.Ax, x1, P.b) < 1e-13
    np.testing.assert_allclose(x1, x0, rtol=0, atol=1e-9 * np.abs(x0).max())
    # the tuned handle computes what a handle set to the same tuning computes
    t = mpir_amd.GmresLDLIR(np.float32, ordering=mpir_amd.ORDER_ND)
    t.SetTolerance(1e-12)
    t.SetTuning(r["upd_group"], r["super_tiles"], r["narrow_rows"])
    t.SetSolveSchedule(r["schedule"])
    t.Compute(P.Ap, P.Ai, P.Ax)
    np.testing.assert_array_equal(t.Solve(P.b), x1)
    # a new analysis of the tuned handle keeps the choice
    s.Compute(P.Ap, P.Ai, P.Ax)
    np.testing.assert_array_equal(s.Solve(P.b), x1)


def test_autotune_batch(gpu, oracle):
    systems = mpir_amd.batch_config(nsys=6, nrhs=2, n1_lo=500, n1_hi=4000)
    b = mpir_amd.BatchLDLIR(np.float32)
    b.SetTolerance(1e-12)
    b.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax) for q in systems])
    r = b.Autotune(reps=1)
    print(r, flush=True)
    assert r["candidates"] == 21 and r["factor_ms"] <= r["factor_ms_before"]
    xs = b.Solve([q.B for q in systems])
    for i, q in enumerate(systems):
        P = q.kkt
        for c in range(2):
            assert backward_error(oracle, P.Ap, P.Ai, P.Ax, xs[i][:, c], q.B[:, c]) < 1e-13
