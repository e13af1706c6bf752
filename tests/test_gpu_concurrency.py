"""Several host threads on one GPU at once.

* Handles solving concurrently with the dependency-driven (persistent)
  triangular solves: every solve completes and equals the same handle's
  result alone, bit for bit. The persistent launches of different streams
  share the device; a task class that no resident workgroup serves would
  leave the others waiting (a bounded wait, reported as MPIR_ERR_INTERNAL).
* The QDLDL drop-ins are reentrant, as the reference's functions are
  (qdldl.hpp:22-23): threads call QDLDL_solve at once on their own factors.
  A thread that solves again with the same factor pattern reuses its plan.
"""
import threading

import numpy as np
import pytest

import mpir_amd

pytestmark = pytest.mark.gpu


def _run_threads(fns):
    errors = [None] * len(fns)
    out = [None] * len(fns)

    def wrap(i):
        try:
            out[i] = fns[i]()
        except BaseException as e:  # noqa: BLE001 - reported below
            errors[i] = e

    ts = [threading.Thread(target=wrap, args=(i,)) for i in range(len(fns))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    for e in errors:
        if e is not None:
            raise e
    return out


@pytest.mark.parametrize("uf", [np.float32, np.float64])
def test_concurrent_dag_solves_equal_solo(gpu, uf):
    sizes = [24, 28, 32, 20]
    solvers, rhs, solo = [], [], []
    for i, k in enumerate(sizes):
        P = mpir_amd.kkt_config(1, k=k, seed=i)
        s = mpir_amd.GmresLDLIR(uf, ordering=mpir_amd.ORDER_ND)
        s.SetSolveSchedule(mpir_amd.SCHED_DAG)
        s.SetTolerance(1e-12)
        s.Compute(P.Ap, P.Ai, P.Ax)
        solvers.append(s)
        rhs.append(P.b)
        solo.append(s.Solve(P.b))

    def work(i):
        return [solvers[i].Solve(rhs[i]) for _ in range(4)]

    results = _run_threads([lambda i=i: work(i) for i in range(len(sizes))])
    for i, xs in enumerate(results):
        for x in xs:
            np.testing.assert_array_equal(x, solo[i])


def test_concurrent_qdldl_solve_reentrant_and_plan_reuse(gpu, oracle):
    cases = []
    for i, k in enumerate([10, 12, 14, 16]):
        P = mpir_amd.kkt_config(1, k=k, seed=i)
        s = mpir_amd.GmresLDLIR(np.float64)
        s.Compute(P.Ap, P.Ai, P.Ax)
        Lp, Li, Lx, D, Dinv = s.export_factor()
        b = np.random.default_rng(i).uniform(-1, 1, P.n)
        xo = oracle.solve(Lp.astype(np.int32), Li, Lx, Dinv, b)
        cases.append((P.n, Lp, Li.astype(np.int64), Lx, Dinv, b, xo))

    def work(i):
        n, Lp, Li, Lx, Dinv, b, xo = cases[i]
        builds0 = mpir_amd.qdldl_solve_plan_builds()
        xs = []
        for rep in range(3):
            # the second and third calls: same pattern, values scaled
            f = 1.0 + rep
            x = b.copy()
            mpir_amd.QDLDL_solve(n, Lp, Li, Lx, Dinv * f, x)
            xs.append((x, xo * f))
        return xs, mpir_amd.qdldl_solve_plan_builds() - builds0

    out = _run_threads([lambda i=i: work(i) for i in range(len(cases))])
    for xs, builds in out:
        assert builds == 1  # one plan per thread, reused by its later calls
        for x, xo in xs:
            np.testing.assert_allclose(x, xo, rtol=1e-10, atol=1e-10 * np.abs(xo).max())
