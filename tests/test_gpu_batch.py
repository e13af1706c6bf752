"""Batches of independent systems on the GPU (through the C ABI).

Bars:
* a system's solution, iteration counts and stop reason are bit-identical
  alone (a batch of one) and inside any batch, at any position, next to any
  other systems and right-hand-side counts;
* every (system, right-hand side) reaches the normwise backward error the
  single-system GmresLDLIR reaches (< 1e-13 here), and agrees with the
  oracle's GMRES-IR solution on the same permutation within 1e-8;
* the inertia of every factor is the reference's QDLDL_factor return value
  for a quasi-definite K (n1 positive pivots).
"""
import numpy as np
import pytest

import mpir_amd

pytestmark = pytest.mark.gpu


def backward_error(oracle, P, x, b):
    r = b - oracle.matmul(P.Ap, P.Ai, P.Ax, x)
    knorm = np.abs(oracle.matmul(P.Ap, P.Ai, np.abs(P.Ax), np.ones(P.n))).max()
    return np.abs(r).max() / (knorm * np.abs(x).max() + np.abs(b).max())


def small_batch(nsys=5, nrhs=3, seed=11):
    return mpir_amd.batch_config(nsys=nsys, nrhs=nrhs, seed=seed, n1_lo=200, n1_hi=3000)


def solve(systems, rhs_counts, uf=np.float32, tol=1e-12):
    s = mpir_amd.BatchLDLIR(uf)
    s.SetTolerance(tol)
    s.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax) for q in systems])
    rhs = [None if c == 0 else q.B[:, :c] for q, c in zip(systems, rhs_counts)]
    xs = s.Solve(rhs)
    return s, xs, s.stats()


def test_batch_solves_every_system(gpu, oracle):
    systems = small_batch()
    counts = [3, 1, 2, 3, 1]
    s, xs, st = solve(systems, counts)
    for i, (q, c) in enumerate(zip(systems, counts)):
        P = q.kkt
        assert st["positive_pivots"][i] == P.n1
        for r in range(c):
            print(f"system {i} (n={P.n}) rhs {r}: ir {st['ir_iterations'][i, r]} gmres {st['gmres_iterations'][i, r]}",
                  flush=True)
            assert backward_error(oracle, P, xs[i][:, r], q.B[:, r]) < 1e-13
            assert st["stop_reason"][i, r] in (1, 2, 3)
        np.testing.assert_allclose(xs[i], q.X_true[:, :c], rtol=0, atol=1e-6 * np.abs(q.X_true).max())


def test_batch_member_equals_alone_bitwise(gpu):
    systems = small_batch(nsys=6, seed=5)
    counts = [2, 3, 1, 2, 0, 3]
    _, xs, st = solve(systems, counts)
    # each system alone (a batch of one), with its own right-hand-side count
    for i in (0, 1, 3, 5):
        _, xa, sa = solve([systems[i]], [counts[i]])
        np.testing.assert_array_equal(xa[0], xs[i])
        c = counts[i]
        np.testing.assert_array_equal(sa["ir_iterations"][0, :c], st["ir_iterations"][i, :c])
        np.testing.assert_array_equal(sa["gmres_iterations"][0, :c], st["gmres_iterations"][i, :c])
        np.testing.assert_array_equal(sa["stop_reason"][0, :c], st["stop_reason"][i, :c])
    # a different batch: reversed order, other neighbours, more slots
    order = [5, 3, 1, 0]
    _, xr, _ = solve([systems[i] for i in order], [3, 2, 3, 2])
    for j, i in enumerate(order):
        c = counts[i]
        np.testing.assert_array_equal(xr[j][:, :c], xs[i])
    assert xs[4] is None


def test_batch_split_panel_steps_match_bitwise(gpu):
    # the batch's factorization with every panel-step launch split gives the
    # solutions of single launches bit for bit
    systems = small_batch(nsys=6, seed=7)
    counts = [2, 1, 3, 1, 2, 1]
    rhs = [q.B[:, :c] for q, c in zip(systems, counts)]
    s = mpir_amd.BatchLDLIR(np.float32)
    s.SetTolerance(1e-12)
    s.SetPanelSplit(0)
    s.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax) for q in systems])
    x0 = s.Solve(rhs)
    s.SetPanelSplit(1)
    s.Factor([q.kkt.Ax for q in systems])
    x1 = s.Solve(rhs)
    for a, b in zip(x0, x1):
        np.testing.assert_array_equal(a, b)


def test_batch_matches_single_system_solver(gpu, oracle):
    systems = small_batch(nsys=3, nrhs=1, seed=3)
    _, xs, st = solve(systems, [1, 1, 1], tol=1e-10)
    for i, q in enumerate(systems):
        P = q.kkt
        g = mpir_amd.GmresLDLIR(np.float32, ordering=mpir_amd.ORDER_ND)
        g.SetTolerance(1e-10)
        g.Compute(P.Ap, P.Ai, P.Ax)
        x1 = g.Solve(q.B[:, 0])
        rc, xo, info = oracle.gmres_ir(P.Ap, P.Ai, P.Ax, q.B[:, 0], perm=g.perm(), uf=np.float32, tol=1e-10)
        assert rc == 0
        scale = np.abs(xo).max()
        assert np.abs(xs[i][:, 0] - xo).max() / scale < 1e-8
        assert np.abs(x1 - xo).max() / scale < 1e-8
        assert backward_error(oracle, P, xs[i][:, 0], q.B[:, 0]) < 1e-13


def test_batch_fp64_factor_and_cache(gpu, oracle):
    systems = small_batch(nsys=4, nrhs=2, seed=8)
    before = mpir_amd.pattern_cache_info()
    # the same patterns twice in one batch, and again in a second batch
    twice = systems + systems
    _, xs, st = solve(twice, [2] * 8, uf=np.float64)
    after = mpir_amd.pattern_cache_info()
    assert after["entries"] - before["entries"] == 4
    for i in range(4):
        np.testing.assert_array_equal(xs[i], xs[i + 4])
        P = systems[i].kkt
        for r in range(2):
            assert backward_error(oracle, P, xs[i][:, r], systems[i].B[:, r]) < 1e-13
    _, _, _ = solve(systems[:2], [1, 1], uf=np.float64)
    final = mpir_amd.pattern_cache_info()
    assert final["entries"] == after["entries"] and final["hits"] >= after["hits"] + 2


def test_batch_zero_pivot_names_the_system(gpu):
    systems = small_batch(nsys=3, nrhs=1, seed=2)
    s = mpir_amd.BatchLDLIR(np.float64)
    P = systems[1].kkt
    Ax = P.Ax.copy()
    # an all-zero K block: the first column's diagonal and its couplings
    Ax[P.Ap[0]:P.Ap[1]] = 0.0
    for j in range(P.n):
        for q in range(P.Ap[j], P.Ap[j + 1]):
            if P.Ai[q] == 0:
                Ax[q] = 0.0
    with pytest.raises(mpir_amd.MpirError, match="system 1"):
        s.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax if i != 1 else Ax) for i, q in enumerate(systems)])


def test_batch_bad_input_rejected(gpu):
    systems = small_batch(nsys=2, nrhs=1, seed=4)
    s = mpir_amd.BatchLDLIR(np.float32)
    s.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax) for q in systems])
    with pytest.raises(ValueError):
        s.Solve([systems[0].B[:, 0]])
    with pytest.raises(mpir_amd.MpirError):
        s.Solve([None, None])


@pytest.mark.parametrize("uf", [np.float32, np.float64])
def test_multi_rhs_preconditioner_equals_single_bitwise(gpu, uf):
    # M^{-1} on several right-hand sides at once (the factor read once for up
    # to 8 of them) gives exactly the single-vector results; 3 and 5 exercise
    # the groups of 2 + 1 and 4 + 1
    import torch
    P = mpir_amd.kkt_config(1, k=20)
    s = mpir_amd.GmresLDLIR(uf, ordering=mpir_amd.ORDER_ND)
    s.Compute(P.Ap, P.Ai, P.Ax)
    lib = mpir_amd._lib.load()
    gen = torch.Generator().manual_seed(5)
    for nrhs in (8, 3, 5):
        X = torch.rand(nrhs, P.n, dtype=torch.float64, generator=gen).to("cuda")
        one = X.clone()
        for r in range(nrhs):
            assert lib.mpir_gmres_ldl_precond_dev(s.handle, mpir_amd.C.c_void_p(one[r].data_ptr())) == 0
        many = X.clone()
        assert lib.mpir_gmres_ldl_precond_multi_dev(s.handle, mpir_amd.C.c_void_p(many.data_ptr()), nrhs) == 0
        torch.cuda.synchronize()
        print(f"nrhs {nrhs}: max |x| {one.abs().max().item():.3e}", flush=True)
        assert torch.equal(one, many)


@pytest.mark.parametrize("uf,sched", [(np.float32, mpir_amd.SCHED_LEVELS), (np.float32, mpir_amd.SCHED_DAG),
                                      (np.float64, mpir_amd.SCHED_DAG)])
def test_batch_graph_replays_equal_eager_bitwise(gpu, oracle, uf, sched):
    # the whole solve as one captured HIP graph: the first solve of a kind
    # runs kernel by kernel, the next ones replay the graph; every replay
    # (with eager work of the same handle in between) gives the eager bits,
    # and so does a handle that never captures
    systems = small_batch(nsys=4, nrhs=3, seed=5)
    counts = [3, 1, 0, 2]
    rhs = [None if c == 0 else q.B[:, :c] for q, c in zip(systems, counts)]
    s = mpir_amd.BatchLDLIR(uf)
    s.SetTolerance(1e-12)
    s.SetSolveSchedule(sched)
    s.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax) for q in systems])
    x0 = s.Solve(rhs)
    st0 = s.stats()
    assert st0["graph_nodes"] == 0 and st0["arnoldi_steps"] > 0
    for rep in range(3):
        x = s.Solve(rhs)
        st = s.stats()
        assert st["graph_nodes"] > 100, st["graph_nodes"]
        for a, b in zip(x, x0):
            if a is not None:
                np.testing.assert_array_equal(a, b)
        for key in ("ir_iterations", "gmres_iterations", "stop_reason", "final_rel_residual"):
            np.testing.assert_array_equal(st[key], st0[key])
        assert st["arnoldi_steps"] == st0["arnoldi_steps"] and st["precond_sweeps"] == st0["precond_sweeps"]
        # eager work of the handle between replays: another solve shape
        s.Solve([q.B[:, :1] for q in systems])
    e = mpir_amd.BatchLDLIR(uf)
    e.SetTolerance(1e-12)
    e.SetSolveSchedule(sched)
    e.SetGraph(False)
    e.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax) for q in systems])
    for _ in range(2):
        xe = e.Solve(rhs)
        assert e.stats()["graph_nodes"] == 0
        for a, b in zip(xe, x0):
            if a is not None:
                np.testing.assert_array_equal(a, b)
    for i, (q, c) in enumerate(zip(systems, counts)):
        for r in range(c):
            assert backward_error(oracle, q.kkt, x0[i][:, r], q.B[:, r]) < 1e-13
    # new values on the same batch: the graph reads them at its next replay
    s.Compute([(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax * 1.01) for q in systems])
    s.Solve(rhs)
    x2 = s.Solve(rhs)
    assert s.stats()["graph_nodes"] > 0
    for i, (q, c) in enumerate(zip(systems, counts)):
        for r in range(c):
            P = q.kkt
            rr = q.B[:, r] - 1.01 * oracle.matmul(P.Ap, P.Ai, P.Ax, x2[i][:, r])
            kn = 1.01 * np.abs(oracle.matmul(P.Ap, P.Ai, np.abs(P.Ax), np.ones(P.n))).max()
            assert np.abs(rr).max() / (kn * np.abs(x2[i][:, r]).max() + np.abs(q.B[:, r]).max()) < 1e-13
