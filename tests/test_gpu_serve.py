"""The serving engine on the GPU: continuous batching of a request stream.

Bars:
* a request's solution is bit-identical alone and in a mixed stream (any
  batch, any lane, any neighbours);
* every request of a stream reaches a normwise backward error < 1e-13
  against its own values and right-hand sides;
* admission keeps the bytes in flight within the budget, runs at most
  `lanes` batches at once, and admits no request before it arrives.
"""
import numpy as np
import pytest

import mpir_amd

pytestmark = pytest.mark.gpu


def backward_error(Ap, Ai, Ax, x, b):
    import scipy.sparse as sp
    n = len(Ap) - 1
    U = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    K = (U + sp.triu(U, 1).T).tocsr()
    r = b - K @ x
    kn = abs(K).sum(axis=1).max()
    return float(np.abs(r).max() / (kn * np.abs(x).max() + np.abs(b).max()))


@pytest.fixture(scope="module")
def pool():
    return mpir_amd.serving_patterns(count=6, k_lo=6, k_hi=14, seed=1)


def make_server(pool, **kw):
    srv = mpir_amd.Server(np.float32, tol=1e-12, keep_x=True, **kw)
    for P in pool:
        srv.add_pattern(P.Ap, P.Ai, P.Ax, P.n1)
    return srv


def test_server_stream_solves_every_request(gpu, pool):
    srv = make_server(pool, lanes=2, max_batch=8)
    t, pat, nr, seeds = mpir_amd.poisson_stream(2000.0, 0.02, len(pool), seed=4)
    out = srv.run(pat, seeds, nr, t)
    summ = srv.summary()
    print(f"requests {len(t)} batches {summ['batches']} max backlog {summ['max_backlog']} "
          f"concurrent {summ['max_concurrent']}", flush=True)
    assert summ["completed"] == len(t) and summ["failed"] == 0
    assert (out["status"] == 0).all()
    assert (out["t_admit"] >= out["t_arrive"]).all() and (out["t_done"] >= out["t_admit"]).all()
    assert summ["max_concurrent"] <= 2
    for i in range(len(t)):
        Ax, B, X = srv.request_data(int(pat[i]), int(seeds[i]), int(nr[i]))
        P = pool[pat[i]]
        x = srv.solution(i)
        assert x.shape == B.shape
        for c in range(B.shape[1]):
            assert backward_error(P.Ap, P.Ai, Ax, x[:, c], B[:, c]) < 1e-13


def test_server_alone_equals_mix_bitwise(gpu, pool):
    t, pat, nr, seeds = mpir_amd.poisson_stream(4000.0, 0.01, len(pool), seed=9)
    assert len(t) >= 20
    mix = make_server(pool, lanes=3, max_batch=5)
    mix.run(pat, seeds, nr, t)
    assert mix.summary()["batches"] >= 4
    alone = make_server(pool, lanes=1, max_batch=1)
    for i in (0, 3, len(t) // 2, len(t) - 1):
        alone.run(pat[i:i + 1], seeds[i:i + 1], nr[i:i + 1], np.zeros(1))
        np.testing.assert_array_equal(alone.solution(0), mix.solution(i))


def test_server_admission_by_footprint(gpu, pool):
    t, pat, nr, seeds = mpir_amd.poisson_stream(3000.0, 0.02, len(pool), seed=2)
    # a budget of a few requests' footprint forces small batches
    budget = 40 << 20
    srv = make_server(pool, lanes=2, max_batch=64, budget_bytes=budget)
    out = srv.run(pat, seeds, nr, t)
    summ = srv.summary()
    assert summ["completed"] == len(t)
    assert summ["peak_inflight_bytes"] <= budget or summ["batches"] == len(t)
    assert summ["batches"] > 2
    # a batch's members were all admitted at once and finish together
    for b in np.unique(out["batch"]):
        sel = out["batch"] == b
        assert np.ptp(out["t_admit"][sel]) == 0 and np.ptp(out["t_done"][sel]) == 0
