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


def caller_request(P, rng, nrhs):
    """Values and right-hand sides a caller brings: off-diagonals of H and A
    times 1 + 0.05u, diag(H) times 1.05 + 0.05v (H stays diagonally
    dominant), -D times 1 + 0.05u; B = K X_true."""
    n = P.n
    cols = np.repeat(np.arange(n), np.diff(P.Ap))
    diag = P.Ai == cols
    u = rng.uniform(-1, 1, len(P.Ax))
    f = np.where(diag & (cols < P.n1), 1.05 + 0.025 * (u + 1), 1.0 + 0.05 * u)
    Ax = P.Ax * f
    X = rng.uniform(-1, 1, (n, nrhs))
    U = sp_upper(P.Ap, P.Ai, Ax)
    B = U @ X + U.T @ X - U.diagonal()[:, None] * X
    return Ax, B, X


def sp_upper(Ap, Ai, Ax):
    import scipy.sparse as sp
    n = len(Ap) - 1
    return sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))


def test_submit_poll_caller_data_from_threads(gpu, pool):
    # three client threads submit requests with their own values and
    # right-hand sides and wait for them; every result solves its own system,
    # and equals the same request solved alone, bit for bit
    import threading
    srv = make_server(pool, lanes=3, max_batch=6)
    rng = np.random.default_rng(21)
    work = []
    for i in range(24):
        p = int(rng.integers(len(pool)))
        Ax, B, X = caller_request(pool[p], rng, int(rng.integers(1, 5)))
        work.append((p, Ax, B if B.shape[1] > 1 else B[:, 0], X))
    results = [None] * len(work)
    errors = []

    def client(c):
        try:
            ids = [(i, srv.submit(work[i][0], work[i][1], work[i][2])) for i in range(c, len(work), 3)]
            for i, rid in ids:
                done, x, info = srv.poll(rid, timeout_ms=-1)
                assert done and info["status"] == 0 and info["latency_ms"] >= info["queue_ms"] >= 0
                results[i] = (x, info)
        except BaseException as e:  # noqa: BLE001
            errors.append(e)

    ts = [threading.Thread(target=client, args=(c,)) for c in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not errors, errors
    assert srv.summary()["batches"] >= 2
    for (p, Ax, B, X), (x, info) in zip(work, results):
        P = pool[p]
        Bm, xm = B.reshape(P.n, -1), x.reshape(P.n, -1)
        assert xm.shape == Bm.shape
        for c in range(Bm.shape[1]):
            assert backward_error(P.Ap, P.Ai, Ax, xm[:, c], Bm[:, c]) < 1e-13
        assert info["rel_residual_max"] < 1e-10
    alone = make_server(pool, lanes=1, max_batch=1)
    for i in (0, 7, 23):
        p, Ax, B, _ = work[i]
        done, x, _ = alone.poll(alone.submit(p, Ax, B), timeout_ms=-1)
        np.testing.assert_array_equal(x, results[i][0])


def test_submit_rejects_bad_requests(gpu, pool):
    srv = make_server(pool, lanes=1, max_batch=2)
    P = pool[0]
    with pytest.raises(ValueError):
        srv.submit(0, P.Ax[:-1], P.b)
    with pytest.raises(ValueError):
        srv.submit(len(pool) + 3, P.Ax, P.b)
    # a request whose right-hand side holds a NaN fails alone; the other
    # request of the same batch is solved
    b = P.b.copy()
    b[3] = np.nan
    bad = srv.submit(0, P.Ax, b)
    good = srv.submit(0, P.Ax, P.b)
    with pytest.raises(mpir_amd.MpirError):
        srv.poll(bad, timeout_ms=-1)
    done, x, info = srv.poll(good, timeout_ms=-1)
    assert done and backward_error(P.Ap, P.Ai, P.Ax, x, P.b) < 1e-13


@pytest.mark.slow
def test_serving_pool_of_64_patterns(gpu):
    # the serving workload's pool: 3D k^3 Laplacians, k = 16 .. 48
    # (n = 5 120 .. 138 240), a short Poisson stream at 2000 requests/s
    pool = mpir_amd.serving_patterns()
    assert len(pool) == 64 and pool[0].n == 5120 and pool[-1].n == 138240
    srv = mpir_amd.Server(np.float32, lanes=3, max_batch=64, tol=1e-10, keep_x=True)
    srv.add_patterns(pool, threads=16)
    t, pat, nr, seeds = mpir_amd.poisson_stream(2000.0, 0.25, len(pool), seed=5)
    out = srv.run(pat, seeds, nr, t)
    summ = srv.summary()
    lat = out["latency_ms"]
    print(f"{len(t)} requests, {summ['batches']} batches, {len(t) / summ['wall_ms'] * 1e3:.0f} req/s, "
          f"p50 {np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms", flush=True)
    assert summ["completed"] == len(t) and summ["failed"] == 0
    assert (out["status"] == 0).all() and (out["rel_residual_max"] <= 1e-10).all()
    # every pattern size class, checked against its own values
    for i in range(0, len(t), max(1, len(t) // 40)):
        Ax, B, X = srv.request_data(int(pat[i]), int(seeds[i]), int(nr[i]))
        P = pool[pat[i]]
        x = srv.solution(i)
        for c in range(B.shape[1]):
            assert backward_error(P.Ap, P.Ai, Ax, x[:, c], B[:, c]) < 1e-13
    alone = mpir_amd.Server(np.float32, lanes=1, max_batch=1, tol=1e-10, keep_x=True)
    alone.add_patterns(pool, threads=16)
    for i in (0, len(t) // 2, len(t) - 1):
        alone.run(pat[i:i + 1], seeds[i:i + 1], nr[i:i + 1], np.zeros(1))
        np.testing.assert_array_equal(alone.solution(0), srv.solution(i))
