"""The oracle (oracle/mpir_oracle.c) pinned before anything is compared with it.

* golden vectors written by tests/golden/make_golden.py from the reference's
  own qdldl.hpp / ops.hpp (bit-exact);
* the compiled reference headers on fresh random inputs (bit-exact);
* ops.hpp's doctest known answers (include/MPIR/ops.hpp:63-351);
* brute-force symbolic elimination, dense reconstruction and dense solves;
* GMRES-IR closed forms: K = I, and an exact FP64 factor (one Krylov step).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from tests.oracle_lib import permute_upper

GOLDEN = Path(__file__).resolve().parent / "golden"
QDLDL_CASES = ["qdldl_small_natural", "qdldl_small_amd", "qdldl_k6_amd"]


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def test_fixture_manifest():
    import hashlib
    man = json.loads((GOLDEN / "manifest.json").read_text())
    for fname, digest in man["files"].items():
        assert hashlib.sha256((GOLDEN / fname).read_bytes()).hexdigest() == digest, fname


@pytest.mark.parametrize("case", QDLDL_CASES)
def test_etree_matches_reference_golden(oracle, case):
    g = load(case)
    tot, Lnz, et = oracle.etree(g["Ap"], g["Ai"])
    assert tot == int(g["etree_sum"])
    np.testing.assert_array_equal(Lnz, g["Lnz"])
    np.testing.assert_array_equal(et, g["etree"])


@pytest.mark.parametrize("case", QDLDL_CASES)
@pytest.mark.parametrize("tag,dt", [("f64", np.float64), ("f32", np.float32)])
def test_factor_and_solve_match_reference_golden(oracle, case, tag, dt):
    g = load(case)
    ret, F = oracle.factor(g["Ap"], g["Ai"], g["Ax"], dt)
    assert ret == int(g[f"{tag}_ret"])
    for key in ("Lp", "Li", "Lx", "D", "Dinv"):
        np.testing.assert_array_equal(F[key], g[f"{tag}_{key}"], err_msg=key)
    x = oracle.solve(F["Lp"], F["Li"], F["Lx"], F["Dinv"], g["b"], np.float64)
    np.testing.assert_array_equal(x, g[f"{tag}_solve_f64"])
    if dt == np.float32:
        x32 = oracle.solve(F["Lp"], F["Li"], F["Lx"], F["Dinv"], g["b"], np.float32)
        np.testing.assert_array_equal(x32, g[f"{tag}_solve_f32"])


@pytest.mark.parametrize("case", QDLDL_CASES)
def test_matmul_matches_reference_golden(oracle, case):
    g = load(case)
    np.testing.assert_array_equal(oracle.matmul(g["Ap"], g["Ai"], g["Ax"], g["b"]), g["matmul"])


def test_ops_match_reference_golden(oracle):
    g = load("ops_vectors")
    for key in ("v_rand", "w_rand", "v_odd", "w_odd", "v_big", "v_tiny", "v_mixed"):
        assert oracle.dnrm2(g[key]) == g[f"dnrm2_{key}"], key
        assert oracle.infnrm(g[key]) == g[f"infnrm_{key}"], key
    assert oracle.dot(g["v_rand"], g["w_rand"]) == g["dot_rand"]
    assert oracle.dot(g["v_odd"], g["w_odd"]) == g["dot_odd"]


def test_ops_doctest_known_answers(oracle):
    # include/MPIR/ops.hpp:202-223 (VectorDot), 289-316 (Dnrm2), 331-351 (InfNrm)
    assert oracle.dot([1.1, 2.2, 3.3], [4.4, 5.5, 6.6]) == pytest.approx(38.72)
    assert oracle.dot([1.5, 2.5, 3.5], [4.5, 5.5, 6.5]) == pytest.approx(43.25)
    assert oracle.dot([], []) == 0.0
    assert oracle.dnrm2([3.0, 4.0]) == pytest.approx(5.0)
    assert oracle.dnrm2([1.0, 2.0, 2.0]) == pytest.approx(3.0)
    assert oracle.dnrm2([0.0, 0.0, 0.0]) == 0.0
    assert oracle.dnrm2([5.5]) == pytest.approx(5.5)
    v = [3.14, 2.71, 1.41]
    assert oracle.dnrm2(v) == pytest.approx(np.sqrt(np.dot(v, v)))
    assert oracle.dnrm2([]) == 0.0
    assert oracle.infnrm([3.0, -4.0, 5.5, -6.1]) == pytest.approx(6.1)
    assert oracle.infnrm([-1.1, -2.2, -3.3]) == pytest.approx(3.3)
    assert oracle.infnrm([0.0, 0.0, 0.0]) == 0.0
    assert oracle.infnrm([7.5]) == pytest.approx(7.5)
    assert oracle.infnrm([-3.14, 2.71, -1.41, 5.89]) == pytest.approx(5.89)
    # scaled sum of squares survives values whose squares overflow / underflow
    assert oracle.dnrm2([3e200, 4e200]) == pytest.approx(5e200)
    assert oracle.dnrm2([3e-200, 4e-200]) == pytest.approx(5e-200)


def random_qd(n1, m, density, seed):
    """Random quasi-definite upper CSC: SPD H (diagonally dominant), random A, -D."""
    rng = np.random.default_rng(seed)
    n = n1 + m
    K = np.zeros((n, n))
    H = (rng.random((n1, n1)) < density) * rng.standard_normal((n1, n1))
    H = np.triu(H, 1)
    H = H + H.T
    H += np.diag(np.abs(H).sum(1) + 1.0)
    K[:n1, :n1] = H
    A = (rng.random((m, n1)) < density) * rng.standard_normal((m, n1))
    K[n1:, :n1] = A
    K[:n1, n1:] = A.T
    K[n1:, n1:] = -np.diag(rng.uniform(0.1, 1.0, m))
    U = np.triu(K)
    Ap = np.zeros(n + 1, np.int32)
    rows, vals = [], []
    for j in range(n):
        r = np.nonzero(U[:, j])[0]
        rows.extend(r)
        vals.extend(U[r, j])
        Ap[j + 1] = len(rows)
    return K, Ap, np.array(rows, np.int32), np.array(vals)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_compiled_reference_random(oracle, ref, seed):
    K, Ap, Ai, Ax = random_qd(20 + 3 * seed, 7 + seed, 0.15, seed)
    assert oracle.etree(Ap, Ai)[0] == ref.etree(Ap, Ai)[0]
    for dt in (np.float64, np.float32):
        r1, F1 = oracle.factor(Ap, Ai, Ax, dt)
        r2, F2 = ref.factor(Ap, Ai, Ax, dt)
        assert r1 == r2
        for key in ("Lp", "Li", "Lx", "D", "Dinv"):
            np.testing.assert_array_equal(F1[key], F2[key])
        b = np.random.default_rng(seed + 100).uniform(-1, 1, len(Ap) - 1)
        np.testing.assert_array_equal(oracle.solve(F1["Lp"], F1["Li"], F1["Lx"], F1["Dinv"], b),
                                      ref.solve(F2["Lp"], F2["Li"], F2["Lx"], F2["Dinv"], b))
        np.testing.assert_array_equal(oracle.matmul(Ap, Ai, Ax, b), ref.matmul(Ap, Ai, Ax, b))
    v = np.random.default_rng(seed).standard_normal(37 + seed)
    w = np.random.default_rng(seed + 1).standard_normal(37 + seed)
    assert oracle.dot(v, w) == ref.dot(v, w)
    assert oracle.dnrm2(v) == ref.dnrm2(v)


@pytest.mark.parametrize("seed", range(4))
def test_etree_brute_force(oracle, seed):
    # boolean elimination on the dense pattern: L pattern, counts, parents
    K, Ap, Ai, Ax = random_qd(12, 5, 0.2, 50 + seed)
    n = K.shape[0]
    pat = K != 0
    for j in range(n):
        rows = [i for i in range(j + 1, n) if pat[i, j]]
        for a in rows:
            for b in rows:
                pat[a, b] = True
    Lnz = np.array([np.count_nonzero(pat[j + 1:, j]) for j in range(n)])
    parent = np.array([(j + 1 + np.argmax(pat[j + 1:, j])) if Lnz[j] else -1 for j in range(n)])
    tot, oLnz, oet = oracle.etree(Ap, Ai)
    np.testing.assert_array_equal(oLnz, Lnz)
    np.testing.assert_array_equal(oet, parent)
    assert tot == Lnz.sum()


@pytest.mark.parametrize("seed", range(4))
def test_factor_reconstructs_and_solves(oracle, seed):
    K, Ap, Ai, Ax = random_qd(25, 9, 0.2, 80 + seed)
    n = K.shape[0]
    ret, F = oracle.factor(Ap, Ai, Ax, np.float64)
    assert ret == 25  # inertia of a quasi-definite matrix: n1 positive pivots
    L = np.eye(n)
    for j in range(n):
        for q in range(F["Lp"][j], F["Lp"][j + 1]):
            L[F["Li"][q], j] = F["Lx"][q]
    err = np.linalg.norm(K - L @ np.diag(F["D"]) @ L.T) / np.linalg.norm(K)
    assert err < 1e-14
    b = np.random.default_rng(seed).uniform(-1, 1, n)
    x = oracle.solve(F["Lp"], F["Li"], F["Lx"], F["Dinv"], b)
    np.testing.assert_allclose(x, np.linalg.solve(K, b), rtol=1e-11, atol=1e-12)


def test_zero_pivot_and_bad_input(oracle):
    # [[1, 1], [1, 1]] has an exact zero pivot
    ret, _ = oracle.factor(np.array([0, 1, 3], np.int32), np.array([0, 0, 1], np.int32), np.array([1.0, 1.0, 1.0]))
    assert ret == -1
    # empty column / entry below the diagonal
    assert oracle.etree(np.array([0, 1, 1], np.int32), np.array([0], np.int32))[0] == -1
    assert oracle.etree(np.array([0, 2, 3], np.int32), np.array([0, 1, 1], np.int32))[0] == -1


def identity_upper(n):
    return np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n)


def test_gmres_identity_closed_form(oracle):
    Ap, Ai, Ax = identity_upper(9)
    b = np.linspace(-1, 1, 9)
    rc, x, info = oracle.gmres_ir(Ap, Ai, Ax, b)
    assert rc == 0
    np.testing.assert_array_equal(x, b)
    # every PrecondGmres call ends after one Arnoldi step (happy breakdown):
    # applications of M^{-1} = one initial solve + one per step, per call
    assert info["gmres_iterations"] == info["precond"] / 2
    assert info["gmres_iterations"] <= 2
    assert info["stop_reason"] in (1, 3)  # small update or exactly zero residual


def test_gmres_exact_factor_one_step(oracle):
    # FP64 factor of K: M^{-1} K = I to rounding, every PrecondGmres call stops after one step
    K, Ap, Ai, Ax = random_qd(30, 10, 0.15, 7)
    b = K @ np.random.default_rng(1).uniform(-1, 1, K.shape[0])
    rc, x, info = oracle.gmres_ir(Ap, Ai, Ax, b, uf=np.float64, tol=1e-10)
    assert rc == 0
    calls = info["precond"]
    # calls = 1 initial solve per PrecondGmres + 1 per Arnoldi step
    assert info["gmres_iterations"] == calls / 2
    eta = np.abs(b - K @ x).max() / (np.abs(K).sum(1).max() * np.abs(x).max() + np.abs(b).max())
    assert eta < 1e-15


def test_gmres_ir_cfg0_fp32_factor(oracle):
    import mpir_amd
    P = mpir_amd.kkt_config(0)
    perm = mpir_amd.amd_order(P.Ap, P.Ai)
    rc, x, info = oracle.gmres_ir(P.Ap, P.Ai, P.Ax, P.b, perm=perm, uf=np.float32, tol=1e-12)
    assert rc == 0 and info["factor_ret"] == P.n1
    y = oracle.matmul(P.Ap, P.Ai, P.Ax, x)
    knorm = np.abs(oracle.matmul(P.Ap, P.Ai, np.abs(P.Ax), np.ones(P.n))).max()
    eta = np.abs(P.b - y).max() / (knorm * np.abs(x).max() + np.abs(P.b).max())
    assert eta < 1e-14
    assert np.abs(x - P.x_true).max() < 1e-9


def test_permute_helper_matches_reference_pattern(oracle):
    import mpir_amd
    P = mpir_amd.kkt_config(0)
    g = load("cfg0_symbolic")
    Cp, Ci, Cx = permute_upper(P.Ap, P.Ai, P.Ax, g["perm"])
    tot, Lnz, et = oracle.etree(Cp, Ci)
    assert tot == int(g["etree_sum"])
    np.testing.assert_array_equal(Lnz, g["Lnz"])
    np.testing.assert_array_equal(et, g["etree"])


def test_luir_identity_closed_form(oracle):
    # K = I: x0 = FACTPREC(b); each correction is the residual rounded to FP32,
    # so x reaches b exactly after a few steps and the zero residual ends
    # the loop (the reference would divide by it)
    Ap, Ai, Ax = identity_upper(7)
    b = np.linspace(-1, 1, 7) / 3.0
    ret, x, steps, rr = oracle.luir(Ap, Ai, Ax, b, np.float32)
    assert ret == 7
    np.testing.assert_array_equal(x, b)
    assert 2 <= steps <= 4 and rr == 0.0
    # FP64 factor: x0 is already exact, the first residual is zero
    ret, x, steps, rr = oracle.luir(Ap, Ai, Ax, b, np.float64)
    np.testing.assert_array_equal(x, b)
    assert steps == 1 and rr == 0.0


def test_luir_exact_factor_and_maxiter_zero(oracle):
    # an FP64 factor: corrections are at the rounding level of x, so the
    # ||d||/||x|| <= 10 eps test (or the 0.5 ratio) ends the loop within two
    # steps; maxiter = 0 is a single solve in double
    K, Ap, Ai, Ax = random_qd(30, 10, 0.15, 11)
    b = K @ np.random.default_rng(2).uniform(-1, 1, K.shape[0])
    ret, x, steps, rr = oracle.luir(Ap, Ai, Ax, b, np.float64)
    assert ret == 30 and 1 <= steps <= 2 and rr < 1e-14
    ret, x0, steps0, rr0 = oracle.luir(Ap, Ai, Ax, b, np.float64, maxiter=0)
    assert steps0 == 0 and rr0 < 1e-13
    F = oracle.factor(Ap, Ai, Ax, np.float64)[1]
    np.testing.assert_array_equal(x0, oracle.solve(F["Lp"], F["Li"], F["Lx"], F["Dinv"], b))


def test_luir_fp32_factor_refines_to_fp64(oracle):
    # the stopping rules on a system whose FP32 factor is a contraction: the
    # corrections shrink geometrically until the 0.5 ratio or 10 eps test
    import mpir_amd
    P = mpir_amd.kkt_config(0)
    Cp, Ci, Cx = permute_upper(P.Ap, P.Ai, P.Ax, mpir_amd.amd_order(P.Ap, P.Ai))
    y = np.random.default_rng(3).uniform(-1, 1, P.n)
    b = oracle.matmul(Cp, Ci, Cx, y)
    ret, x, steps, rr = oracle.luir(Cp, Ci, Cx, b, np.float32, maxiter=20)
    assert ret == P.n1
    assert 2 <= steps < 20 and rr < 1e-12
    np.testing.assert_allclose(x, y, rtol=0, atol=1e-9)


# ---------------------------------------------------------------------- FGPILU
def ilu_dense(n, Up, Ui, Ux, Lx):
    """Dense L (rows of L on the indices of U's columns, mirrored) and U."""
    L, U = np.zeros((n, n)), np.zeros((n, n))
    for c in range(n):
        for q in range(Up[c], Up[c + 1]):
            U[Ui[q], c] = Ux[q]
            L[c, Ui[q]] = Lx[q]
    return L, U


def ilu_pattern(n, Up, Ui):
    pat = np.zeros((n, n), bool)
    for c in range(n):
        pat[Ui[Up[c]:Up[c + 1]], c] = True
    return pat


def upper_csc(K):
    n = K.shape[0]
    Ap, rows, vals = np.zeros(n + 1, np.int32), [], []
    for j in range(n):
        r = np.nonzero(np.triu(K)[:, j])[0]
        rows.extend(r)
        vals.extend(K[r, j])
        Ap[j + 1] = len(rows)
    return Ap, np.array(rows, np.int32), np.array(vals)


def laplacian_2d(m):
    T = 4 * np.eye(m) - np.eye(m, k=1) - np.eye(m, k=-1)
    return np.kron(np.eye(m), T) - np.kron(np.eye(m, k=1) + np.eye(m, k=-1), np.eye(m))


def test_fgpilu_pattern_is_a_boolean_power(oracle):
    # Compute's pattern loop (gmres_ir.hpp:369-408): column col of the next
    # pattern = rows of A's upper columns row1 for the rows row1 of column col,
    # i.e. the level-k pattern is the structure of B^(k+1), B = triu(A) != 0
    K, Ap, Ai, Ax = random_qd(20, 8, 0.15, 3)
    B = (np.triu(K) != 0).astype(np.int64)
    n = K.shape[0]
    for k in range(4):
        Up, Ui, Ux, Lx, Di, info = oracle.fgpilu(Ap, Ai, Ax, ilu_k=k, max_sweeps=0)
        pat = np.zeros((n, n), bool)
        for c in range(n):
            rows = Ui[Up[c]:Up[c + 1]]
            assert np.all(np.diff(rows) > 0)
            pat[rows, c] = True
        np.testing.assert_array_equal(pat, np.linalg.matrix_power(B, k + 1) > 0)
        # no sweeps: the initial guesses are A on the pattern
        assert info["sweeps"] == 0
        Ld, Ud = ilu_dense(n, Up, Ui, Ux, Lx)
        np.testing.assert_array_equal(Ud, np.where(pat, np.triu(K), 0))
        np.testing.assert_array_equal(Di, 1 / np.diag(K))


def test_fgpilu_initial_residual_closed_form(oracle):
    # NLResNorm (:801-824) of the initial guesses: the sum over the pattern
    # (both triangles) of |A - L U| with L = U = A there, L's diagonal included
    K, Ap, Ai, Ax = random_qd(15, 6, 0.2, 4)
    n = K.shape[0]
    Up, Ui, Ux, Lx, Di, info = oracle.fgpilu(Ap, Ai, Ax, ilu_k=1, max_sweeps=0)
    Ld, Ud = ilu_dense(n, Up, Ui, Ux, Lx)
    pat = ilu_pattern(n, Up, Ui)
    assert np.count_nonzero(pat) > np.count_nonzero(np.triu(K))  # fill entries count too
    R = np.abs(K - Ld @ Ud)
    expect = R[pat].sum() + R.T[pat & ~np.eye(n, dtype=bool)].sum()
    assert info["nl_res0"] == pytest.approx(expect, rel=1e-13)


def test_fgpilu_tridiagonal_is_exact_ldlt(oracle):
    # no fill: ILU(0) = the exact factorization, from the textbook recurrence
    # d_1 = a_1, l_{i+1} = b_i / d_i, d_{i+1} = a_{i+1} - l_{i+1} b_i
    n = 14
    a = 4 + 0.1 * np.arange(n)
    b = -1 + 0.05 * np.arange(n - 1)
    K = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    Ap, Ai, Ax = upper_csc(K)
    d, l = np.zeros(n), np.zeros(n)
    d[0] = a[0]
    for i in range(n - 1):
        l[i + 1] = b[i] / d[i]
        d[i + 1] = a[i + 1] - l[i + 1] * b[i]
    Up, Ui, Ux, Lx, Di, info = oracle.fgpilu(Ap, Ai, Ax, ilu_k=0, max_sweeps=10, tol=1e-10)
    Ld, Ud = ilu_dense(n, Up, Ui, Ux, Lx)
    np.testing.assert_allclose(np.diag(Ud), d, rtol=1e-15)
    np.testing.assert_allclose(np.diag(Ud, 1), b, rtol=1e-15)
    np.testing.assert_allclose(np.diag(Ld, -1), l[1:], rtol=1e-15)
    np.testing.assert_array_equal(np.diag(Ld), np.ones(n))
    np.testing.assert_allclose(Di, 1 / d, rtol=1e-15)
    # the in-order sweep is exact at once; the second changes nothing
    assert info["sweeps"] == 2 and info["nl_res"] < 1e-13


@pytest.mark.parametrize("seed", range(3))
def test_fgpilu_band_pattern_is_the_lu_factorization(oracle, seed):
    # a full band holds all of its LU fill, so ILU(k) on it is the LU
    # factorization without pivoting (unique; it exists for a strictly
    # diagonally dominant matrix): L U = K, L unit lower, U upper. The
    # diagonal is positive in the first 10 rows, negative after, as in a
    # quasi-definite matrix.
    rng = np.random.default_rng(20 + seed)
    n, w = 15, 3
    K = np.zeros((n, n))
    for d in range(1, w + 1):
        v = rng.uniform(-1, 1, n - d)
        K += np.diag(v, d) + np.diag(v, -d)
    K += np.diag(np.where(np.arange(n) < 10, 1.0, -1.0) * (np.abs(K).sum(1) + 0.5))
    Ap, Ai, Ax = upper_csc(K)
    for k in (0, 2):
        Up, Ui, Ux, Lx, Di, info = oracle.fgpilu(Ap, Ai, Ax, ilu_k=k, max_sweeps=3)
        Ld, Ud = ilu_dense(n, Up, Ui, Ux, Lx)
        np.testing.assert_allclose(Ld @ Ud, K, rtol=0, atol=1e-14 * np.abs(K).max())
        np.testing.assert_array_equal(np.diag(Ld), np.ones(n))
        np.testing.assert_allclose(Di, 1 / np.diag(Ud), rtol=1e-15)
        assert np.count_nonzero(Di > 0) == 10


def test_fgpilu_pattern_misses_fill_below_the_diagonal_path(oracle):
    # the pattern follows increasing paths only: an arrow matrix (row 0 full)
    # fills completely in LU, but ILU(k) keeps its upper pattern for every k
    n = 8
    K = np.eye(n) * 4
    K[0, 1:] = K[1:, 0] = 1
    Ap, Ai, Ax = upper_csc(K)
    for k in (0, 3, n):
        Up, Ui, *_ = oracle.fgpilu(Ap, Ai, Ax, ilu_k=k, max_sweeps=0)
        np.testing.assert_array_equal(ilu_pattern(n, Up, Ui), np.triu(K) != 0)


def test_fgpilu_ilu0_matches_a_on_the_pattern_only(oracle):
    # ILU(0) of the 2-D Laplacian: (L U)_ij = A_ij on the pattern, dropped fill elsewhere
    K = laplacian_2d(8)
    n = K.shape[0]
    Ap, Ai, Ax = upper_csc(K)
    Up, Ui, Ux, Lx, Di, info = oracle.fgpilu(Ap, Ai, Ax, ilu_k=0, max_sweeps=10)
    Ld, Ud = ilu_dense(n, Up, Ui, Ux, Lx)
    pat = (K != 0)
    E = Ld @ Ud - K
    assert np.abs(E[pat]).max() < 1e-14
    assert np.abs(E[~pat]).max() > 1e-2
    # M^{-1} = U^{-1} L^{-1} (TriangularSolve, :743-774) against dense solves
    x = np.random.default_rng(5).uniform(-1, 1, n)
    y = oracle.fgpilu_solve(Up, Ui, Ux, Lx, Di, x)
    np.testing.assert_allclose(y, np.linalg.solve(Ud, np.linalg.solve(Ld, x)), rtol=1e-13, atol=1e-14)


def test_fgpilu_fp32_factor_and_sweep_limit(oracle):
    # float factors round the FP64 ones; MAX_SWEEPS = 1 stops after one sweep
    K = laplacian_2d(6)
    Ap, Ai, Ax = upper_csc(K)
    r64 = oracle.fgpilu(Ap, Ai, Ax, uf=np.float64, ilu_k=1, max_sweeps=10)
    r32 = oracle.fgpilu(Ap, Ai, Ax, uf=np.float32, ilu_k=1, max_sweeps=1)
    assert r32[5]["sweeps"] == 1
    np.testing.assert_allclose(r32[2], r64[2], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(r32[3], r64[3], rtol=1e-6, atol=1e-6)
    assert np.all(r32[2] == r32[2].astype(np.float32))


def test_gmres_ilu_ir_cfg0(oracle):
    # the reference's shipped path (Compute with ILU_K = 0, example.cpp:144-147)
    # on the small KKT system: FP32 ILU(0) preconditioned GMRES-IR
    import mpir_amd
    P = mpir_amd.kkt_config(0)
    rc, x, info = oracle.gmres_ilu_ir(P.Ap, P.Ai, P.Ax, P.b, uf=np.float32, ilu_k=0, max_sweeps=5, tol=1e-12)
    assert rc == 0 and info["sweeps"] == 2
    assert info["precond"] == info["gmres_iterations"] + info["ir_iterations"] + 1
    assert np.abs(x - P.x_true).max() < 1e-6
