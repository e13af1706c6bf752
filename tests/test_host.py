"""Host-side pieces of libmpir_hip.so (no GPU work): the C ABI exports, the
workload generator, the AMD ordering and the QDLDL_etree drop-in."""
import json
import re
from pathlib import Path

import numpy as np
import pytest

import mpir_amd
from mpir_amd import _lib
from tests.oracle_lib import permute_upper

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"


def declared_symbols():
    text = (ROOT / "include" / "mpir_hip.h").read_text()
    return sorted(set(re.findall(r"\b(mpir_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_generator_is_deterministic_and_pinned(oracle):
    import hashlib
    P = mpir_amd.kkt_config(0)
    man = json.loads((GOLDEN / "manifest.json").read_text())
    h = hashlib.sha256()
    for a in (P.Ap, P.Ai, P.Ax, P.x_true, P.b):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == man["cfg0_generator_sha256"]
    assert (P.n, P.n1, int(P.Ap[-1])) == (1280, 1024, 4032)
    # b = K x_true exactly as the reference's MatrixMultiply sums it
    np.testing.assert_array_equal(oracle.matmul(P.Ap, P.Ai, P.Ax, P.x_true), P.b)


def test_generator_workload_sizes():
    P = mpir_amd.kkt_config(1, k=16)
    n1 = 16 ** 3
    assert P.n1 == n1 and P.n == n1 + n1 // 4
    # every constraint row has 4 distinct grid columns plus its -D diagonal
    cols = np.diff(P.Ap)[n1:]
    assert np.all(cols == 5)
    assert np.all(P.Ax[P.Ap[n1 + 1:] - 1] == -1e-4)


def test_amd_valid_deterministic_and_reduces_fill(oracle):
    P = mpir_amd.kkt_config(0)
    p1 = mpir_amd.amd_order(P.Ap, P.Ai)
    p2 = mpir_amd.amd_order(P.Ap, P.Ai)
    np.testing.assert_array_equal(p1, p2)
    assert sorted(p1.tolist()) == list(range(P.n))
    g = dict(np.load(GOLDEN / "cfg0_symbolic.npz", allow_pickle=False))
    np.testing.assert_array_equal(p1, g["perm"])  # own-path regression pin
    Cp, Ci, _ = permute_upper(P.Ap, P.Ai, P.Ax, p1)
    fill_amd = oracle.etree(Cp, Ci)[0]
    fill_nat = oracle.etree(P.Ap, P.Ai)[0]
    assert fill_amd < 0.25 * fill_nat


def test_amd_fill_close_to_scipy_minimum_degree(oracle):
    # SuperLU's multiple minimum degree on A^T + A as an independent yardstick
    sp = pytest.importorskip("scipy.sparse")
    spl = pytest.importorskip("scipy.sparse.linalg")
    P = mpir_amd.kkt_config(1, k=12)
    p = mpir_amd.amd_order(P.Ap, P.Ai)
    Cp, Ci, _ = permute_upper(P.Ap, P.Ai, P.Ax, p)
    ours = oracle.etree(Cp, Ci)[0]
    U = sp.csc_matrix((P.Ax, P.Ai, P.Ap), shape=(P.n, P.n))
    K = (U + sp.triu(U, 1).T).tocsc()
    lu = spl.splu(K, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))
    mmd = lu.L.nnz - P.n
    assert ours < 1.15 * mmd


@pytest.mark.parametrize("case", ["qdldl_small_natural", "qdldl_small_amd", "qdldl_k6_amd"])
def test_qdldl_etree_dropin_bit_exact(case):
    g = dict(np.load(GOLDEN / f"{case}.npz", allow_pickle=False))
    n = len(g["Ap"]) - 1
    for dt in (np.int32, np.int64):
        Ap, Ai = g["Ap"].astype(dt), g["Ai"].astype(dt)
        work, Lnz, et = (np.zeros(n, dt) for _ in range(3))
        tot = mpir_amd.QDLDL_etree(n, Ap, Ai, work, Lnz, et)
        assert tot == int(g["etree_sum"])
        np.testing.assert_array_equal(Lnz, g["Lnz"])
        np.testing.assert_array_equal(et, g["etree"])


def test_qdldl_etree_error_codes(capfd):
    w, l, e = (np.zeros(2, np.int32) for _ in range(3))
    # empty column -> -1
    assert mpir_amd.QDLDL_etree(2, np.array([0, 1, 1], np.int32), np.array([0], np.int32), w, l, e) == -1
    # entry below the diagonal -> -1 and the reference's message
    assert mpir_amd.QDLDL_etree(2, np.array([0, 2, 3], np.int32), np.array([0, 1, 1], np.int32), w, l, e) == -1
    assert "not upper triangular" in capfd.readouterr().out


def test_qdldl_etree_overflow_code():
    # row 0 coupled to every column: L is dense, sum(Lnz) = n (n - 1) / 2
    n = 65600
    Ap = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    Ap[1:] -= 1
    Ai = np.empty(2 * n - 1, np.int64)
    Ai[0] = 0
    Ai[1::2] = 0
    Ai[2::2] = np.arange(1, n)
    work, Lnz, et = (np.zeros(n, np.int32) for _ in range(3))
    assert mpir_amd.QDLDL_etree(n, Ap.astype(np.int32), Ai.astype(np.int32), work, Lnz, et) == -2
    w64, l64, e64 = (np.zeros(n, np.int64) for _ in range(3))
    assert mpir_amd.QDLDL_etree(n, Ap, Ai, w64, l64, e64) == n * (n - 1) // 2


DROPIN = Path(__file__).resolve().parent / "cpp" / "dropin_test"


def test_cpp_dropin_headers_without_a_device():
    # include/mpir/*.hpp compiled into a reference-style caller: the host-side
    # QDLDL_etree works, every GPU entry point fails loudly without a device
    if mpir_amd.device_ok():
        pytest.skip("a gfx950 device is present; the GPU test runs the full program")
    import subprocess
    assert DROPIN.exists(), "tests/cpp/dropin_test is built by mpir_amd.build"
    proc = subprocess.run([str(DROPIN), "--host"], capture_output=True, text=True, timeout=60)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "MPIR_ERR_NODEVICE" in proc.stdout


PLAN_CHECK = Path(__file__).resolve().parent / "cpp" / "plan_check"


def test_factor_and_solve_schedules_for_every_tuning():
    # tests/cpp/plan_check.cpp: for several update groupings, super-block
    # edges and narrow-front limits, every trailing-update entry of every
    # panel step is applied exactly once and in time, the solve tasks cover
    # each front once, the dependency lists are ordered; planted mistakes
    # (a dropped or doubled update task, a dropped narrow task) are caught
    import subprocess
    assert PLAN_CHECK.exists(), "tests/cpp/plan_check is built by mpir_amd.build"
    proc = subprocess.run([str(PLAN_CHECK)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "planted mistakes caught" in proc.stdout and "plan_check ok" in proc.stdout


MTX = Path(__file__).resolve().parent / "golden" / "kkt_cfg0.mtx"


def test_matrix_market_reader_matches_the_generator():
    # the reference driver's read (transposed triplets, sorted by column then
    # row) of a lower-stored symmetric file gives back the upper CSC exactly
    n, Ap, Ai, Ax, sym = mpir_amd.read_mtx(MTX)
    P = mpir_amd.kkt_config(0)
    assert sym and n == P.n
    np.testing.assert_array_equal(Ap, P.Ap)
    np.testing.assert_array_equal(Ai, P.Ai)
    np.testing.assert_array_equal(Ax, P.Ax)


def test_matrix_market_reader_rejects_bad_files(tmp_path):
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(mpir_amd.MpirError, match="coordinate"):
        mpir_amd.read_mtx(bad)
    bad.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n1 1 1.0\n4 1 2.0\n")
    with pytest.raises(mpir_amd.MpirError, match="out of range"):
        mpir_amd.read_mtx(bad)
    with pytest.raises(mpir_amd.MpirError, match="cannot open"):
        mpir_amd.read_mtx(tmp_path / "missing.mtx")


def test_mtx_driver_without_a_device():
    if mpir_amd.device_ok():
        pytest.skip("a gfx950 device is present; the GPU test runs the driver")
    import subprocess
    exe = Path(mpir_amd.__file__).resolve().parent / "mtx_solve"
    assert exe.exists(), "mpir_amd/mtx_solve is built by mpir_amd.build"
    proc = subprocess.run([str(exe), str(MTX)], capture_output=True, text=True, timeout=60)
    assert proc.returncode != 0 and "no gfx950 device" in proc.stderr


def test_batch_generator_sizes_and_local_constraints():
    systems = mpir_amd.batch_config(nsys=16, nrhs=2, seed=3, n1_lo=2e3, n1_hi=6e4)
    assert len(systems) == 16
    for q in systems:
        P = q.kkt
        assert P.n1 == q.k ** q.dim and P.n - P.n1 == P.n1 // 4
        assert q.B.shape == (P.n, 2) and q.X_true.shape == (P.n, 2)
        # constraint columns: 3-5 entries plus -D on the diagonal, all within
        # two grid steps of the row's first cell
        for j in range(P.n1, P.n, max(1, (P.n - P.n1) // 50)):
            rows = P.Ai[P.Ap[j]:P.Ap[j + 1]]
            assert 4 <= len(rows) <= 6 and rows[-1] == j
            cells = rows[:-1]
            coords = np.stack([(cells // q.k ** a) % q.k for a in range(q.dim)], axis=1)
            assert np.abs(coords[:, None, :] - coords[None, :, :]).max() <= 4
    again = mpir_amd.batch_config(nsys=16, nrhs=2, seed=3, n1_lo=2e3, n1_hi=6e4)
    for a, b in zip(systems, again):
        np.testing.assert_array_equal(a.kkt.Ax, b.kkt.Ax)
        np.testing.assert_array_equal(a.B, b.B)


def test_batch_plan_is_the_union_of_the_systems_plans():
    # the joined plan over the forest has the systems' fill and flops, and
    # as many levels as the deepest system
    systems = mpir_amd.batch_config(nsys=5, nrhs=1, seed=7, n1_lo=300, n1_hi=4000)
    pats = [(q.kkt.Ap, q.kkt.Ai) for q in systems]
    joined = mpir_amd.batch_plan_info(pats)
    alone = [mpir_amd.batch_plan_info([p]) for p in pats]
    assert joined["n"] == sum(q.kkt.n for q in systems)
    assert joined["nnz_l"] == sum(a["nnz_l"] for a in alone)
    assert list(joined["nnz_l_each"]) == [a["nnz_l"] for a in alone]
    assert abs(joined["factor_flops"] - sum(a["factor_flops"] for a in alone)) <= len(alone)
    assert joined["nsupernodes"] == sum(a["nsupernodes"] for a in alone)
    assert joined["nlevels"] == max(a["nlevels"] for a in alone)
    assert joined["panel_entries"] == sum(a["panel_entries"] for a in alone)
    assert joined["panel_tasks"] == sum(a["panel_tasks"] for a in alone)
    for a in alone:
        assert a["nnz_l"] > 0 and a["bytes_each"][0] > 0


def test_batch_plan_rejects_bad_patterns():
    P = mpir_amd.kkt_config(0)
    Ai = P.Ai.copy()
    Ai[P.Ap[5]] = 7  # below the diagonal of column 5
    with pytest.raises(mpir_amd.MpirError, match="below the diagonal"):
        mpir_amd.batch_plan_info([(P.Ap, P.Ai), (P.Ap, Ai)])
