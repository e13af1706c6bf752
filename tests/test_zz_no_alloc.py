"""mpir_gmres_ldl_precond_dev_noalloc (no device allocation) through the C
ABI: bit-identical to mpir_gmres_ldl_precond_dev, on calls in a row whose
sizes shrink and grow again (the static work memory keeps the previous
call's contents), for FP32 and FP64 factors."""
import ctypes as C

import numpy as np
import pytest

import mpir_amd

pytestmark = pytest.mark.gpu


def _handle(uf, k, seed):
    P = mpir_amd.kkt_config(1, k=k, seed=seed)
    s = mpir_amd.GmresLDLIR(uf, ordering=mpir_amd.ORDER_ND)
    s.Compute(P.Ap, P.Ai, P.Ax)
    return s, P.n


def _apply(lib, fn, s, x):
    import torch
    d = torch.from_numpy(x).to("cuda:0")
    rc = fn(s.handle, C.c_void_p(d.data_ptr()))
    torch.cuda.synchronize()
    return rc, d.cpu().numpy()


@pytest.mark.parametrize("uf", [np.float32, np.float64])
def test_noalloc_equals_precond_dev_on_calls_in_a_row(gpu, uf):
    lib = mpir_amd.load()
    big, n_big = _handle(uf, 24, 1)
    small, n_small = _handle(uf, 12, 2)
    rng = np.random.default_rng(7)
    for s, n in ((big, n_big), (small, n_small), (big, n_big)):
        x = rng.uniform(-1, 1, n)
        rc1, y1 = _apply(lib, lib.mpir_gmres_ldl_precond_dev, s, x.copy())
        rc2, y2 = _apply(lib, lib.mpir_gmres_ldl_precond_dev_noalloc, s, x.copy())
        print(f"n={n}: rc {rc1} {rc2}", flush=True)
        assert rc1 == 0 and rc2 == 0, mpir_amd.last_error()
        assert np.isfinite(y2).all()
        np.testing.assert_array_equal(y2, y1)


def test_noalloc_rejects_what_it_cannot_do(gpu):
    lib = mpir_amd.load()
    s = mpir_amd.GmresLDLIR(np.float32)  # no factor yet
    import torch
    d = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    assert lib.mpir_gmres_ldl_precond_dev_noalloc(s.handle, C.c_void_p(d.data_ptr())) < 0
    # the 64^3 factor: either its solves fit the 64 MiB of static work
    # memory (then the result equals precond_dev's), or the call says how
    # many bytes they need
    big, n = _handle(np.float32, 64, 0)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    rc1, y1 = _apply(lib, lib.mpir_gmres_ldl_precond_dev, big, x.copy())
    rc2, y2 = _apply(lib, lib.mpir_gmres_ldl_precond_dev_noalloc, big, x.copy())
    msg = mpir_amd.last_error()
    print(f"64^3: rc {rc2}: {msg if rc2 else 'fits'}", flush=True)
    assert rc1 == 0
    if rc2 == 0:
        np.testing.assert_array_equal(y2, y1)
    else:
        assert rc2 == -100 and "bytes of work memory" in msg
