"""Write the golden fixtures of tests/golden/ from the reference's own code.

    make -C oracle ref            # builds oracle/_ref/libqdldl_ref.so from
                                  # /root/reference/include/MPIR as shipped
    python tests/golden/make_golden.py

Inputs come from the deterministic KKT generator of mpir_amd (host code) and
are stored in the fixtures, so each fixture is self-contained: inputs and the
outputs of QDLDL_etree / QDLDL_factor / QDLDL_solve / MatrixMultiply /
VectorDot / Dnrm2 / InfNrm as compiled from the reference headers.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import mpir_amd  # noqa: E402
from tests.oracle_lib import Reference, permute_upper  # noqa: E402


def sha256_arrays(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def qdldl_case(ref: Reference, kkt, perm, seed: int) -> dict:
    Ap, Ai, Ax = permute_upper(kkt.Ap, kkt.Ai, kkt.Ax, perm) if perm is not None else (kkt.Ap, kkt.Ai, kkt.Ax)
    out = dict(Ap=Ap, Ai=Ai, Ax=Ax)
    tot, Lnz, et = ref.etree(Ap, Ai)
    out.update(etree_sum=np.int64(tot), Lnz=Lnz, etree=et)
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, len(Ap) - 1)
    out["b"] = b
    for tag, dt in (("f64", np.float64), ("f32", np.float32)):
        ret, F = ref.factor(Ap, Ai, Ax, dt)
        out[f"{tag}_ret"] = np.int64(ret)
        for key in ("Lp", "Li", "Lx", "D", "Dinv"):
            out[f"{tag}_{key}"] = F[key]
        out[f"{tag}_solve_f64"] = ref.solve(F["Lp"], F["Li"], F["Lx"], F["Dinv"], b, np.float64)
        if dt == np.float32:
            out[f"{tag}_solve_f32"] = ref.solve(F["Lp"], F["Li"], F["Lx"], F["Dinv"], b, np.float32)
    out["matmul"] = ref.matmul(Ap, Ai, Ax, b)
    return out


def main() -> int:
    ref = Reference.load()
    if ref is None:
        print("oracle/_ref/libqdldl_ref.so is missing: run `make -C oracle ref` first", file=sys.stderr)
        return 1
    manifest = {"command": "make -C oracle ref && python tests/golden/make_golden.py", "files": {}}

    # QDLDL on a small 2D KKT system, natural and AMD orders
    small = mpir_amd.gen_kkt(dim=2, k=12, h_shift=1e-2, m=36, a_mode=0, cnt_lo=3, cnt_hi=3, d_value=1e-2, seed=3)
    perm = mpir_amd.amd_order(small.Ap, small.Ai)
    for name, pm in (("qdldl_small_natural", None), ("qdldl_small_amd", perm)):
        np.savez_compressed(HERE / f"{name}.npz", **qdldl_case(ref, small, pm, seed=11))

    # 3D system with grid-neighbour constraint rows (the 64^3 workload's family)
    k3 = mpir_amd.gen_kkt(dim=3, k=6, h_shift=1e-3, m=54, a_mode=1, cnt_lo=4, cnt_hi=4, d_value=1e-4, seed=5)
    np.savez_compressed(HERE / "qdldl_k6_amd.npz",
                        **qdldl_case(ref, k3, mpir_amd.amd_order(k3.Ap, k3.Ai), seed=12))

    # the n = 1280 workload: generator digest, etree and counts of its AMD order
    cfg0 = mpir_amd.kkt_config(0)
    p0 = mpir_amd.amd_order(cfg0.Ap, cfg0.Ai)
    Cp, Ci, Cx = permute_upper(cfg0.Ap, cfg0.Ai, cfg0.Ax, p0)
    tot, Lnz, et = ref.etree(Cp, Ci)
    y = ref.matmul(cfg0.Ap, cfg0.Ai, cfg0.Ax, cfg0.x_true)
    np.savez_compressed(HERE / "cfg0_symbolic.npz", perm=p0, etree_sum=np.int64(tot), Lnz=Lnz, etree=et,
                        matmul_xtrue=y)
    manifest["cfg0_generator_sha256"] = sha256_arrays(cfg0.Ap, cfg0.Ai, cfg0.Ax, cfg0.x_true, cfg0.b)

    # ops.hpp vector kernels
    rng = np.random.default_rng(7)
    vecs = {
        "v_rand": rng.standard_normal(1000),
        "w_rand": rng.standard_normal(1000),
        "v_odd": rng.uniform(-1, 1, 7),
        "w_odd": rng.uniform(-1, 1, 7),
        "v_big": rng.standard_normal(64) * 1e200,
        "v_tiny": rng.standard_normal(64) * 1e-200,
        "v_mixed": np.concatenate([rng.standard_normal(5) * 1e150, np.zeros(3), rng.standard_normal(5) * 1e-150]),
    }
    out = dict(vecs)
    for key, v in vecs.items():
        out[f"dnrm2_{key}"] = np.float64(ref.dnrm2(v))
        out[f"infnrm_{key}"] = np.float64(ref.infnrm(v))
    out["dot_rand"] = np.float64(ref.dot(vecs["v_rand"], vecs["w_rand"]))
    out["dot_odd"] = np.float64(ref.dot(vecs["v_odd"], vecs["w_odd"]))
    np.savez_compressed(HERE / "ops_vectors.npz", **out)

    for f in sorted(HERE.glob("*.npz")):
        manifest["files"][f.name] = hashlib.sha256(f.read_bytes()).hexdigest()
    (HERE / "manifest.json").write_text(json.dumps(manifest, indent=2) + "\n")
    print(json.dumps(manifest, indent=2))
    return 0


if __name__ == "__main__":
    sys.exit(main())
