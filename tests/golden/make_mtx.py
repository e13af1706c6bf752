"""Writes tests/golden/kkt_cfg0.mtx: the 2D KKT workload (n = 1280) as a
symmetric Matrix Market file stored by its lower triangle, the layout the
reference's driver (examples/example.cpp:87-138) reads. Values with 17
significant digits, so they round-trip exactly.

    python tests/golden/make_mtx.py
"""
from pathlib import Path
import sys

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import mpir_amd  # noqa: E402


def main() -> None:
    P = mpir_amd.kkt_config(0)
    lines = ["%%MatrixMarket matrix coordinate real symmetric",
             "% KKT [H A^T; A -D], H = 2D 5-point Laplacian 32x32 + 1e-2 I, A 256x1024 (3 N(0,1) per row),",
             "% D = 1e-2 I; lower triangle, 1-based",
             f"{P.n} {P.n} {len(P.Ai)}"]
    # entry (i, j) of the upper CSC (i <= j) is stored as the lower entry (j, i)
    for j in range(P.n):
        for q in range(P.Ap[j], P.Ap[j + 1]):
            lines.append(f"{j + 1} {P.Ai[q] + 1} {float(P.Ax[q]):.17g}")
    (HERE / "kkt_cfg0.mtx").write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
