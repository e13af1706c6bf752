"""Build libmpir_hip.so (HIP for gfx950 + host C++) in-tree.

    python -m mpir_amd.build            # compile everything, always
    python -m mpir_amd.build --jobs 8

Objects go to _build/ at the repository root; the shared library is written
to mpir_amd/libmpir_hip.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build" / "mpir_hip"
LIB = PKG / "libmpir_hip.so"
DROPIN_SRC = ROOT / "tests" / "cpp" / "dropin.cpp"
DROPIN_BIN = ROOT / "tests" / "cpp" / "dropin_test"
PLANCHECK_SRC = ROOT / "tests" / "cpp" / "plan_check.cpp"
PLANCHECK_BIN = ROOT / "tests" / "cpp" / "plan_check"
DRIVER_SRC = CSRC / "driver" / "mtx_solve.cpp"
DRIVER_BIN = PKG / "mtx_solve"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-g", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-munsafe-fp-atomics", "-fvisibility=hidden", "-fvisibility-inlines-hidden",
]
# Host code: no contraction so host sums round exactly as written.
CXX_FLAGS = ["-O2", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-sign-compare",
             "-fvisibility=hidden", "-fvisibility-inlines-hidden"]

HOST_SOURCES = [
    CSRC / "host" / "generators.cpp",
    CSRC / "host" / "ordering.cpp",
    CSRC / "host" / "ordering_nd.cpp",
    CSRC / "host" / "symbolic.cpp",
    CSRC / "host" / "mtx.cpp",
    CSRC / "plan.cpp",
    CSRC / "batch_plan.cpp",
    CSRC / "host" / "ilu.cpp",
]
HIP_SOURCES = [
    CSRC / "kernels" / "factor.hip",
    CSRC / "kernels" / "solve.hip",
    CSRC / "kernels" / "spmv.hip",
    CSRC / "kernels" / "ilu.hip",
    CSRC / "gmres.hip",
    CSRC / "batch.hip",
    CSRC / "ilu_solver.hip",
    CSRC / "noalloc.hip",
    CSRC / "serve.cpp",
    CSRC / "capi.cpp",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def _obj(src: Path) -> Path:
    rel = src.relative_to(CSRC)
    return BUILD / (str(rel).replace("/", "__") + ".o")


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    """Compile every source and link the shared library; returns its path."""
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    cmds = []
    for src in HOST_SOURCES:
        cmds.append([CXX, *CXX_FLAGS, "-c", str(src), "-o", str(_obj(src))])
    for src in HIP_SOURCES:
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(_obj(src))])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, c) for c in cmds]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    objs = [str(_obj(s)) for s in HOST_SOURCES + HIP_SOURCES]
    # only the C ABI of include/mpir_hip.h is exported (hidden visibility),
    # and the library binds its own references to itself (-Bsymbolic), so
    # two builds loaded into one process never call into each other
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    # the C++ drop-in test program (a reference-style caller of include/mpir/*.hpp)
    _run([CXX, "-std=c++20", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(DROPIN_SRC), f"-L{PKG}", "-lmpir_hip",
          "-Wl,-rpath,$ORIGIN/../../mpir_amd", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(DROPIN_BIN)])
    # host-only checker of the factor and solve schedules (linked against the
    # host objects: it needs no device and no exported symbol)
    plan_objs = [str(_obj(CSRC / f)) for f in ("plan.cpp", "host/symbolic.cpp", "host/ordering.cpp",
                                                 "host/ordering_nd.cpp", "host/generators.cpp")]
    _run([CXX, *[f for f in CXX_FLAGS if f != "-fPIC"], f"-I{CSRC}", f"-I{CSRC / 'host'}", str(PLANCHECK_SRC),
          *plan_objs, "-o", str(PLANCHECK_BIN)])
    # the Matrix Market driver (examples/example.cpp of the reference)
    _run([CXX, "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(DRIVER_SRC), f"-L{PKG}", "-lmpir_hip",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(DRIVER_BIN)])
    if verbose:
        print(f"built {LIB}")
    return LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    build(args.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
