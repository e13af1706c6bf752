"""Benchmark: GMRES-IR with an LDL^T preconditioner on the 64^3 KKT workload.

Workload (one step = numeric factorization + GMRES-IR solve of one RHS):
  K = [H A^T; A -D], H = 3D 7-point Laplacian on 64^3 + 1e-3 I (n1 = 262144),
  A = 65536 x 262144 with 4 entries per row (a uniform column and its +x, +y,
  +z grid neighbours), values N(0,1), D = 1e-4 I; n = 327680. Nested
  dissection ordering (default) and relaxed supernodes; symbolic analysis
  done once before timing (reused per pattern).
  Headline: FP32 factor, FP64 working/residual precision, tolerance 1e-10.
  Alongside: the same solve with an FP64 factor (the plain method).

Inputs (values and right-hand side) are resident in HBM when timing starts.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
# HBM bytes one FP32 factorization of the 64^3 workload moves, measured with
# rocprofv3 (FETCH_SIZE x 2 + WRITE_SIZE, the factors from a 1 GiB copy;
# profiles/hbm_traffic_64cubed.md); a bench run cannot collect counters
FACTOR_TRAFFIC_64 = 23.97e9


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz_upper: int) -> int:
    nnz_f = 2 * nnz_upper - n
    return 12 * nnz_f + 4 * (n + 1) + 16 * n


def precond_bytes(n: int, nnz_l: int, s_f: int, s_x: int = 8) -> int:
    return 2 * (s_f + 4) * nnz_l + 8 * (n + 1) + s_f * n + 4 * s_x * n


def backward_error(P, x) -> float:
    import scipy.sparse as sp
    U = sp.csc_matrix((P.Ax, P.Ai, P.Ap), shape=(P.n, P.n))
    K = (U + sp.triu(U, 1).T).tocsr()
    r = P.b - K @ x
    knorm = abs(K).sum(axis=1).max()
    return float(np.abs(r).max() / (knorm * np.abs(x).max() + np.abs(P.b).max()))


def run_path(mpir, torch, P, uf, tol, steps, warmup, dev_ax, dev_b, dev_x, ordering, ilu=None, autotune=False):
    s = mpir.GmresLDLIR(uf, ordering=ordering)
    if ilu is not None:
        s.SetPreconditioner(mpir.PREC_FGPILU, *ilu)
    s.SetTolerance(tol)
    t0 = time.perf_counter()
    s.Analyze(P.Ap, P.Ai)
    analyze_s = time.perf_counter() - t0
    log(f"{np.dtype(uf).name}: analyze {analyze_s:.2f}s, warmup {warmup}")
    tuned = None
    if autotune:
        # setup, not timed: the level-tile autotuner on this pattern's factor
        s.factor_dev(dev_ax.data_ptr())
        tuned = s.Autotune(reps=3)
        log(f"{np.dtype(uf).name}: autotune {tuned}")
    for _ in range(warmup):
        s.factor_dev(dev_ax.data_ptr())
        s.solve_dev(dev_b.data_ptr(), dev_x.data_ptr())
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for i in range(steps):
        a = time.perf_counter()
        s.factor_dev(dev_ax.data_ptr())
        s.solve_dev(dev_b.data_ptr(), dev_x.data_ptr())
        lat.append((time.perf_counter() - a) * 1e3)
        log(f"{np.dtype(uf).name}: step {i} {lat[-1]:.1f} ms")
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    st = s.stats()
    x = dev_x.cpu().numpy()
    return s, dict(total_s=total, lat_ms=lat, stats=st, analyze_s=analyze_s, x=x, perm=s.perm(), tuned=tuned)


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    k = (len(v) - 1) * q
    lo, hi = int(np.floor(k)), int(np.ceil(k))
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def host_facts() -> dict:
    """Core count, CPU model and frequency governor of this host."""
    facts = {"nproc": os.cpu_count()}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                facts["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    gov = Path("/sys/devices/system/cpu/cpu0/cpufreq/scaling_governor")
    facts["governor"] = gov.read_text().strip() if gov.exists() else "not exposed"
    return facts


def cpu_baseline(P, perm, uf=np.float32, tol: float = 1e-10) -> tuple:
    """The oracle (the plain C restatement of the reference path: QDLDL_etree
    + QDLDL_factor<UF> of the scaled, permuted matrix, then GmresLDLIR::Solve
    with QDLDL_solve as M^-1) on the benchmark workload itself, with the GPU
    run's permutation, once, on one core (the calling thread pinned with
    sched_setaffinity; the reference path is serial). Returns the record and
    the oracle's solution."""
    from tests.oracle_lib import Oracle
    orc = Oracle()
    facts = host_facts()
    old = os.sched_getaffinity(0)
    core = max(old)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        rc, xo, info = orc.gmres_ir(P.Ap, P.Ai, P.Ax, P.b, perm=perm, uf=uf, tol=tol)
        wall = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    assert rc == 0, "oracle GMRES-IR failed"
    solve_s = info["factor_s"] + info["solve_s"]
    name = "FP32" if uf == np.float32 else "FP64"
    rec = {
        "value": round(1.0 / solve_s, 6),
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"the full workload once (n={P.n}): oracle {name}-factor GMRES-IR, same permutation as the GPU, "
                   f"tol {tol:g}: factor {info['factor_s']:.1f} s, refinement {info['solve_s']:.2f} s "
                   f"({int(info['ir_iterations'])} IR / {int(info['gmres_iterations'])} GMRES iterations); "
                   f"pinned to core {core}; host {facts.get('model', '?')}, {facts['nproc']} cores, "
                   f"governor {facts['governor']}"),
        "factor_s": round(info["factor_s"], 3), "refine_s": round(info["solve_s"], 3), "wall_s": round(wall, 3),
        "ir_iterations": int(info["ir_iterations"]), "gmres_iterations": int(info["gmres_iterations"]),
        "stop_reason": int(info["stop_reason"]), "host": facts,
    }
    return rec, xo


def precision_sweep(mpir, k: int, tol: float) -> int:
    """The precision form of the workload: factor precision FP64 / FP32 / FP16
    / BF16 (16-bit: FP32 accumulation, D^{-1} in FP32), the reference's
    diagonal scaling and Ruiz equilibration, D in {1e-2, 1e-4, 1e-6}. One
    JSON line per case: iterations, backward error, forward error against
    the generator's x_true and against the FP64-factor solution."""
    cases = [(np.float64, "f64"), (np.float32, "f32"), (np.float16, "f16"), ("bf16", "bf16")]
    for d in (1e-2, 1e-4, 1e-6):
        P = mpir.kkt_config(1, k=k, d_value=d)
        x64 = None
        for uf, name in cases:
            for scaling, sname in ((mpir.SCALE_DIAG, "diag"), (mpir.SCALE_RUIZ, "ruiz")):
                s = mpir.GmresLDLIR(uf, ordering=mpir.ORDER_ND, scaling=scaling)
                s.SetTolerance(tol)
                row = {"workload": f"KKT 3D {k}^3", "D": d, "factor": name, "scaling": sname}
                try:
                    s.Compute(P.Ap, P.Ai, P.Ax)
                    x = s.Solve(P.b)
                except mpir.MpirError as e:
                    row["error"] = str(e)
                    print(json.dumps(row), flush=True)
                    continue
                st = s.stats()
                if x64 is None and name == "f64":
                    x64 = x
                row.update({"ir_iterations": st["ir_iterations"], "gmres_iterations": st["gmres_iterations"],
                            "stop_reason": st["stop_reason"], "ruiz_sweeps": st["ruiz_sweeps"],
                            "inertia_ok": st["positive_pivots"] == P.n1, "bwd_err": backward_error(P, x),
                            "fwd_err_true": float(np.abs(x - P.x_true).max() / np.abs(P.x_true).max()),
                            "fwd_err_fp64": float(np.abs(x - x64).max() / np.abs(x64).max()),
                            "factor_ms": round(st["factor_ms"], 3), "solve_ms": round(st["solve_ms"], 3)})
                print(json.dumps(row), flush=True)
    return 0


def ilu_solve_bytes(n: int, nnz: int, s_f: int) -> int:
    """Algorithmic bytes of one FGPILU M^{-1}: the L pass reads the column
    pointers, row indices and L values of every pattern entry; the U pass the
    row pointers, (column, entry) pairs and U values of the strictly-upper
    entries and 1 / U_jj; x is read and written by both passes."""
    strict = nnz - n
    return (4 + s_f) * nnz + (8 + s_f) * strict + 8 * (n + 1) + s_f * n + 4 * 8 * n


def fgpilu_form(mpir, torch, args) -> int:
    """The reference's shipped preconditioner on the benchmark workload:
    GmresLDLIR::Compute's FGPILU (ILU(k) pattern of the matrix in its given
    order, MAX_SWEEPS = 5 as examples/example.cpp:144-147 calls it) under the
    same GMRES-IR (10 IR x 10 GMRES iterations, tol 1e-10). One JSON line per
    (k, factor precision): factor / solve time, iterations, stop reason, the
    backward error reached, and the M^{-1} kernel against HBM bandwidth."""
    P = mpir.kkt_config(1, k=args.k)
    dev = torch.device("cuda:0")
    dev_ax = torch.from_numpy(P.Ax).to(dev)
    dev_b = torch.from_numpy(P.b).to(dev)
    dev_x = torch.zeros(P.n, dtype=torch.float64, device=dev)
    for ilu_k in (0, 1):
        for uf, name, s_f in ((np.float32, "f32", 4), (np.float64, "f64", 8)):
            s, r = run_path(mpir, torch, P, uf, args.tol, args.steps, args.warmup, dev_ax, dev_b, dev_x,
                            mpir.ORDER_NATURAL, ilu=(ilu_k, 5))
            st = r["stats"]
            pre_ms = s.time_kernels("precond", 20)
            fac_ms = s.time_kernels("factor", 3)
            nb = ilu_solve_bytes(P.n, st["nnz_l"], s_f)
            row = {"workload": f"KKT 3D {args.k}^3 (n={P.n}), natural order", "preconditioner": f"FGPILU ILU({ilu_k})",
                   "factor": name, "max_sweeps": 5, "sweeps": st["ilu_sweeps"],
                   "nl_res0": st["ilu_nl_residual0"], "nl_res": st["ilu_nl_residual"],
                   "pattern_nnz": st["nnz_l"], "levels": st["nlevels"], "analyze_s": round(r["analyze_s"], 3),
                   "ms_per_solve": round(1e3 * r["total_s"] / args.steps, 3),
                   "factor_ms": round(st["factor_ms"], 3), "solve_ms": round(st["solve_ms"], 3),
                   "ir_iterations": st["ir_iterations"], "gmres_iterations": st["gmres_iterations"],
                   "stop_reason": st["stop_reason"], "bwd_err": backward_error(P, r["x"]),
                   "positive_pivots": st["positive_pivots"], "n1": P.n1,
                   "kernels": {"factor_ms": round(fac_ms, 4), "precond_ms": round(pre_ms, 4),
                               "precond_GBs": round(nb / (pre_ms * 1e-3) / 1e9, 2),
                               "precond_frac_hbm": round(nb / (pre_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
            print(json.dumps(row), flush=True)
    return 0


def batched(mpir, torch, args) -> int:
    """The batched form: `nsys` independent KKT systems (2D/3D grids, n1
    log-uniform in [2e3, 6e4], m = n1/4 local constraint rows of 3-5 entries,
    D = 1e-3 I), `nrhs` right-hand sides each; FP32 factor, FP64 GMRES-IR to
    1e-10, all systems in one block-diagonal solve. One step = numeric
    factorization of every system + GMRES-IR of every (system, RHS)."""
    log(f"generating {args.nsys} systems x {args.nrhs} RHS")
    systems = mpir.batch_config(nsys=args.nsys, nrhs=args.nrhs, seed=args.seed)
    dev = torch.device("cuda:0")
    ax = [torch.from_numpy(q.kkt.Ax).to(dev) for q in systems]
    bt = [torch.from_numpy(np.ascontiguousarray(q.B.T)).to(dev) for q in systems]
    xt = [torch.zeros_like(b) for b in bt]
    s = mpir.BatchLDLIR(np.float32, ordering=mpir.ORDER_ND, threads=16)
    s.SetTolerance(args.tol)
    t0 = time.perf_counter()
    s.Analyze([(q.kkt.Ap, q.kkt.Ai) for q in systems])
    analyze_s = time.perf_counter() - t0
    tot = s.stats()["totals"]
    log(f"analyze {analyze_s:.1f}s: n={tot['n']} nnz_l={tot['nnz_l']} flops={tot['factor_flops']:.3e} "
        f"levels={tot['nlevels']}")
    nr = [args.nrhs] * len(systems)
    axp = [a.data_ptr() for a in ax]
    bp = [b.data_ptr() for b in bt]
    xp = [x.data_ptr() for x in xt]
    s.SetSolveSchedule(mpir.SCHED_DAG if args.batch_schedule == "dag" else mpir.SCHED_LEVELS)
    tuned = None
    if args.autotune:
        # setup, not timed: the level-tile autotuner on this batch's factor,
        # after one solve (M^{-1} is then timed on all right-hand sides)
        s.factor_dev(axp)
        s.solve_dev(nr, bp, xp)
        tuned = s.Autotune(reps=2)
        log(f"autotune: {tuned}")
    # warm-up: the first solve runs kernel by kernel, the solves after it
    # replay the captured graph of the whole solve
    for _ in range(max(1, args.warmup)):
        s.factor_dev(axp)
        s.solve_dev(nr, bp, xp)
    torch.cuda.synchronize()
    lat, steps_ms = [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        a = time.perf_counter()
        s.factor_dev(axp)
        fac_ms = (time.perf_counter() - a) * 1e3
        s.solve_dev(nr, bp, xp)
        steps_ms.append((time.perf_counter() - a) * 1e3)
        st = s.stats()
        # a system is done when its last right-hand side has stopped
        lat.extend((fac_ms + st["done_ms"].max(axis=1)).tolist())
        log(f"step {i}: {steps_ms[-1]:.1f} ms (factor {fac_ms:.1f} ms)")
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    st = s.stats()
    pairs = len(systems) * args.nrhs
    # backward error of every (system, RHS)
    import scipy.sparse as sp
    eta = []
    for q, x in zip(systems, xt):
        X = x.cpu().numpy().T
        U = sp.csc_matrix((q.kkt.Ax, q.kkt.Ai, q.kkt.Ap), shape=(q.kkt.n, q.kkt.n))
        K = (U + sp.triu(U, 1).T).tocsr()
        R = q.B - K @ X
        kn = abs(K).sum(axis=1).max()
        eta.extend((np.abs(R).max(axis=0) / (kn * np.abs(X).max(axis=0) + np.abs(q.B).max(axis=0))).tolist())
    t_fac = s.time_kernels("factor", 1)
    t_prec = s.time_kernels("precond", 3)
    t_prec8 = s.time_kernels("precond_multi", 3)
    t_spmv = s.time_kernels("spmv", 10)
    nnz_upper = sum(int(q.kkt.Ap[-1]) for q in systems)
    b_prec = precond_bytes(tot["n"], tot["nnz_l"], 4)
    b_spmv = spmv_bytes(tot["n"], nnz_upper)
    value = pairs * args.steps / total
    out = {
        "metric": "solves/s (system-RHS pairs) at bwd err <=1e-10, batched form; p50/p99 ms per system",
        "value": round(value, 2), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-factor/f64", "data": "synthetic",
        "config": {"workload": f"{len(systems)} KKT systems (2D/3D grids, n1 log-uniform 2e3..6e4, m=n1/4 local "
                               f"constraint rows of 3-5 entries, D=1e-3) x {args.nrhs} RHS, ND ordering, FP32 LDL^T "
                               f"+ FP64 GMRES-IR, tol {args.tol:g}",
                   "systems": len(systems), "rhs_per_system": args.nrhs, "n_total": tot["n"],
                   "nnz_l": tot["nnz_l"], "factor_flops": tot["factor_flops"], "analyze_s": round(analyze_s, 2)},
        "bwd_err_max": max(eta),
        "p50_ms": round(pct(lat, 0.5), 3), "p99_ms": round(pct(lat, 0.99), 3),
        "gmres_iterations_mean": float(st["gmres_iterations"].mean()),
        "ir_iterations_max": int(st["ir_iterations"].max()),
        "inertia_ok": bool(all(int(p) == q.kkt.n1 for p, q in zip(st["positive_pivots"], systems))),
        "graph_nodes": st["graph_nodes"], "arnoldi_steps": st["arnoldi_steps"],
        "precond_sweeps": st["precond_sweeps"], "autotune": tuned,
        "kernels": {"factor_ms": round(t_fac, 3), "factor_TFLOPs": round(tot["factor_flops"] / t_fac / 1e9, 3),
                    "precond_slot_ms": round(t_prec, 3), "precond_GBs": round(b_prec / t_prec / 1e6, 1),
                    "precond_all_rhs_ms": round(t_prec8, 3),
                    "spmv_slot_ms": round(t_spmv, 4), "spmv_GBs": round(b_spmv / t_spmv / 1e6, 1)},
    }
    tf = tot["factor_flops"] / t_fac / 1e9
    out["roofline"] = {"bound": "mfma", "achieved": round(tf, 3), "peak": PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                       "frac": round(tf / PEAK_TFLOPS["f32"], 4), "traffic": None,
                       "kernel": "numeric LDL^T factorization of the whole batch"}
    out["roofline_solve"] = {"bound": "hbm", "achieved": round(b_prec / t_prec / 1e6, 1), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(b_prec / t_prec / 1e6 / HBM_PEAK_GBS, 4),
                             "kernel": "M^-1 application to one right-hand side of every system"}
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = batched_cpu_baseline(mpir, systems, tot, args)
    print(json.dumps(out), flush=True)
    return 0


def batched_cpu_baseline(mpir, systems, tot, args, seconds_budget: float = 20.0) -> dict:
    """The oracle (one core) on a bounded sample of the batch: the smallest
    systems, one RHS each, FP32 factor + FP64 GMRES-IR; its factor rate
    (flop/s) and solve rate (nnz(L) per second per RHS) scale the sample to
    the whole batch."""
    from tests.oracle_lib import Oracle
    orc = Oracle()
    order = sorted(range(len(systems)), key=lambda i: systems[i].kkt.n)
    fl, lz, tf, ts, used = 0.0, 0, 0.0, 0.0, 0
    t_begin = time.perf_counter()
    for i in order:
        q = systems[i]
        g = mpir.GmresLDLIR(np.float32, ordering=mpir.ORDER_ND)
        g.Analyze(q.kkt.Ap, q.kkt.Ai)
        st = g.stats()
        perm = g.perm()
        rc, x, info = orc.gmres_ir(q.kkt.Ap, q.kkt.Ai, q.kkt.Ax, q.B[:, 0], perm=perm, uf=np.float32, tol=args.tol)
        assert rc == 0
        fl += st["factor_flops"]
        lz += st["nnz_l"]
        tf += info["factor_s"]
        ts += info["solve_s"]
        used += 1
        if time.perf_counter() - t_begin > seconds_budget:
            break
    f_rate = fl / tf
    s_rate = lz / ts
    est = tot["factor_flops"] / f_rate + args.nrhs * tot["nnz_l"] / s_rate
    pairs = len(systems) * args.nrhs
    return {"value": round(pairs / est, 3), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"oracle FP32-factor GMRES-IR on the {used} smallest systems (1 RHS each, {tf + ts:.1f} s): "
                       f"factor {f_rate / 1e9:.2f} GFLOP/s, solve {s_rate / 1e6:.1f} M nnz(L)/s per RHS; scaled to "
                       f"the batch's {tot['factor_flops']:.3e} flops and {args.nrhs} x {tot['nnz_l']} nnz(L)")}


def serving(mpir, args) -> int:
    """The serving form: Poisson arrivals over a pool of 64 KKT patterns (3D
    k^3 Laplacians, k = 16..48, m = k^3/4 rows of 4 entries, D = 1e-3 I);
    each request carries fresh values (5% perturbation of its pattern's base
    values) and 1-4 right-hand sides; FP32 factor, FP64 GMRES-IR to 1e-10.
    Symbolic analysis once per pattern; continuous batching with admission by
    device-memory footprint. Runs the stream at the offered rate, then at 60%
    of the throughput it reached, and reports both."""
    log("registering the 64 serving patterns")
    pats = mpir.serving_patterns(count=64)
    srv = mpir.Server(np.float32, lanes=args.lanes, budget_bytes=int(args.budget_gb * (1 << 30)),
                      max_batch=args.max_batch, tol=args.tol, keep_x=False, threads=16)
    t0 = time.perf_counter()
    srv.add_patterns(pats, threads=16)
    analyze_s = time.perf_counter() - t0
    log(f"symbolic analysis of 64 patterns: {analyze_s:.1f} s")
    # warm-up (code objects; every lane's device buffers grown to full
    # batches, so that the timed streams do not wait on reallocations)
    tw, pw, nw, sw = mpir.poisson_stream(args.rate, 0.3, len(pats), seed=args.seed + 99)
    srv.run(pw, sw, nw, tw)

    def one(rate, seconds, seed):
        t, pat, nr, seeds = mpir.poisson_stream(rate, seconds, len(pats), seed=seed)
        log(f"stream: {len(t)} requests at {rate:.0f}/s over {seconds:g} s")
        out = srv.run(pat, seeds, nr, t)
        summ = srv.summary()
        ok = out["status"] == 0
        lat = out["latency_ms"][ok]
        wall_s = summ["wall_ms"] / 1e3
        sizes = np.bincount(out["batch"][out["batch"] >= 0])
        res = {"offered_rate": rate, "seconds": seconds, "requests": int(len(t)), "completed": summ["completed"],
               "failed": summ["failed"], "wall_s": round(wall_s, 3),
               "throughput_req_s": round(summ["completed"] / wall_s, 2),
               "solves_s": round(float(nr[ok].sum()) / wall_s, 2),
               "p50_ms": round(pct(lat.tolist(), 0.5), 3), "p99_ms": round(pct(lat.tolist(), 0.99), 3),
               "queue_p50_ms": round(pct((out["t_admit"] - out["t_arrive"])[ok].tolist(), 0.5), 3),
               "max_backlog": summ["max_backlog"], "batches": summ["batches"],
               "mean_batch": round(float(sizes.mean()), 2) if len(sizes) else 0.0,
               "peak_inflight_GB": round(summ["peak_inflight_bytes"] / 1e9, 2),
               "max_concurrent_batches": summ["max_concurrent"], "device_regrowths": summ["device_regrowths"],
               "phase_ms_per_batch": {k: round(summ[k + "_ms"] / max(1, summ["batches"]), 2)
                                      for k in ("prepare", "plan", "plan_join", "plan_upload", "factor", "solve")},
               "rel_residual_max": float(out["rel_residual_max"][ok].max()) if ok.any() else None,
               "gmres_iterations_max": int(out["gmres_max"][ok].max()) if ok.any() else None}
        log(f"stream done: {res}")
        return res

    hot = one(args.rate, args.serve_seconds, args.seed)
    steady = one(max(1.0, 0.6 * hot["throughput_req_s"]), args.serve_seconds, args.seed + 1)
    # backward error of a verified sample (solutions kept)
    chk = mpir.Server(np.float32, lanes=args.lanes, budget_bytes=int(args.budget_gb * (1 << 30)),
                      max_batch=args.max_batch, tol=args.tol, keep_x=True, threads=16)
    chk.add_patterns(pats, threads=16)
    t, pat, nr, seeds = mpir.poisson_stream(args.rate, 0.03, len(pats), seed=args.seed + 2)
    chk.run(pat, seeds, nr, t)
    import scipy.sparse as sp
    eta = []
    for i in range(len(t)):
        Ax, B, X = chk.request_data(int(pat[i]), int(seeds[i]), int(nr[i]))
        P = pats[pat[i]]
        U = sp.csc_matrix((Ax, P.Ai, P.Ap), shape=(P.n, P.n))
        K = (U + sp.triu(U, 1).T).tocsr()
        x = chk.solution(i)
        R = B - K @ x
        kn = abs(K).sum(axis=1).max()
        eta.extend((np.abs(R).max(axis=0) / (kn * np.abs(x).max(axis=0) + np.abs(B).max(axis=0))).tolist())
    out = {
        "metric": "requests/s and p50/p99 ms per request at bwd err <=1e-10, serving form",
        "value": hot["throughput_req_s"], "unit": "requests/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32-factor/f64", "data": "synthetic",
        "config": {"workload": "Poisson arrivals over 64 KKT patterns (3D k^3 Laplacians k=16..48, m=k^3/4 rows of "
                               "4 entries, D=1e-3), fresh 5%-perturbed values and 1-4 RHS per request, FP32 LDL^T + "
                               f"FP64 GMRES-IR, tol {args.tol:g}",
                   "lanes": args.lanes, "max_batch": args.max_batch, "budget_GB": args.budget_gb,
                   "pattern_analysis_s": round(analyze_s, 2)},
        "offered": hot, "steady": steady,
        "verified_sample": {"requests": int(len(t)), "pairs": len(eta), "bwd_err_max": max(eta) if eta else None},
    }
    print(json.dumps(out), flush=True)
    return 0


TASK_KINDS = {0: "small", 1: "fx", 2: "fu", 3: "br", 4: "bx", 5: "fm", 6: "bm"}


def precond_timeline(mpir, args) -> int:
    """Where one M^-1 application's time goes, per task kind and per level of
    the supernodal tree: the device task timeline of the dependency-driven
    solves (trace_precond) joined with the plan's supernode levels. Prints a
    table per direction, then one JSON line with the spans."""
    P = mpir.kkt_config(1, k=args.k)
    order = {"amd": mpir.ORDER_AMD, "nd": mpir.ORDER_ND}[args.ordering]
    s = mpir.GmresLDLIR(np.float32, ordering=order)
    s.Compute(P.Ap, P.Ai, P.Ax)
    ms = s.time_kernels("precond", 20)
    s.trace_precond()  # the first traced application warms the trace buffer
    tr, nf = s.trace_precond()
    sn = s.supernode_info()
    kind = (tr[:, 0] & 15).astype(int)
    sup = (tr[:, 0] >> 4).astype(int)
    drawn = tr[:, 1].astype(np.int64)
    waited = tr[:, 2].astype(np.int64)
    done = (tr[:, 3] >> 16).astype(np.int64)
    wg = (tr[:, 3] & 0xFFFF).astype(int)
    out = {"precond_ms_events": round(ms, 4), "tasks": int(len(tr)), "forward_tasks": int(nf)}
    for part, sl in (("forward", slice(0, nf)), ("backward", slice(nf, len(tr)))):
        k, sp, dr, wt, dn = kind[sl], sup[sl], drawn[sl], waited[sl], done[sl]
        t0 = dr.min()
        span = (dn.max() - t0) / 100.0
        run = (dn - wt) / 100.0
        busy = run.sum() / (span * (wg[sl].max() + 1))
        print(f"## {part}: span {span:.1f} us, {len(k)} tasks, {wg[sl].max() + 1} workgroups, "
              f"workgroups running a task {100 * busy:.0f} % of the span", flush=True)
        print("| kind | tasks | wait mean us | run mean us | run p90 us | run sum us |\n|---|---:|---:|---:|---:|---:|")
        for kk in sorted(set(k)):
            m = k == kk
            print(f"| {TASK_KINDS[kk]} | {m.sum()} | {((wt[m] - dr[m]) / 100).mean():.2f} | {run[m].mean():.2f} | "
                  f"{np.percentile(run[m], 90):.2f} | {run[m].sum():.1f} |")
        print("\n| level | tasks | first start us | last end us | level span us | run sum us | max cols | max rows |\n"
              "|---:|---:|---:|---:|---:|---:|---:|---:|")
        lev = sn["level"][sp]
        for lv in sorted(set(lev)):
            m = lev == lv
            a, b = (wt[m].min() - t0) / 100.0, (dn[m].max() - t0) / 100.0
            print(f"| {lv} | {m.sum()} | {a:.1f} | {b:.1f} | {b - a:.1f} | {run[m].sum():.1f} | "
                  f"{sn['ncols'][sp[m]].max()} | {sn['nrows'][sp[m]].max()} |")
        print(flush=True)
        out[part + "_span_us"] = round(span, 1)
        out[part + "_busy"] = round(float(busy), 3)
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=64, help="grid side (64 = the benchmark workload)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--ordering", choices=["amd", "nd"], default="nd",
                    help="fill-reducing ordering (nd = nested dissection, amd = approximate minimum degree)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-factor comparison path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-fp64", action="store_true", help="also time the oracle's FP64-factor path")
    ap.add_argument("--sweep", action="store_true",
                    help="precision form instead of the bench line: f64/f32/f16/bf16 factors x scaling x D scale")
    ap.add_argument("--batched", action="store_true",
                    help="batched form instead of the bench line: many systems x several RHS in one solve")
    ap.add_argument("--nsys", type=int, default=512, help="systems of the batched form")
    ap.add_argument("--nrhs", type=int, default=8, help="right-hand sides per system of the batched form")
    ap.add_argument("--batch-schedule", choices=["dag", "levels"], default="dag",
                    help="triangular-solve schedule of the batched form (before any autotuning)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the level-tile autotuner on the factor before timing (setup, not timed)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--fgpilu", action="store_true",
                    help="the reference's shipped FGPILU preconditioner on the workload (one JSON line per case)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only a 1 GiB device-to-device copy (x3): the reference kernel for HBM counter scaling")
    ap.add_argument("--precond-timeline", action="store_true",
                    help="task timeline of one M^-1 application per kind and tree level (FP32 factor)")
    ap.add_argument("--serve", action="store_true",
                    help="serving form instead of the bench line: Poisson request stream with continuous batching")
    ap.add_argument("--rate", type=float, default=2000.0, help="offered request rate of the serving form (1/s)")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="arrival window of the serving form (s)")
    ap.add_argument("--lanes", type=int, default=3, help="concurrent batches of the serving form")
    ap.add_argument("--max-batch", type=int, default=64, help="requests per batch of the serving form")
    ap.add_argument("--budget-gb", type=float, default=128.0, help="device-memory budget of the serving form")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: the solve path runs on a single device "
              "(a multi-GPU run would shard whole systems, or subtrees of one system's elimination "
              "tree, over the devices); run with --gpus 1", file=sys.stderr)
        return 2
    import torch

    import mpir_amd as mpir
    if not mpir.device_ok():
        print("no gfx950 device", file=sys.stderr)
        return 1
    if args.copy_calib:
        # a plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for
        # 1 GiB moved each way scales the counters of the solver's kernels
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "launches": 3}), flush=True)
        return 0
    if args.sweep:
        return precision_sweep(mpir, args.k, args.tol)
    if args.batched:
        return batched(mpir, torch, args)
    if args.fgpilu:
        return fgpilu_form(mpir, torch, args)
    if args.serve:
        return serving(mpir, args)
    if args.precond_timeline:
        return precond_timeline(mpir, args)
    log(f"generating the {args.k}^3 KKT workload")
    P = mpir.kkt_config(1, k=args.k)
    dev = torch.device("cuda:0")
    dev_ax = torch.from_numpy(P.Ax).to(dev)
    dev_b = torch.from_numpy(P.b).to(dev)
    dev_x = torch.zeros(P.n, dtype=torch.float64, device=dev)

    order = {"amd": mpir.ORDER_AMD, "nd": mpir.ORDER_ND}[args.ordering]
    s32, r32 = run_path(mpir, torch, P, np.float32, args.tol, args.steps, args.warmup, dev_ax, dev_b, dev_x, order,
                        autotune=args.autotune)
    eta32 = backward_error(P, r32["x"])
    st = r32["stats"]
    log(f"f32: {st}")
    # kernel-level timing with HIP events on the solver's stream
    t_prec = s32.time_kernels("precond", 5)
    t_spmv = s32.time_kernels("spmv", 20)
    t_fac = s32.time_kernels("factor", 1)
    nnz_upper = int(P.Ap[-1])
    b_prec = precond_bytes(P.n, st["nnz_l"], 4)
    b_spmv = spmv_bytes(P.n, nnz_upper)
    res = {
        "precond_ms": t_prec, "precond_GBs": b_prec / t_prec / 1e6,
        "spmv_ms": t_spmv, "spmv_GBs": b_spmv / t_spmv / 1e6,
        "factor_ms": t_fac, "factor_TFLOPs": st["factor_flops"] / t_fac / 1e9,
    }
    log(f"kernels: {res}")
    del s32

    r64 = None
    if not args.no_fp64:
        s64, r64 = run_path(mpir, torch, P, np.float64, args.tol, max(1, args.steps // 2), args.warmup, dev_ax, dev_b,
                            dev_x, order, autotune=args.autotune)
        r64["eta"] = backward_error(P, r64["x"])
        del s64

    value = args.steps / r32["total_s"]
    out = {
        "metric": "solves/s + HBM GB/s at bwd err <=1e-10, mixed-IR vs FP64 factor; p50/p99 ms",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["total_s"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-factor/f64",
        "data": "synthetic",
        "config": {"workload": f"KKT 3D {args.k}^3 Laplacian + {P.n - P.n1} constraint rows (n={P.n}), "
                               f"{args.ordering.upper()} ordering, FP32 LDL^T + FP64 GMRES-IR, tol 1e-10, 1 RHS",
                   "n": P.n, "nnz_upper": nnz_upper, "nnz_l": st["nnz_l"], "factor_flops": st["factor_flops"]},
        "bwd_err": eta32,
        "p50_ms": round(pct(r32["lat_ms"], 0.5), 3),
        "p99_ms": round(pct(r32["lat_ms"], 0.99), 3),
        "ir_iterations": st["ir_iterations"],
        "gmres_iterations": st["gmres_iterations"],
        "kernels": {k: round(v, 4) for k, v in res.items()},
        "autotune": r32["tuned"],
    }
    if r64 is not None:
        v64 = len(r64["lat_ms"]) / r64["total_s"]
        out["fp64_factor"] = {"value": round(v64, 4), "bwd_err": r64["eta"],
                              "p50_ms": round(pct(r64["lat_ms"], 0.5), 3),
                              "ir_iterations": r64["stats"]["ir_iterations"],
                              "gmres_iterations": r64["stats"]["gmres_iterations"],
                              "factor_ms": round(r64["stats"]["factor_ms"], 3),
                              "solve_ms": round(r64["stats"]["solve_ms"], 3)}
        out["mixed_vs_fp64"] = round(value / v64, 4)
    # dominant kernel group of a solve: the numeric factorization (MFMA-bound
    # dense fronts); the M^-1 application and the SpMV (HBM-bound) alongside
    tf = res["factor_TFLOPs"]
    out["roofline"] = {"bound": "mfma", "achieved": round(tf, 3), "peak": PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                       "frac": round(tf / PEAK_TFLOPS["f32"], 4),
                       "traffic": FACTOR_TRAFFIC_64 if args.k == 64 and order == mpir.ORDER_ND else None,
                       "kernel": "numeric LDL^T factorization (FP32; F = sum Lnz_j (Lnz_j + 2) + n per launch sequence)"}
    out["roofline_solve"] = {"bound": "hbm", "achieved": round(res["precond_GBs"], 2), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(res["precond_GBs"] / HBM_PEAK_GBS, 4),
                             "kernel": "M^-1 application (L, D^-1, L^T passes)"}
    out["roofline_spmv"] = {"bound": "hbm", "achieved": round(res["spmv_GBs"], 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(res["spmv_GBs"] / HBM_PEAK_GBS, 4),
                            "kernel": "FP64 CSR SpMV (35 MB working set, fits the Infinity Cache)"}
    # forward errors: against the generator's exact solution and against the
    # FP64-factor GPU solution (and, below, the oracle's CPU solution)
    out["fwd_err"] = float(np.abs(r32["x"] - P.x_true).max() / np.abs(P.x_true).max())
    if r64 is not None:
        out["fwd_err_vs_fp64_factor"] = float(np.abs(r32["x"] - r64["x"]).max() / np.abs(r64["x"]).max())
    if not args.no_cpu:
        log("cpu baseline: the oracle on the full workload, one core")
        rec, xo = cpu_baseline(P, r32["perm"], np.float32, args.tol)
        out["cpu_baseline"] = rec
        out["fwd_err_vs_oracle"] = float(np.abs(r32["x"] - xo).max() / np.abs(xo).max())
        out["oracle_bwd_err"] = backward_error(P, xo)
        if args.cpu_fp64:
            rec64, _ = cpu_baseline(P, r32["perm"], np.float64, args.tol)
            out["cpu_baseline_fp64_factor"] = rec64
        else:
            out["cpu_baseline_fp64_factor"] = ("not run by default (the FP64 oracle factor of this workload "
                                               "takes as long again); bench.py --cpu-fp64 runs it")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
