#!/usr/bin/env python3
"""Benchmark of the L-sweeps preconditioned Helmholtz solve on one MI355X.

Metric: solves/s and p50/p99 latency at true relative residual <= 1e-6,
mixed precision (FP32 local factors and sweeps under FP64 flexible GMRES)
against the plain FP64 path (FP64 factors, the reference GMRES at tol 1e-8).

A step is one solve of the workload (one right-hand side) with the factors
resident on the device and the right-hand side already in HBM.  Factorisation
is timed separately, as the reference does ("Time to factorize system" vs
"Time to run the GMRES", src/Problem.cpp:143-144).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lens1600|const400]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "l-sweeps-2d_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "solves/s + p50/p99 latency at rel. residual 1e-6: mixed/batched vs plain FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q)) if len(v) else float("nan")


def run_path(L, torch, cfg, precision, mode, tol, max_iter, steps, warmup, restart=30):
    p = L.dd_params(domain=cfg["domain"], ncells=cfg["ncells"], nels=cfg["nels"], omega=cfg["omega"],
                    min_wl_in_pml=cfg["min_wl_in_pml"], precision=precision, velocity=cfg["velocity"],
                    vel_domain=cfg["vel_domain"])
    t0 = time.perf_counter()
    dd = L.DDCheckerboard(p)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    dd.factorize()
    dd.sync()
    t_fact = time.perf_counter() - t0
    b = torch.zeros(dd.N, dtype=torch.complex128, device="cuda:0")
    x = torch.zeros_like(b)
    dd.assemble_rhs_dev(cfg["sources"], b.data_ptr())
    dd.sync()
    for _ in range(warmup):
        dd.gmres_dev(b.data_ptr(), x.data_ptr(), mode=mode, tol=tol, max_iter=max_iter, restart=restart)
    dd.sync()
    torch.cuda.synchronize()
    lat = []
    infos = []
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        info = dd.gmres_dev(b.data_ptr(), x.data_ptr(), mode=mode, tol=tol, max_iter=max_iter, restart=restart)
        dd.sync()
        lat.append(time.perf_counter() - t0)
        infos.append(info)
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t_all
    # per-launch profile of the batched local solves (outside the timed region)
    y = torch.zeros_like(b)
    prof = dd.profile_precond_dev(b.data_ptr(), y.data_ptr())
    st = dd.stats()
    res = dict(setup_s=t_setup, factorize_s=t_fact, solve_s=t_all / steps, solves_per_s=steps / t_all,
               p50_ms=1e3 * pct(lat, 50), p99_ms=1e3 * pct(lat, 99),
               iterations=infos[-1]["iterations"], precond_applies=infos[-1]["precond_applies"],
               true_residual=max(i["true_residual"] for i in infos), estimate=infos[-1]["estimate"],
               dofs=dd.N, local_solves_per_precond=st["solves_per_precond"], factor_bytes=st["factor_bytes"],
               profile=prof)
    x_host = x.cpu().numpy()
    del dd
    torch.cuda.empty_cache()
    return res, x_host


def cpu_baseline(cfg, iterations, seconds=20.0):
    """The oracle's local SuperLU solve on one interior cell, timed on the host
    (1 core), extrapolated to one reference-mode GMRES solve."""
    import numpy as np
    from oracle.geometry import Checkerboard, ConstantMedium, CSVMedium
    t0 = time.perf_counter()
    med = ConstantMedium() if cfg["velocity"] is None else CSVMedium(cfg["velocity"], [list(d) for d in cfg["vel_domain"]])
    cb = Checkerboard([list(cfg["domain"][0]), list(cfg["domain"][1])], cfg["ncells"], cfg["nels"], cfg["omega"],
                      cfg["min_wl_in_pml"], med, assemble=False)
    nR, nC = cfg["ncells"]
    c = cb.cell(nR // 2, nC // 2)
    c.assemble()
    tf = time.perf_counter()
    c.factorize()
    t_fact = time.perf_counter() - tf
    rng = np.random.default_rng(0)
    b = rng.standard_normal(c.size) + 1j * rng.standard_normal(c.size)
    ts = []
    while sum(ts) < seconds * 0.5 and len(ts) < 200:
        t1 = time.perf_counter()
        c.solve(b)
        ts.append(time.perf_counter() - t1)
    t_loc = float(np.median(ts))
    n = max(nR, nC)
    per_precond = 4 * (3 * n - 2) ** 2
    t_solve = (iterations + 2) * per_precond * t_loc
    return dict(value=1.0 / t_solve, unit="solves/s", cores=1, kind="port",
                sample=(f"oracle (numpy/scipy SuperLU) on cell ({nR // 2},{nC // 2}) of the {cfg['name']} "
                        f"workload: factorize 1 of {nR * nC} cells ({t_fact:.2f} s) and {len(ts)} local solves "
                        f"(median {t_loc * 1e3:.1f} ms), extrapolated to {iterations}+2 preconditioner "
                        f"applications x {per_precond} local solves; total sample "
                        f"{time.perf_counter() - t0:.1f} s"),
                local_solve_ms=t_loc * 1e3, local_factor_s=t_fact)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lens1600")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--paths", default="mixed,plain")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(sharding across GPUs is design-only in this build)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import lsweeps_amd as L
    from lsweeps_amd.workloads import CONFIGS

    cfg = CONFIGS[args.workload]()
    log(f"workload {cfg['name']}: grid {cfg['grid']}, board {cfg['ncells']}, omega {cfg['omega']:.3f}")
    paths = args.paths.split(",")
    out = {}
    xs = {}
    if "plain" in paths:
        log("plain FP64 path (reference GMRES, tol %.0e)" % cfg["tol"])
        out["plain"], xs["plain"] = run_path(L, torch, cfg, L.LSW_PREC_FP64, 0, cfg["tol"], cfg["max_iter"],
                                             args.steps, args.warmup)
        log("  ", {k: v for k, v in out["plain"].items() if k != "profile"})
    if "mixed" in paths:
        log("mixed path (FP32 factors/sweeps, FP64 flexible GMRES, tol 1e-6)")
        out["mixed"], xs["mixed"] = run_path(L, torch, cfg, L.LSW_PREC_FP32, 1, 1e-6, 60, args.steps, args.warmup)
        log("  ", {k: v for k, v in out["mixed"].items() if k != "profile"})
    head = out.get("mixed") or out["plain"]
    prof = head["profile"]
    ms = prof["fwd_ms"] + prof["bwd_ms"]
    nl = prof["fwd_launches"] + prof["bwd_launches"]
    byts = prof["fwd_bytes"] + prof["bwd_bytes"]
    achieved = (byts / nl) / (ms / nl * 1e-3) / 1e9 if ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_nd_fwd + k_nd_bwd (batched ND local solves)",
            "bytes_per_launch": byts / nl, "ms_per_launch": ms / nl, "launches_per_precond": nl}
    cpu = None
    if not args.no_cpu_baseline:
        iters = out["plain"]["iterations"] if "plain" in out else head["iterations"]
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(cfg, iters)
        log("  ", cpu)
    line = {
        "metric": METRIC,
        "value": round(head["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["solve_s"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 factors/sweeps + c128 Krylov" if "mixed" in out else "c128",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "grid": list(cfg["grid"]), "board": list(cfg["ncells"]),
                   "omega": round(cfg["omega"], 4), "sources": len(cfg["sources"]), "dofs": head["dofs"],
                   "path": "mixed" if "mixed" in out else "plain"},
        "p50_ms": round(head["p50_ms"], 3),
        "p99_ms": round(head["p99_ms"], 3),
        "dofs_per_s": round(head["dofs"] * head["solves_per_s"], 1),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    for k, v in out.items():
        line[k] = {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items() if kk != "profile"}
        line[k]["profile"] = v["profile"]
    if "mixed" in out and "plain" in out:
        line["speedup_mixed_vs_plain"] = round(out["mixed"]["solves_per_s"] / out["plain"]["solves_per_s"], 3)
        xm, xp = xs["mixed"], xs["plain"]
        line["mixed_vs_plain_rel_l2"] = float(np.linalg.norm(xm - xp) / np.linalg.norm(xp))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
