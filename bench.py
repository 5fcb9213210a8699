#!/usr/bin/env python3
"""Benchmark of the L-sweeps preconditioned Helmholtz solve on one MI355X.

Metric: solves/s and p50/p99 latency at true relative residual <= 1e-6,
mixed precision / batched against plain FP64.  Paths measured on the same
model:

  plain_ref  FP64 factors and sweeps, the reference GMRES at its tolerance
             (the parity path; its logged estimate is not the true residual)
  plain      FP64 factors and sweeps, flexible GMRES to true residual 1e-6
  mixed      FP32 factors and sweeps, FP64 flexible GMRES to 1e-6  (value)
  batched    mixed, 16 point sources along y = 0.1 solved as one block

A step is one solve (one right-hand side; for `batched` one block, counted as
16 solves) with the factors resident in HBM and the right-hand side already
on the device.  Factorisation is timed separately, as the reference does
("Time to factorize system" / "Time to run the GMRES", src/Problem.cpp:143-144).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lens1600|const400]
       python bench.py --workload serve    (request stream through the server)
       python bench.py --workload pack48   (48 uneven problems at once)
       python bench.py --workload layered  (2400x800 layered model, 128 sources as one block)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "l-sweeps-2d_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "solves/s + p50/p99 latency at rel. residual 1e-6: mixed/batched vs plain FP64"
BATCH = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q)) if len(v) else float("nan")


def make_dd(L, cfg, precision):
    p = L.dd_params(domain=cfg["domain"], ncells=cfg["ncells"], nels=cfg["nels"], omega=cfg["omega"],
                    min_wl_in_pml=cfg["min_wl_in_pml"], precision=precision, velocity=cfg["velocity"],
                    vel_domain=cfg["vel_domain"])
    t0 = time.perf_counter()
    dd = L.DDCheckerboard(p)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    dd.factorize()
    dd.sync()
    return dd, t_setup, time.perf_counter() - t0


def time_single(dd, torch, cfg, mode, tol, max_iter, steps, warmup):
    b = torch.zeros(dd.N, dtype=torch.complex128, device="cuda:0")
    x = torch.zeros_like(b)
    # torch fills on its own stream; the library's default stream is not
    # ordered with it
    torch.cuda.synchronize()
    dd.assemble_rhs_dev(cfg["sources"], b.data_ptr())
    dd.sync()
    for _ in range(warmup):
        dd.gmres_dev(b.data_ptr(), x.data_ptr(), mode=mode, tol=tol, max_iter=max_iter, restart=30)
    dd.sync()
    lat, infos = [], []
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        infos.append(dd.gmres_dev(b.data_ptr(), x.data_ptr(), mode=mode, tol=tol, max_iter=max_iter, restart=30))
        dd.sync()
        lat.append(time.perf_counter() - t0)
    t_all = time.perf_counter() - t_all
    return dict(solve_s=t_all / steps, solves_per_s=steps / t_all, p50_ms=1e3 * pct(lat, 50),
                p99_ms=1e3 * pct(lat, 99), iterations=infos[-1]["iterations"],
                precond_applies=infos[-1]["precond_applies"],
                true_residual=max(i["true_residual"] for i in infos), estimate=infos[-1]["estimate"]), \
        x.cpu().numpy()


def time_batched(dd, torch, cfg, tol, steps, warmup):
    import numpy as np
    xs = (np.arange(BATCH) + 0.5) / BATCH
    srcs = [(float(v), 0.1) for v in xs]
    B = torch.zeros(BATCH * dd.N, dtype=torch.complex128, device="cuda:0")
    X = torch.zeros_like(B)
    torch.cuda.synchronize()
    dd.assemble_rhs_block_dev(srcs, B.data_ptr())
    dd.sync()
    for _ in range(warmup):
        dd.gmres_block_dev(BATCH, B.data_ptr(), X.data_ptr(), tol=tol, max_iter=60, restart=30)
    dd.sync()
    lat, infos = [], []
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        infos = dd.gmres_block_dev(BATCH, B.data_ptr(), X.data_ptr(), tol=tol, max_iter=60, restart=30)
        dd.sync()
        lat.append(time.perf_counter() - t0)
    t_all = time.perf_counter() - t_all
    return dict(block=BATCH, block_s=t_all / steps, solves_per_s=steps * BATCH / t_all,
                p50_ms=1e3 * pct(lat, 50), p99_ms=1e3 * pct(lat, 99),
                iterations_max=max(i["iterations"] for i in infos),
                true_residual=max(i["true_residual"] for i in infos))


def profile(dd, torch):
    b = torch.randn(dd.N, dtype=torch.complex128, device="cuda:0")
    y = torch.zeros_like(b)
    torch.cuda.synchronize()
    return dd.profile_precond_dev(b.data_ptr(), y.data_ptr())


def roofline(prof):
    ms = prof["fwd_ms"] + prof["bwd_ms"]
    nl = prof["fwd_launches"] + prof["bwd_launches"]
    byts = prof["fwd_bytes"] + prof["bwd_bytes"]
    achieved = (byts / nl) / (ms / nl * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_nd_packed + k_nd_wide (batched ND local solves, one M^-1 application)",
            "bytes_per_launch": round(byts / nl, 1), "ms_per_launch": round(ms / nl, 5),
            "launches_per_precond": nl, "fwd": prof}


def load_traffic(workload):
    """HBM bytes per solve launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (profiles/traffic_<workload>.json), if present."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(cfg, iterations, seconds=20.0):
    """The oracle's local SuperLU solve on one interior cell, timed on the host
    (1 core), extrapolated to one reference-mode GMRES solve."""
    import numpy as np
    from oracle.geometry import Checkerboard, ConstantMedium, CSVMedium
    t0 = time.perf_counter()
    med = ConstantMedium() if cfg["velocity"] is None else CSVMedium(cfg["velocity"],
                                                                      [list(d) for d in cfg["vel_domain"]])
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
    while sum(ts) < seconds * 0.75 and len(ts) < 3000:
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
                        f"applications x {per_precond} local solves of the reference GMRES; total sample "
                        f"{time.perf_counter() - t0:.1f} s"),
                local_solve_ms=round(t_loc * 1e3, 3), local_factor_s=round(t_fact, 3))


def _problem_params(L, p, precision):
    return L.dd_params(domain=p["domain"], ncells=p["ncells"], nels=p["nels"], omega=p["omega"],
                       min_wl_in_pml=p["min_wl_in_pml"], precision=precision, velocity=p["velocity"],
                       vel_domain=p["vel_domain"], general=p.get("general", False))


def plain_request(L, torch, params, sources, tol=1e-6, max_iter=200):
    """The plain FP64 path for one request, as the reference runs a problem:
    set up and factorise (no reuse across requests, src/Problem.cpp:70-91),
    then each source solved on its own (FP64 flexible GMRES to the true
    residual).  Returns (seconds, worst true residual, max iterations)."""
    t0 = time.perf_counter()
    dd = L.DDCheckerboard(params)
    dd.factorize()
    b = torch.zeros(dd.N, dtype=torch.complex128, device="cuda:0")
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    worst, iters = 0.0, 0
    for s in sources:
        dd.assemble_rhs_dev([s], b.data_ptr())
        info = dd.gmres_dev(b.data_ptr(), x.data_ptr(), mode=1, tol=tol, max_iter=max_iter, restart=30)
        worst = max(worst, info["true_residual"])
        iters = max(iters, info["iterations"])
    dd.sync()
    t = time.perf_counter() - t0
    del dd
    return t, worst, iters


def run_stream(L, srv, pids, stream, t_scale=1.0):
    """Submits the requests of `stream` at their arrival times (scaled) from
    this thread; returns the per-request results and the wall time."""
    import threading
    ids = [None] * len(stream)
    t0 = time.perf_counter()
    res = [None] * len(stream)

    def waiter():
        for k in range(len(stream)):
            while ids[k] is None:
                time.sleep(0.001)
            res[k] = srv.wait(ids[k])

    th = threading.Thread(target=waiter)
    th.start()
    for k, (ta, key, srcs) in enumerate(stream):
        dt = ta * t_scale - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids[k] = srv.submit(pids[key], srcs, want_solution=False)
    th.join()
    return res, time.perf_counter() - t0


def serve_bench(args):
    """Serving workload: a Poisson stream over layered models (seeds 0..M-1)
    x 4 frequencies (6, 8, 10, 12 points per wavelength), 1-64 sources per
    request near the surface.  Measures the plain FP64 path one request at a
    time on a sample of the stream, then the server (FP32 factors, FP64
    GMRES, continuous batching) at 0.8x the plain capacity and at saturation."""
    import numpy as np
    import torch
    import lsweeps_amd as L
    from lsweeps_amd.workloads import config_serve_problem, serve_requests, SERVE_PPW
    M, F = args.serve_models, len(SERVE_PPW)
    probs = {}
    t0 = time.perf_counter()
    for m in range(M):
        for f in range(F):
            probs[(m, f)] = config_serve_problem(m, f)
    log(f"serve: {len(probs)} problems generated in {time.perf_counter() - t0:.1f} s")

    def sources(key, xf):
        p = probs[key]
        (ax, bx), (ay, by) = p["domain"]
        h = (by - ay) / (p["grid"][1] - 1)
        return [(ax + v * (bx - ax), by - 10 * h) for v in xf]

    raw = serve_requests(args.serve_requests, rate=1.0, seed=0, models=M, freqs=F)
    stream = [(t, (m, f), sources((m, f), xf)) for (t, m, f, xf) in raw]
    # plain path on a bounded sample
    rng = np.random.default_rng(1)
    sample = rng.choice(len(stream), size=min(args.plain_sample, len(stream)), replace=False)
    pt, pn, pres = 0.0, 0, 0.0
    for k in sample:
        _, key, srcs = stream[k]
        t, worst, it = plain_request(L, torch, _problem_params(L, probs[key], L.LSW_PREC_FP64), srcs)
        pt += t
        pn += len(srcs)
        pres = max(pres, worst)
        log(f"  plain request {k}: {probs[key]['name']} {len(srcs)} sources {t:.2f} s (iters {it})")
        torch.cuda.empty_cache()
    plain_req_s = pt / len(sample)
    plain = dict(solves_per_s=pn / pt, s_per_request=plain_req_s, requests=len(sample), solves=pn,
                 worst_true_residual=pres)
    log("plain", plain)
    # one server for both phases: the offered load of 0.8x the plain capacity
    # (cold cache: problems load on first use), then the whole stream at once
    out = {}
    t0 = time.perf_counter()
    srv = L.Server(budget_bytes=args.serve_budget, slots=args.serve_slots, restart=args.serve_restart,
                   tol=1e-6, max_iter=200, precision=L.LSW_PREC_FP32, workers=args.serve_workers)
    pids = {key: srv.add_problem(_problem_params(L, p, L.LSW_PREC_FP32)) for key, p in probs.items()}
    log(f"serve: {len(pids)} problems registered in {time.perf_counter() - t0:.1f} s")
    prev = srv.stats()
    for label, scale, st in (("load0.8", plain_req_s / 0.8, stream[:args.serve_load_requests]),
                             ("saturated", 0.0, stream[args.serve_load_requests:])):
        res, wall = run_stream(L, srv, pids, st, t_scale=scale)
        stats = srv.stats()
        lat = [r["latency_s"] for r in res]
        bad = [r for r in res if r["status"] != 0]
        nsol = sum(r["nsrc"] for r in res)
        out[label] = dict(requests=len(res), solves=nsol, wall_s=wall, solves_per_s=nsol / wall,
                          p50_s=pct(lat, 50), p99_s=pct(lat, 99), failed=len(bad),
                          worst_true_residual=max(r["max_true_residual"] for r in res),
                          mean_interarrival_s=scale,
                          **{k: stats[k] - prev[k] for k in ("loads", "evictions", "steps", "merged_steps")},
                          max_active=stats["max_active"], peak_reserved_gb=stats["peak_reserved_bytes"] / 1e9,
                          worker_s={k: round(stats[k] - prev[k], 3) for k in ("setup_s", "factor_s", "block_s",
                                                                              "step_s")})
        prev = stats
        log(label, out[label])
        for r in bad[:5]:
            log(f"  failed request (problem {r['problem']}): status {r['status']} {r['message']}")
    srv.close()
    torch.cuda.empty_cache()
    line = {
        "metric": METRIC, "value": round(out["saturated"]["solves_per_s"], 4), "unit": "solves/s",
        "n_gpus": 1, "steps": out["saturated"]["requests"], "warmup": 0,
        "ms_per_step": round(1e3 * out["saturated"]["wall_s"] / max(1, out["saturated"]["requests"]), 3),
        "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 factors/sweeps + c128 Krylov", "data": "synthetic",
        "config": {"workload": f"serve_{M}models_x{F}freq_{len(stream)}req", "budget_gb": args.serve_budget / 1e9,
                   "slots": args.serve_slots, "restart": args.serve_restart, "workers": args.serve_workers},
        "p50_ms": round(1e3 * out["saturated"]["p50_s"], 1), "p99_ms": round(1e3 * out["saturated"]["p99_s"], 1),
        "plain": plain, "served": out,
        "speedup_served_vs_plain": round(out["saturated"]["solves_per_s"] / plain["solves_per_s"], 2),
    }
    print(json.dumps(line))


def pack_bench(args):
    """The uneven batch of 48 problems: every problem submitted at once to
    the server (FP32 factors, FP64 GMRES, continuous batching across
    problems) against the plain FP64 path one problem at a time on a
    sample."""
    import numpy as np
    import torch
    import lsweeps_amd as L
    from lsweeps_amd.workloads import config_pack48
    t0 = time.perf_counter()
    probs = config_pack48(seed=0, count=args.pack_count)
    log(f"pack: {len(probs)} problems generated in {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(2)
    sample = rng.choice(len(probs), size=min(args.plain_sample, len(probs)), replace=False)
    pt, pn, pres = 0.0, 0, 0.0
    for k in sample:
        p = probs[k]
        t, worst, it = plain_request(L, torch, _problem_params(L, p, L.LSW_PREC_FP64), p["sources"])
        pt += t
        pn += len(p["sources"])
        pres = max(pres, worst)
        log(f"  plain {p['name']}: {len(p['sources'])} sources {t:.2f} s (iters {it})")
        torch.cuda.empty_cache()
    plain = dict(solves_per_s=pn / pt, problems=len(sample), solves=pn, worst_true_residual=pres)
    log("plain", plain)
    srv = L.Server(budget_bytes=args.serve_budget, slots=16, restart=args.serve_restart, tol=1e-6, max_iter=200,
                   precision=L.LSW_PREC_FP32, workers=args.serve_workers)
    pids = [srv.add_problem(_problem_params(L, p, L.LSW_PREC_FP32)) for p in probs]
    stream = [(0.0, k, p["sources"]) for k, p in enumerate(probs)]
    res, wall = run_stream(L, srv, pids, stream)
    stats = srv.stats()
    srv.close()
    nsol = sum(r["nsrc"] for r in res)
    lat = [r["latency_s"] for r in res]
    packed = dict(problems=len(res), solves=nsol, wall_s=wall, solves_per_s=nsol / wall,
                  p50_s=pct(lat, 50), p99_s=pct(lat, 99), failed=sum(r["status"] != 0 for r in res),
                  worst_true_residual=max(r["max_true_residual"] for r in res),
                  dofs=int(sum(p["grid"][0] * p["grid"][1] for p in probs)),
                  **{k: stats[k] for k in ("loads", "evictions", "steps")},
                  peak_reserved_gb=stats["peak_reserved_bytes"] / 1e9,
                  worker_s={k: round(stats[k], 3) for k in ("setup_s", "factor_s", "block_s", "step_s")})
    log("packed", packed)
    for r in res:
        if r["status"] != 0:
            log(f"  failed request (problem {r['problem']}): status {r['status']} {r['message']}")
    line = {
        "metric": METRIC, "value": round(packed["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1,
        "steps": len(probs), "warmup": 0, "ms_per_step": round(1e3 * wall / len(probs), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 factors/sweeps + c128 Krylov", "data": "synthetic",
        "config": {"workload": f"pack{len(probs)}_seed0", "budget_gb": args.serve_budget / 1e9,
                   "workers": args.serve_workers, "restart": args.serve_restart},
        "p50_ms": round(1e3 * packed["p50_s"], 1), "p99_ms": round(1e3 * packed["p99_s"], 1),
        "plain": plain, "packed": packed,
        "speedup_packed_vs_plain": round(packed["solves_per_s"] / plain["solves_per_s"], 2),
    }
    print(json.dumps(line))


def layered_bench(args):
    """The layered 3x1 workload (2400 x 800 nodes, 12 x 4 cells, general
    geometry, seed 0) with 128 sources 10 cells below the top: the sources as
    one block through the slot engine (FP32 factors, FP64 GMRES, `slots`
    right-hand sides in flight) against the plain FP64 path one source at a
    time (a sample of sources, timed per solve).  Factorisation is reported
    separately for both."""
    import numpy as np
    import torch
    import lsweeps_amd as L
    from lsweeps_amd.workloads import config_layered_3x1
    cfg = config_layered_3x1(seed=0, nsrc=args.layered_sources)
    log(f"layered: grid {cfg['grid']}, board {cfg['ncells']}, omega {cfg['omega']:.2f}, {len(cfg['sources'])} sources")
    # plain FP64: one DD, one source at a time
    t0 = time.perf_counter()
    dd = L.DDCheckerboard(_problem_params(L, cfg, L.LSW_PREC_FP64))
    dd.factorize()
    dd.sync()
    t_fact64 = time.perf_counter() - t0
    b = torch.zeros(dd.N, dtype=torch.complex128, device="cuda:0")
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    sample = cfg["sources"][:: max(1, len(cfg["sources"]) // args.plain_sample)][: args.plain_sample]
    ts, worst, its = [], 0.0, []
    for k, src in enumerate(sample):
        dd.assemble_rhs_dev([src], b.data_ptr())
        dd.sync()
        t1 = time.perf_counter()
        info = dd.gmres_dev(b.data_ptr(), x.data_ptr(), mode=1, tol=1e-6, max_iter=200, restart=30)
        dd.sync()
        if k > 0:  # the first solve builds the plans
            ts.append(time.perf_counter() - t1)
        worst = max(worst, info["true_residual"])
        its.append(info["iterations"])
    del dd, b, x
    torch.cuda.empty_cache()
    plain = dict(solves_per_s=len(ts) / sum(ts), s_per_solve=float(np.mean(ts)), factorize_s=t_fact64,
                 sample=len(sample), iterations=its, worst_true_residual=worst)
    log("plain", plain)
    srv = L.Server(budget_bytes=args.serve_budget, slots=args.serve_slots, restart=args.serve_restart, tol=1e-6,
                   max_iter=300, precision=L.LSW_PREC_FP32, workers=1)
    pid = srv.add_problem(_problem_params(L, cfg, L.LSW_PREC_FP32))
    t0 = time.perf_counter()
    w = srv.wait(srv.submit(pid, cfg["sources"][:1], want_solution=False))  # load + factorise
    t_load = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = srv.wait(srv.submit(pid, cfg["sources"], want_solution=False))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    mixed = dict(solves_per_s=len(cfg["sources"]) / wall, wall_s=wall, load_s=t_load, status=r["status"],
                 max_iterations=r["max_iterations"], worst_true_residual=r["max_true_residual"],
                 steps=st["steps"], max_active=st["max_active"], warmup_status=w["status"])
    log("mixed block", mixed)
    n_dofs = int(np.prod(cfg["grid"]))
    line = {
        "metric": METRIC, "value": round(mixed["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(1e3 * wall, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 factors/sweeps + c128 Krylov", "data": "synthetic",
        "config": {"workload": cfg["name"], "grid": list(cfg["grid"]), "board": list(cfg["ncells"]),
                   "sources": len(cfg["sources"]), "slots": args.serve_slots, "restart": args.serve_restart,
                   "geometry": "general"},
        "plain": plain, "mixed_block": mixed,
        "speedup_vs_plain": round(mixed["solves_per_s"] / plain["solves_per_s"], 2),
        "dofs": n_dofs,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lens1600")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip timing alternative launch shapes per DD and block width before the timed solves")
    ap.add_argument("--paths", default="plain_ref,plain,mixed,batched")
    ap.add_argument("--serve-models", type=int, default=8)
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-load-requests", type=int, default=6)
    ap.add_argument("--serve-budget", type=float, default=240e9)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-restart", type=int, default=8)
    ap.add_argument("--serve-workers", type=int, default=6)
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--pack-count", type=int, default=48)
    ap.add_argument("--layered-sources", type=int, default=128)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with one 1 GiB device-to-device copy (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(sharding across GPUs is design-only in this build)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serve":
        serve_bench(args)
        return
    if args.workload == "pack48":
        pack_bench(args)
        return
    if args.workload == "layered":
        layered_bench(args)
        return
    import numpy as np
    import torch
    import lsweeps_amd as L
    from lsweeps_amd.workloads import CONFIGS

    cfg = CONFIGS[args.workload]()
    paths = args.paths.split(",")
    log(f"workload {cfg['name']}: grid {cfg['grid']}, board {cfg['ncells']}, omega {cfg['omega']:.3f}")
    out, xs, prof, tune = {}, {}, {}, {}
    if "plain_ref" in paths or "plain" in paths:
        dd, ts, tf = make_dd(L, cfg, L.LSW_PREC_FP64)
        log(f"FP64 setup {ts:.2f} s, factorize {tf:.3f} s")
        if args.autotune:
            t0 = time.perf_counter()
            tune["fp64_x1"] = dd.autotune(1)
            log(f"  autotune {time.perf_counter() - t0:.1f} s", tune)
        if "plain_ref" in paths:
            out["plain_ref"], xs["plain_ref"] = time_single(dd, torch, cfg, 0, cfg["tol"], cfg["max_iter"],
                                                            args.steps, args.warmup)
            out["plain_ref"].update(factorize_s=tf, setup_s=ts, gmres="reference", tol=cfg["tol"])
            log("  plain_ref", out["plain_ref"])
        if "plain" in paths:
            out["plain"], xs["plain"] = time_single(dd, torch, cfg, 1, 1e-6, 60, args.steps, args.warmup)
            out["plain"].update(factorize_s=tf, setup_s=ts, gmres="flexible", tol=1e-6)
            log("  plain", out["plain"])
        prof["fp64"] = profile(dd, torch)
        st64 = dd.stats()
        del dd
        torch.cuda.empty_cache()
    if "mixed" in paths or "batched" in paths:
        dd, ts, tf = make_dd(L, cfg, L.LSW_PREC_FP32)
        log(f"FP32 setup {ts:.2f} s, factorize {tf:.3f} s")
        if args.autotune:
            t0 = time.perf_counter()
            tune["fp32_x1"] = dd.autotune(1)
            if "batched" in paths:
                tune["fp32_x16"] = dd.autotune(16)
            log(f"  autotune {time.perf_counter() - t0:.1f} s", tune)
        if "mixed" in paths:
            out["mixed"], xs["mixed"] = time_single(dd, torch, cfg, 1, 1e-6, 60, args.steps, args.warmup)
            out["mixed"].update(factorize_s=tf, setup_s=ts, gmres="flexible", tol=1e-6)
            log("  mixed", out["mixed"])
        if "batched" in paths:
            out["batched"] = time_batched(dd, torch, cfg, 1e-6, max(1, args.steps // 2), min(args.warmup, 1))
            log("  batched", out["batched"])
        prof["fp32"] = profile(dd, torch)
        st32 = dd.stats()
        del dd
        torch.cuda.empty_cache()
    head_name = next((k for k in ("mixed", "plain", "plain_ref") if k in out), None)
    if head_name is None:
        log("no single-source path in --paths: results above, no bench line")
        return
    head = out[head_name]
    roof = roofline(prof["fp32"] if "fp32" in prof else prof["fp64"])
    cpu = None
    if not args.no_cpu_baseline:
        iters = out["plain_ref"]["iterations"] if "plain_ref" in out else head["iterations"]
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(cfg, iters)
        log("  ", cpu)
    n_dofs = int(np.prod(cfg["grid"]))
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
        "dtype": "c64 factors/sweeps + c128 Krylov" if head_name == "mixed" else "c128",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "grid": list(cfg["grid"]), "board": list(cfg["ncells"]),
                   "omega": round(cfg["omega"], 4), "sources": len(cfg["sources"]), "path": head_name,
                   "residual": "true relative residual <= 1e-6"},
        "p50_ms": round(head["p50_ms"], 3),
        "p99_ms": round(head["p99_ms"], 3),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    for k, v in out.items():
        line[k] = {kk: (float(f"{vv:.6g}") if isinstance(vv, float) else vv) for kk, vv in v.items()}
    if "fp64" in prof:
        line["roofline_fp64"] = roofline(prof["fp64"])
    if "plain" in out:
        base = out["plain"]["solves_per_s"]
        for k in ("mixed", "batched"):
            if k in out:
                line[f"speedup_{k}_vs_plain"] = round(out[k]["solves_per_s"] / base, 3)
    if "mixed" in xs and "plain" in xs:
        line["mixed_vs_plain_rel_l2"] = float(np.linalg.norm(xs["mixed"] - xs["plain"]) /
                                              np.linalg.norm(xs["plain"]))
    if "plain_ref" in xs and "plain" in xs:
        line["plain_ref_vs_plain_rel_l2"] = float(np.linalg.norm(xs["plain_ref"] - xs["plain"]) /
                                                  np.linalg.norm(xs["plain"]))
    if tune:
        line["launch_shapes"] = tune
    line["dofs"] = n_dofs
    line["dofs_per_s"] = round(n_dofs * head["solves_per_s"], 1)
    traffic = load_traffic(cfg["name"])
    if traffic:
        line["roofline"]["traffic"] = traffic["bytes_per_launch"]
        line["roofline"]["traffic_source"] = traffic["source"]
    print(json.dumps(line))
    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
