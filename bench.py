#!/usr/bin/env python3
"""Benchmark of the L-sweeps preconditioned Helmholtz solve on one MI355X.

Metric: solves/s and p50/p99 latency at true relative residual <= 1e-6,
mixed precision / batched against plain FP64.

Default line (python bench.py): the uneven batch of 48 problems from seed 0
(lsweeps_amd.workloads.config_pack48).
  value   mixed precision (FP32 factors and sweeps, FP64 flexible GMRES), every
          problem in flight at once through the continuous-batching server.
          A step is one pass over the pack: a new server (empty factor
          cache), every problem registered, set up, factorised and all its
          sources solved; --warmup passes untimed, then --steps passes timed.
  plain   the plain FP64 path, one problem at a time as the reference runs a
          problem (set-up, factorisation, then each source by FP64 flexible
          GMRES), over all 48 problems.
Beside it: the 1600^2 lens workload (plain_ref = FP64 with the reference
GMRES, plain, mixed, a 16-source block), the roofline of the batched local
solves, and the CPU baselines (the oracle on host processes).

Other lines: --workload lens1600 | const400 (one problem), pack48 (the pack
alone), layered (2400x800, 128 sources as one block), serve (a Poisson
request stream).
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


def profile(dd, torch, nrhs=1):
    b = torch.randn(nrhs * dd.N, dtype=torch.complex128, device="cuda:0")
    y = torch.zeros_like(b)
    torch.cuda.synchronize()
    return dd.profile_precond_dev(b.data_ptr(), y.data_ptr(), nrhs=nrhs)


def roofline(prof):
    """HBM roofline of the batched local solves of one M^-1 application
    (algorithmic bytes over the summed HIP-event launch times), with the
    flops beside it: frac_combined = summed per-launch roofline time (the
    larger of its bytes at 8 TB/s and its flops at the matrix-core peak) over
    the summed launch time."""
    ms = prof["fwd_ms"] + prof["bwd_ms"]
    nl = prof["fwd_launches"] + prof["bwd_launches"]
    byts = prof["fwd_bytes"] + prof["bwd_bytes"]
    flops = prof.get("fwd_flops", 0.0) + prof.get("bwd_flops", 0.0)
    achieved = (byts / nl) / (ms / nl * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
           "kernel": "k_nd_packed + k_nd_wide (batched ND local solves, one M^-1 application)",
           "bytes_per_launch": round(byts / nl, 1), "ms_per_launch": round(ms / nl, 5),
           "launches_per_precond": nl, "fwd": prof}
    if flops > 0 and ms > 0:
        out["flops_per_launch"] = round(flops / nl, 1)
        out["achieved_tflops"] = round(flops / (ms * 1e-3) / 1e12, 2)
        out["roofline_ms"] = round(prof["ideal_ms"], 4)
        out["frac_combined"] = round(prof["ideal_ms"] / ms, 4)
    return out


def load_traffic(workload):
    """HBM bytes per solve launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (profiles/traffic_<workload>.json), if present."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(workload, procs=None, **kw):
    """The oracle's restatement timed on P host processes (oracle/cpu_run.py,
    run as a child process so that nothing forks this GPU process): the
    reference's "factorize" and "GMRES" timers plus set-up, on a bounded
    sample of the workload.  P = the host cores this job may use, at most 16
    (one GPU's share of the box)."""
    import subprocess
    if procs is None:
        try:
            procs = len(os.sched_getaffinity(0))
        except AttributeError:
            procs = os.cpu_count() or 1
        procs = max(1, min(16, procs))
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_run.py"), "--workload", workload, "--procs", str(procs)]
    for k, v in kw.items():
        cmd += [f"--{k}", str(v)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


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
    time on a sample of the stream (its capacity in requests/s), then one
    server (FP32 factors, FP64 GMRES, continuous batching, 288 GB admission
    budget) taking the first --serve-load-requests requests as a Poisson
    stream at 0.8x that capacity (value, p50/p99), and the rest as one burst
    (saturation, reported separately)."""
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
        # latency at the offered load of 0.8x the plain capacity (Poisson
        # arrivals); throughput (value) of the saturated burst
        "p50_ms": round(1e3 * out["load0.8"]["p50_s"], 1), "p99_ms": round(1e3 * out["load0.8"]["p99_s"], 1),
        "latency_phase": f"Poisson stream of {out['load0.8']['requests']} requests at 0.8x the plain capacity",
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
                   precision=L.LSW_PREC_FP32, workers=args.serve_workers, batch_order=True)
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


def lens_section(args, L, torch, cfg, paths):
    """The single-problem figures (plain_ref, plain, mixed, batched) of one
    workload with per-launch solve profiles; returns (results, solutions,
    profiles, launch shapes)."""
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
        if "batched" in paths:
            prof["fp32_x16"] = profile(dd, torch, BATCH)
        del dd
        torch.cuda.empty_cache()
    return out, xs, prof, tune


def lens_fields(cfg, out, xs, prof, tune):
    """The single-problem figures as JSON fields."""
    import numpy as np
    f = {}
    for k, v in out.items():
        f[k] = {kk: (float(f"{vv:.6g}") if isinstance(vv, float) else vv) for kk, vv in v.items()}
    if "fp32" in prof or "fp64" in prof:
        f["roofline"] = roofline(prof["fp32"] if "fp32" in prof else prof["fp64"])
        traffic = load_traffic(cfg["name"])
        if traffic and "fp32" in prof:
            f["roofline"]["traffic"] = traffic["bytes_per_launch"]
            f["roofline"]["traffic_source"] = traffic["source"]
    if "fp64" in prof:
        f["roofline_fp64"] = roofline(prof["fp64"])
    if "fp32_x16" in prof:
        f["roofline_block16"] = roofline(prof["fp32_x16"])
        f["roofline_block16"]["kernel"] = ("k_nd_mfma + k_nd_mfma_wide (batched ND local solves of one M^-1 "
                                           "application to a block of 16 sources: FP32 tasks of 8 right-hand "
                                           "sides on the matrix cores)")
        traffic = load_traffic(cfg["name"] + "_block16")
        if traffic:
            f["roofline_block16"]["traffic"] = traffic["bytes_per_launch"]
            f["roofline_block16"]["traffic_source"] = traffic["source"]
    if "plain" in out:
        base = out["plain"]["solves_per_s"]
        for k in ("mixed", "batched"):
            if k in out:
                f[f"speedup_{k}_vs_plain"] = round(out[k]["solves_per_s"] / base, 3)
    if "mixed" in xs and "plain" in xs:
        f["mixed_vs_plain_rel_l2"] = float(np.linalg.norm(xs["mixed"] - xs["plain"]) / np.linalg.norm(xs["plain"]))
    if "plain_ref" in xs and "plain" in xs:
        f["plain_ref_vs_plain_rel_l2"] = float(np.linalg.norm(xs["plain_ref"] - xs["plain"]) /
                                               np.linalg.norm(xs["plain"]))
    if tune:
        f["launch_shapes"] = tune
    f["dofs"] = int(np.prod(cfg["grid"]))
    return f


def lens_bench(args):
    """One workload's single-problem line (--workload lens1600 or const400)."""
    import torch
    import lsweeps_amd as L
    from lsweeps_amd.workloads import CONFIGS
    cfg = CONFIGS[args.workload]()
    out, xs, prof, tune = lens_section(args, L, torch, cfg, args.paths.split(","))
    head_name = next((k for k in ("mixed", "plain", "plain_ref") if k in out), None)
    if head_name is None:
        log("no single-source path in --paths: results above, no bench line")
        return
    head = out[head_name]
    cpu = None
    if not args.no_cpu_baseline:
        iters = out["plain_ref"]["iterations"] if "plain_ref" in out else head["iterations"]
        log("cpu baseline (oracle on host processes) ...")
        cpu = cpu_baseline("lens1600" if cfg["name"].startswith("lens") else "const400", iterations=iters)
        log("  ", cpu)
    line = {
        "metric": METRIC, "value": round(head["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["solve_s"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 factors/sweeps + c128 Krylov" if head_name == "mixed" else "c128", "data": "synthetic",
        "config": {"workload": cfg["name"], "grid": list(cfg["grid"]), "board": list(cfg["ncells"]),
                   "omega": round(cfg["omega"], 4), "sources": len(cfg["sources"]), "path": head_name,
                   "residual": "true relative residual <= 1e-6"},
        "p50_ms": round(head["p50_ms"], 3), "p99_ms": round(head["p99_ms"], 3),
    }
    line.update(lens_fields(cfg, out, xs, prof, tune))
    line["cpu_baseline"] = _cpu_object(cpu) if cpu else None
    line["dofs_per_s"] = round(line["dofs"] * head["solves_per_s"], 1)
    print(json.dumps(line))
    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()


def _cpu_object(r):
    """The cpu_baseline object of the bench line from a cpu_run result."""
    if "gmres_s" in r and "iterations" in r and "solves" not in r:
        return {"value": round(r["solves_per_s"], 6), "unit": "solves/s", "cores": r["procs"], "kind": "port",
                "sample": (f"{r['workload']}: the oracle's restatement (SuperLU local solves, reference-mode GMRES) "
                           f"on {r['procs']} processes ("
                           f"{'one per board row, the reference rank layout' if r.get('assign') == 'rows' else 'cells dealt round-robin'}"
                           f") of {r['host_cpus']} host CPUs ({r['cpu']}); measured: "
                           f"{r['measured']}; value = 1 / GMRES seconds"),
                "factorize_s": round(r["factorize_s"], 3), "gmres_s": round(r["gmres_s"], 3),
                "setup_s": round(r["setup_s"], 3), "end_to_end_s": round(r["end_to_end_s"], 3),
                "iterations": r["iterations"], "true_residual": r.get("true_residual")}
    return {"value": round(r["solves_per_s"], 6), "unit": "solves/s", "cores": r.get("procs"), "kind": "port",
            "sample": (f"the oracle's restatement (SuperLU local solves, flexible GMRES) on {r.get('procs')} "
                       f"processes of {r['host_cpus']} host CPUs ({r['cpu']}): {r['measured']}"),
            "solves": r["solves"], "seconds": round(r["seconds"], 3), "factorize_s": round(r["factorize_s"], 3),
            "gmres_s": round(r["gmres_s"], 3), "setup_s": round(r["setup_s"], 3)}


def pack_pass(L, probs, workers, restart, budget, slots=16):
    """One pass over the uneven pack: a new server (empty factor cache), every
    problem registered and its request submitted, every result waited for.
    Returns (wall seconds, per-request results in pack order, server stats).

    Registration (host geometry and the device-bytes estimate) runs on a few
    threads (the library call releases the GIL), largest problems (grid nodes
    x sources) first, and each request is submitted as soon as its problem is
    registered: the workers start on the largest problems while the rest
    register (the server itself then orders by estimated work)."""
    from concurrent.futures import ThreadPoolExecutor
    order = sorted(range(len(probs)), key=lambda k: -probs[k]["grid"][0] * probs[k]["grid"][1] * len(probs[k]["sources"]))
    t0 = time.perf_counter()
    srv = L.Server(budget_bytes=budget, slots=slots, restart=restart, tol=1e-6, max_iter=200,
                   precision=L.LSW_PREC_FP32, workers=workers, batch_order=True)
    try:
        def register(k):
            pid = srv.add_problem(_problem_params(L, probs[k], L.LSW_PREC_FP32))
            return srv.submit(pid, probs[k]["sources"], want_solution=False)

        with ThreadPoolExecutor(8) as ex:
            ids = dict(zip(order, ex.map(register, order)))
        res = [srv.wait(ids[k]) for k in range(len(probs))]
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.close()
    return wall, res, st


def north_star_bench(args):
    """The default line: the 48-problem uneven pack (mixed precision, packed
    and served by continuous batching) against the plain FP64 path one problem
    at a time, with the 1600^2 lens figures, the solve kernels' roofline and
    the CPU baselines beside it."""
    import numpy as np
    import torch
    import lsweeps_amd as L
    from lsweeps_amd.workloads import config_pack48, config_lens_1600
    probs = config_pack48(seed=0, count=args.pack_count)
    nsrc = sum(len(p["sources"]) for p in probs)
    log(f"pack: {len(probs)} problems, {nsrc} sources, {sum(p['grid'][0] * p['grid'][1] for p in probs)} grid nodes")
    # plain FP64: every problem as the reference runs one (set-up, factorise,
    # then each source on its own), one after the other
    t0 = time.perf_counter()
    pt, pres, pits = 0.0, 0.0, []
    for p in ([] if args.pack_passes_only else probs):
        t, worst, it = plain_request(L, torch, _problem_params(L, p, L.LSW_PREC_FP64), p["sources"])
        pt += t
        pres = max(pres, worst)
        pits.append(it)
    torch.cuda.empty_cache()
    if not args.pack_passes_only:
        plain = dict(solves_per_s=nsrc / pt, seconds=pt, problems=len(probs), solves=nsrc,
                     worst_true_residual=pres, max_iterations=max(pits),
                     gmres="flexible FP64, restart 30, one source at a time")
        log("plain", plain, f"({time.perf_counter() - t0:.1f} s)")
    # mixed, packed: W untimed passes, then K timed passes
    for w in range(args.warmup):
        wall, _, _ = pack_pass(L, probs, args.pack_workers, args.serve_restart, args.serve_budget)
        log(f"  warm-up pass {w}: {wall:.2f} s")
    walls, lat, worst, failed, itmax = [], [], 0.0, 0, 0
    st = None
    for k in range(args.steps):
        wall, res, st = pack_pass(L, probs, args.pack_workers, args.serve_restart, args.serve_budget)
        walls.append(wall)
        lat += [r["latency_s"] for r in res]
        failed += sum(r["status"] != 0 for r in res)
        worst = max(worst, max(r["max_true_residual"] for r in res))
        itmax = max(itmax, max(r["max_iterations"] for r in res))
        log(f"  pass {k}: {wall:.2f} s = {nsrc / wall:.2f} solves/s")
    total = sum(walls)
    packed = dict(solves_per_s=args.steps * nsrc / total, seconds_per_pass=total / args.steps, passes=args.steps,
                  failed=failed, worst_true_residual=worst, max_iterations=itmax,
                  loads=st["loads"], evictions=st["evictions"], steps=st["steps"], merged_steps=st["merged_steps"],
                  worker_s={k: round(st[k], 3) for k in ("setup_s", "factor_s", "block_s", "step_s")},
                  workers=args.pack_workers, restart=args.serve_restart, budget_gb=args.serve_budget / 1e9)
    log("packed", packed)
    if args.pack_passes_only:
        # (a profiling run of the packed passes alone: not the bench line)
        print(json.dumps({"workload": f"pack{len(probs)}_seed0", "packed": packed,
                          "p50_ms": round(1e3 * pct(lat, 50), 1), "p99_ms": round(1e3 * pct(lat, 99), 1)}))
        return
    # the 1600^2 lens workload: single-problem figures and the solve roofline
    cfg = config_lens_1600()
    out, xs, prof, tune = lens_section(args, L, torch, cfg, args.paths.split(","))
    lens = lens_fields(cfg, out, xs, prof, tune)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baselines (oracle on host processes) ...")
        cpu = cpu_baseline("pack48", budget=args.cpu_budget)
        log("  pack48", cpu)
        iters = out["plain_ref"]["iterations"] if "plain_ref" in out else 7
        cpu_lens = cpu_baseline("lens1600", iterations=iters)
        log("  lens1600", cpu_lens)
        lens["cpu_baseline"] = _cpu_object(cpu_lens)
        # the reference's own rank layout: one process per checkerboard row
        rows = int(cfg["ncells"][0])
        cpu_rows = cpu_baseline("lens1600", procs=rows, iterations=iters, assign="rows")
        log("  lens1600, one process per row", cpu_rows)
        lens["cpu_baseline_rows"] = _cpu_object(cpu_rows)
    line = {
        "metric": METRIC, "value": round(packed["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 factors/sweeps + c128 Krylov", "data": "synthetic",
        "config": {"workload": f"pack{len(probs)}_seed0", "problems": len(probs), "solves": nsrc,
                   "grid_nodes": int(sum(p["grid"][0] * p["grid"][1] for p in probs)),
                   "step": "one pass over the pack: a new server, every problem registered, set up, factorised "
                           "and all its sources solved",
                   "path": "FP32 factors and sweeps, FP64 flexible GMRES, continuous batching",
                   "residual": "true relative residual <= 1e-6"},
        "p50_ms": round(1e3 * pct(lat, 50), 1), "p99_ms": round(1e3 * pct(lat, 99), 1),
        "speedup_vs_plain": round(packed["solves_per_s"] / plain["solves_per_s"], 3),
        "plain": plain, "packed": packed,
        "roofline": None,
        "cpu_baseline": _cpu_object(cpu) if cpu else None,
        "lens1600": lens,
    }
    # the dominant kernel of the packed path: the FP32 8-right-hand-side
    # solve tasks on the matrix cores, measured on one 16-source application
    # on the 1600^2 lens; the single-source figure beside it
    single, block = lens.pop("roofline", None), lens.pop("roofline_block16", None)
    if single is not None:
        single["workload"] = "the batched ND solves of one single-source M^-1 application on the 1600^2 lens (FP32)"
    if block is not None:
        block["workload"] = ("the batched ND solves of one M^-1 application to 16 sources on the 1600^2 lens "
                             "(FP32), the launch shape the packed and served paths run")
        line["roofline"], line["roofline_single"] = block, single
    else:
        line["roofline"] = single
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="default",
                    help="default (the 48-problem pack with the lens figures), lens1600, const400, pack48, "
                         "layered, serve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip timing alternative launch shapes per DD and block width before the timed solves")
    ap.add_argument("--paths", default="plain_ref,plain,mixed,batched")
    ap.add_argument("--pack-workers", type=int, default=12)
    ap.add_argument("--pack-passes-only", action="store_true",
                    help="default workload: only the packed passes (for profiling; prints no bench line)")
    ap.add_argument("--serve-models", type=int, default=8)
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-load-requests", type=int, default=32)
    ap.add_argument("--serve-budget", type=float, default=None,
                    help="device bytes the server may reserve (default: 288e9 for the serving stream, the HBM "
                         "the serving workload states; 240e9 for the pack)")
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-restart", type=int, default=10)
    ap.add_argument("--serve-workers", type=int, default=6)
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--pack-count", type=int, default=48)
    ap.add_argument("--layered-sources", type=int, default=128)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with one 1 GiB device-to-device copy (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.serve_budget is None:
        args.serve_budget = 288e9 if args.workload == "serve" else 240e9
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(sharding across GPUs is design-only in this build)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "default":
        north_star_bench(args)
    elif args.workload == "serve":
        serve_bench(args)
    elif args.workload == "pack48":
        pack_bench(args)
    elif args.workload == "layered":
        layered_bench(args)
    else:
        lens_bench(args)


if __name__ == "__main__":
    main()
