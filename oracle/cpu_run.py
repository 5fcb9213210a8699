"""CPU baseline: the oracle's restatement of the reference path, timed on P
host processes.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): bench.py runs this module
as a child process (python oracle/cpu_run.py ...) for its cpu_baseline
figures.  It prints one JSON object.

The timers follow Problem::solve (src/Problem.cpp:101-145): "factorize" is
DDCheckerboard::factorize on every rank at once (here: every worker process
factorises its cells with SuperLU, the wall time until all are done);
"GMRES" is GlobalIterativeSolverGMRES::solve (the reference-mode restatement
of oracle/gmres.py, with the local solves of each anti-diagonal on the
workers).  Set-up (geometry and assembly) is reported beside them, as the
reference leaves it untimed.

Workloads:
  const400  the 400 x 400, 4 x 4 board: the whole reference run (set-up,
            factorize, GMRES to its tolerance) is measured.
  lens1600  the 1600 x 1600 lens, 8 x 8 board: set-up, factorize and one
            preconditioner application are measured; the GMRES time is that
            application times (iterations + 2), the reference's application
            count for `--iterations` GMRES steps (plus one matvec per step).
  pack48    the 48 uneven problems: problems of the pack, smallest first,
            each solved completely (set-up, factorize, every source by
            flexible GMRES to true residual 1e-6) until --budget seconds
            have passed; solves/s over that sample.
"""
import os

for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"   # one thread per process: P processes = P cores

import argparse
import importlib.util
import json
import math
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle.geometry import Checkerboard, ConstantMedium, CSVMedium  # noqa: E402
from oracle.gmres import fgmres, gmres_reference  # noqa: E402
from oracle.parallel import CellPool  # noqa: E402
from oracle.sweeps import Preconditioner  # noqa: E402


def _workloads():
    # the synthetic problem generators, loaded by path (the package itself
    # would load the HIP library)
    path = os.path.join(ROOT, "l-sweeps-2d_amd", "lsweeps_amd", "workloads.py")
    spec = importlib.util.spec_from_file_location("_lsw_workloads", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _checkerboard(cfg):
    med = ConstantMedium() if cfg["velocity"] is None else CSVMedium(cfg["velocity"],
                                                                      [list(d) for d in cfg["vel_domain"]])
    return Checkerboard([list(cfg["domain"][0]), list(cfg["domain"][1])], cfg["ncells"], cfg["nels"], cfg["omega"],
                        cfg["min_wl_in_pml"], med, general=cfg.get("general", False))


def run_reference(cfg, procs, assign, full, iterations=None):
    t0 = time.perf_counter()
    cb = _checkerboard(cfg)
    A = cb.global_operator().tocsr()
    b = cb.global_rhs(cfg["sources"])
    t_setup = time.perf_counter() - t0
    out = dict(setup_s=t_setup, cells=len(cb.cells), dofs=int(cb.N))
    with CellPool(cb, procs, assign) as pool:
        out["procs"] = pool.procs
        out["factorize_s"] = pool.factor_s
        P = Preconditioner(cb, solver=pool)
        if full:
            t1 = time.perf_counter()
            x, log = gmres_reference(lambda v: A @ v, P.apply_global, b, tol=cfg["tol"], max_iter=cfg["max_iter"])
            out["gmres_s"] = time.perf_counter() - t1
            out["iterations"] = len(log) - 1
            out["true_residual"] = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
            out["measured"] = "set-up, factorize and the whole reference-mode GMRES"
        else:
            t1 = time.perf_counter()
            P.apply_global(b)
            t_apply = time.perf_counter() - t1
            t1 = time.perf_counter()
            A @ b
            t_mv = time.perf_counter() - t1
            it = int(iterations)
            out["precond_apply_s"] = t_apply
            out["matvec_s"] = t_mv
            out["iterations"] = it
            out["gmres_s"] = (it + 2) * t_apply + (it + 1) * t_mv
            out["measured"] = (f"set-up, factorize and one preconditioner application ({P.nsolves} local solves); "
                               f"GMRES = {it} + 2 applications + {it} + 1 matvecs")
        out["solve_count_per_apply"] = 4 * (3 * max(cb.nR, cb.nC) - 2) ** 2 if cb.nR == cb.nC else None
    out["end_to_end_s"] = out["setup_s"] + out["factorize_s"] + out["gmres_s"]
    out["solves_per_s"] = 1.0 / out["gmres_s"]
    return out


def run_pack(probs, procs, budget):
    order = sorted(range(len(probs)), key=lambda k: probs[k]["grid"][0] * probs[k]["grid"][1])
    t_all = time.perf_counter()
    done, nsol, worst = [], 0, 0.0
    t_f = t_s = t_g = 0.0
    for k in order:
        p = probs[k]
        t0 = time.perf_counter()
        cb = _checkerboard(p)
        A = cb.global_operator().tocsr()
        t_s += time.perf_counter() - t0
        with CellPool(cb, procs) as pool:
            t_f += pool.factor_s
            P = Preconditioner(cb, solver=pool)
            t1 = time.perf_counter()
            here = 0
            for s in p["sources"]:
                b = cb.global_rhs([s])
                x, it, _ = fgmres(lambda v: A @ v, P.apply_global, b, tol=1e-6, restart=30, max_iter=200)
                worst = max(worst, float(np.linalg.norm(b - A @ x) / np.linalg.norm(b)))
                nsol += 1
                here += 1
                if time.perf_counter() - t_all >= budget:
                    break
            t_g += time.perf_counter() - t1
        done.append(f"{p['name']} ({here} of {len(p['sources'])} sources)" if here < len(p["sources"])
                    else p["name"])
        if time.perf_counter() - t_all >= budget:
            break
    total = time.perf_counter() - t_all
    return dict(problems=done, procs=procs, solves=nsol, seconds=total, setup_s=t_s, factorize_s=t_f, gmres_s=t_g,
                solves_per_s=nsol / total, worst_true_residual=worst,
                measured=f"problems of the {len(probs)}-problem pack, smallest first, each set up and factorised, "
                         f"then its sources solved one at a time by flexible GMRES to true residual 1e-6, until "
                         f"{budget:.0f} s had passed: {', '.join(done)}; the smallest problems are the cheapest per "
                         f"source, so this overstates the CPU rate over the whole pack")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True, choices=["const400", "lens1600", "pack48"])
    ap.add_argument("--procs", type=int, default=16)
    ap.add_argument("--assign", default="cells", choices=["cells", "rows"])
    ap.add_argument("--iterations", type=int, default=7)
    ap.add_argument("--budget", type=float, default=20.0)
    args = ap.parse_args()
    W = _workloads()
    t0 = time.perf_counter()
    if args.workload == "const400":
        out = run_reference(W.config_const_400(), args.procs, args.assign, full=True)
    elif args.workload == "lens1600":
        out = run_reference(W.config_lens_1600(), args.procs, args.assign, full=False, iterations=args.iterations)
    else:
        out = run_pack(W.config_pack48(seed=0), args.procs, args.budget)
    out.update(workload=args.workload, assign=args.assign, cpu=_cpu_model(), host_cpus=os.cpu_count(),
               wall_s=time.perf_counter() - t0)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
