#!/usr/bin/env python3
"""Benchmark of the WCSPH step engine on one MI355X.

One step = one Integrator::integrate (src/Integration.cpp:233-303) over the
whole particle set: 2 neighbour builds, 2 x (dSPH_PreStep, Detect_Surface,
dissipation_terms, particle_shift) and every Newmark-Beta sub-iteration
(force + update + residual) until convergence.  The metric is particle-steps
per second = particles x steps / wall time, with the particle state resident
in HBM when the timed region starts.

Prints ONE JSON line with the metric, the roofline of the dominant kernel
(the fused force + Newmark update) and the CPU baseline (the FP64 oracle
restatement built with the reference's compiler flags, timed on the host
cores on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fjsph_amd import abi, cases  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s and %HBM roofline at 4M 3D particles; surrogate vs mesh aero"


def workload(name, n_target):
    if name == "jet4m":
        c = cases.jet_column_3d(n_target=n_target, precision=abi.PREC_MIXED)
        desc = ("3D coflow jet column, D=2 mm liquid jet at 10 m/s in 30 m/s coflow, dx=0.04 mm, "
                "H=2dx, ALE shifting on, Gissler aero from the constant coflow velocity "
                "(mesh and surrogate aero sources not built yet)")
        return c, desc
    if name == "droplet":
        return cases.droplet_3d(precision=abi.PREC_MIXED), "3D droplet D=4 mm, dx=0.05 mm, 20 m/s crossflow"
    if name == "dam2d":
        return cases.dam_break_2d(), "2D dam break 1x2 m column, dx=0.02 m, FP64"
    raise SystemExit(f"unknown workload {name}")


def load_traffic():
    """HBM bytes per force launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE) and committed as profiles/force_update_traffic.json."""
    path = os.path.join(ROOT, "profiles", "force_update_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline(args):
    """FP64 oracle restatement (reference flags -O3 -ffast-math, OpenMP) on a
    bounded sample of the same workload: start-up pass + 2 untimed steps,
    then `cpu_steps` timed steps."""
    from oracle.pyoracle import Oracle
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if args.workload == "jet4m":
        c = cases.jet_column_3d(n_target=args.cpu_particles, precision=abi.PREC_FP64)
    else:
        c, _ = workload(args.workload, args.particles)
    o = Oracle(c, nthreads=nthreads, kind="baseline")
    o.integrate_no_update()
    for _ in range(2):
        o.integrate_step()
    its = []
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        its.append(o.integrate_step()["iterations"])
    dt = time.perf_counter() - t0
    o.close()
    return {
        "value": c.n * args.cpu_steps / dt,
        "unit": "particle-steps/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"{c.name} with {c.n} particles (same dx, H and physics), FP64, steps 3-{2 + args.cpu_steps} "
                   f"after start-up, {args.cpu_steps} timed steps, mean {np.mean(its):.1f} sub-iterations/step; "
                   "cell-list neighbour search, serial dissipation_terms as in the reference"),
        "seconds": dt,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="jet4m")
    ap.add_argument("--particles", type=int, default=4_000_000)
    ap.add_argument("--cpu-particles", type=int, default=60_000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the path shards by "
              "spatial domain decomposition, which is described but not built.", file=sys.stderr)
        sys.exit(2)

    from fjsph_amd.engine import Engine

    case, desc = workload(args.workload, args.particles)
    print(f"[bench] workload {case.name}: {case.n} particles; building engine", file=sys.stderr, flush=True)
    eng = Engine(case)
    eng.integrate_no_update()
    for k in range(args.warmup):
        eng.integrate_step()
    print(f"[bench] warm-up done ({args.warmup} steps)", file=sys.stderr, flush=True)

    eng.enable_timers(True)
    its = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        info = eng.integrate_step()
        its.append(info["iterations"])
    elapsed = time.perf_counter() - t0
    timers = eng.timers()
    pairs, max_count = eng.neighbour_stats()
    value = case.n * args.steps / elapsed

    # roofline of the dominant kernel: fused force + Newmark update
    ms, launches = timers["force_update"]
    avg_s = (ms / 1e3) / max(launches, 1)
    bpp = eng.pass_bytes(abi.TIMER_NAMES.index("force_update"))
    alg_bytes = bpp * case.n
    achieved = alg_bytes / avg_s / 1e9
    traffic = load_traffic()
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": (traffic or {}).get("bytes_per_launch"),
        "kernel": "k_force (fused get_acc_and_Rrho + Newmark update)",
        "bytes_per_particle": bpp,
        "avg_launch_ms": avg_s * 1e3,
        "launches": launches,
    }
    pass_ms = {k: round(v[0] / args.steps, 3) for k, v in timers.items()}

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms, fp64 accumulation/integration",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "particles": case.n,
            "mean_neighbours": pairs / case.n + 1,
            "max_neighbours": max_count + 1,
            "subiterations_per_step": float(np.mean(its)),
            "pass_ms_per_step": pass_ms,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["config"]["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
