#!/usr/bin/env python3
"""Benchmark of the WCSPH step engine on one MI355X.

One step = one Integrator::integrate (src/Integration.cpp:233-303) over the
whole particle set: 2 neighbour builds, 2 x (dSPH_PreStep, Detect_Surface,
dissipation_terms, particle_shift) and every Newmark-Beta sub-iteration
(force + update + residual) until convergence.  The metric is particle-steps
per second = particles x steps / wall time, with the particle state resident
in HBM when the timed region starts.

Prints ONE JSON line with the metric, the roofline of the dominant kernel
(k_force_m, the pair sums of get_acc_and_Rrho), per-pass timings and the CPU
baseline (the FP64 restatement built with the reference's compiler flags,
timed on the host cores on a bounded sample of the same workload).

`--copy-probe BYTES` instead copies a device buffer of BYTES bytes a few
times and exits: the known-size kernel that scales the HBM counters of
rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE).
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
FP32_VECTOR_PEAK_TFLOPS = 157.3


def force_flops_per_pair(dim, ale):
    """FP32 operations of one pair in k_force_m (FMA = 2) as it evaluates
    them, frozen-distance form: R_ji and v_ji with the target's rounding
    residual (4D), build-position difference (2D), r^2 (2D-1), sqrt (1), V_j
    (1), 1/(r^2+eps) (2), gradient (D), pressure factor (2), r.g and v.g
    (2(2D-1)), viscous factor (5).  Without ALE: W'/r (6), surface tension
    (6, every fluid pair), momentum sums (6D) and the continuity sum (2).
    With ALE: W'/r folded to (kc - kc r/2H)^3 (4), momentum sums without the
    surface-tension term (4D; surface tension, 6 + 2D, runs only in the thin
    surface zone and is not counted), vPert_j.g (2D-1), its share of the
    viscous factor (2), the two density sums (6) and sum_j V_j g (2D)."""
    base = 4 * dim + 2 * dim + (2 * dim - 1) + 1 + 1 + 2 + dim + 2 + 2 * (2 * dim - 1) + 5
    if ale:
        return base + 4 + 4 * dim + (2 * dim - 1) + 2 + 6 + 2 * dim
    return base + 6 + 6 + 6 * dim + 2


METRIC = "particle-steps/s and %HBM roofline at 4M 3D particles; surrogate vs mesh aero"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_force_m_jet4m_traffic.json")


AUTOTUNE_CANDIDATES = (128, 192, 256, 384, 512, 768, 1024)


def workload(name, n_target):
    if name == "jet4m":
        c = cases.jet_column_3d(n_target=n_target, precision=abi.PREC_MIXED)
        desc = ("3D coflow jet column, D=2 mm liquid jet at 10 m/s in 30 m/s coflow, dx=0.04 mm, "
                "H=2dx, ALE shifting on, Gissler drag")
        return c, desc
    if name == "droplet":
        return cases.droplet_3d(precision=abi.PREC_MIXED), "3D droplet D=4 mm, dx=0.05 mm, 20 m/s crossflow"
    if name == "dam2d":
        return cases.dam_break_2d(), "2D dam break 1x2 m column, dx=0.02 m, FP64"
    if name == "jicf":
        c = cases.jet_in_crossflow_3d(n_target=1_000_000, precision=abi.PREC_MIXED)
        desc = (f"3D jet in crossflow: D=1 mm pipe with ghost walls (2 layers) injecting at 5 m/s along +y, "
                f"dx=0.03 mm, H=2dx, {c.n_columns} inlet columns, the jet issued to about 1M particles and "
                "growing, Gissler drag in a 50 m/s crossflow")
        return c, desc
    if name == "batched":
        c = cases.batched_droplets(precision=abi.PREC_MIXED)
        desc = ("256 independent 3D droplets, diameters log-uniform in [0.5, 2] mm, dx=0.05 mm, crossflow "
                "5-40 m/s, Gissler drag; each droplet its own simulation (own time step, CFL control, "
                "Newmark-Beta convergence and npd), one launch per pass over all droplets")
        return c, desc
    raise SystemExit(f"unknown workload {name}")


def load_traffic(n):
    """HBM bytes per k_force_m launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor),
    rescaled to n particles."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["bytes_per_particle"] * n, t
    except (OSError, ValueError, KeyError):
        return None, None


def cpu_baseline(args):
    """FP64 restatement (reference flags -O3 -ffast-math, OpenMP) on a
    bounded sample of the same workload: start-up pass + 2 untimed steps,
    then `cpu_steps` timed steps with the reference's serial
    dissipation_terms, then `cpu_steps` more with it parallelised."""
    from oracle.pyoracle import Oracle
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if args.workload == "jet4m":
        c = cases.jet_column_3d(n_target=args.cpu_particles, precision=abi.PREC_FP64)
    elif args.workload == "batched":
        c = cases.batched_droplets(n_drops=8, precision=abi.PREC_FP64)  # drops of the same size law
    elif args.workload == "jicf":
        # the same pipe, inlet and physics, issued to a shorter jet
        c = cases.jet_in_crossflow_3d(n_target=args.cpu_particles, precision=abi.PREC_FP64)
    else:
        c, _ = workload(args.workload, args.particles)
    o = Oracle(c, nthreads=nthreads, kind="baseline")
    o.integrate_no_update()
    # the GPU runs time steps warmup + 2 onwards (warm-up steps, then one
    # step after the tile autotuner): the same window here
    for _ in range(args.warmup + 1):
        o.integrate_step()
    res = {}
    for serial in (True, False):
        o.set_serial_dissipation(serial)
        its = []
        particle_steps = 0
        t0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            # a growing particle set (inlets) counts the particles each step starts with
            particle_steps += o.n
            its.append(o.integrate_step()["iterations"])
        res[serial] = (particle_steps / (time.perf_counter() - t0), float(np.mean(its)))
    o.close()
    return {
        "value": res[True][0],
        "unit": "particle-steps/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"{c.name} with {c.n} particles (same dx, H and physics, constant gas velocity), FP64, "
                   f"{args.cpu_steps} timed steps after start-up + {args.warmup + 1} steps, mean "
                   f"{res[True][1]:.1f} sub-iterations/step; cell-list neighbour search, serial "
                   "dissipation_terms as in the reference"),
        "subiterations_per_step": res[True][1],
        "value_parallel_dissipation": res[False][0],
        "subiterations_per_step_parallel_dissipation": res[False][1],
        "host": _host(),
    }


def _host():
    """The CPU the baseline ran on: model string, logical CPUs listed and
    usable (affinity mask), sockets/cores as /proc/cpuinfo gives them."""
    info = {"nproc": os.cpu_count(), "affinity": None, "cpu_model": None, "logical_cpus_listed": 0,
            "physical_ids": 0, "cores_per_socket": None}
    try:
        aff = sorted(os.sched_getaffinity(0))
        info["affinity"] = len(aff)
        info["affinity_cpus"] = f"{aff[0]}-{aff[-1]}" if aff and aff[-1] - aff[0] + 1 == len(aff) else str(aff)
    except (AttributeError, OSError):
        pass
    try:
        phys = set()
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "processor":
                    info["logical_cpus_listed"] += 1
                elif key == "model name" and info["cpu_model"] is None:
                    info["cpu_model"] = val
                elif key == "physical id":
                    phys.add(val)
                elif key == "cpu cores" and info["cores_per_socket"] is None:
                    info["cores_per_socket"] = int(val)
        info["physical_ids"] = len(phys)
    except OSError:
        pass
    return info


def copy_probe(nbytes):
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}))


def run_engine(case, args, label, autotune=True, block_targets=None, steps=None):
    """Warm up, then time `steps` steps; returns the measurements."""
    from fjsph_amd.engine import Engine

    steps = steps or args.steps
    print(f"[bench] {label}: {case.n} particles; building engine", file=sys.stderr, flush=True)
    eng = Engine(case)
    grouped = getattr(case, "owner", None) is not None
    if grouped:
        # independent simulations: each droplet keeps its own time step, CFL
        # control, Newmark-Beta convergence and npd
        eng.set_groups(case.owner)
    eng.set_reorder_interval(args.reorder_every)
    if block_targets or args.block_targets:
        eng.set_block_targets(block_targets or args.block_targets)
    if getattr(case, "surrogate", None) is not None:
        eng.set_surrogate(case.surrogate)
    eng.integrate_no_update()
    for _ in range(args.warmup):
        eng.integrate_step()
    tune = None
    if autotune and not args.no_autotune:
        # tile blocks chosen by the engine's autotuner on this state (untimed)
        best, ms = eng.autotune_blocks(AUTOTUNE_CANDIDATES, reps=2)
        tune = {"block_targets": best, "ms_per_candidate": {str(k): v for k, v in ms.items()}}
        print(f"[bench] {label}: autotuned tile blocks: {best} targets per block", file=sys.stderr, flush=True)
        eng.integrate_step()
    eng.enable_timers(True)
    its, group_its = [], []
    growing = bool(getattr(case, "inlets", None))
    particle_steps = 0
    w0 = time.monotonic_ns()
    t0 = time.perf_counter()
    for _ in range(steps):
        # a growing particle set (inlets) counts the particles each step starts with
        particle_steps += eng.n if growing else case.n
        its.append(eng.integrate_step()["iterations"])
        if grouped:
            group_its.append([g["iterations"] for g in eng.group_info()])
    elapsed = time.perf_counter() - t0
    w1 = time.monotonic_ns()
    r = {"elapsed": elapsed, "value": particle_steps / elapsed, "its": float(np.mean(its)), "autotune": tune,
         "window_ns": [w0, w1],
         "particles_end": eng.n,
         "steps": steps, "block_targets": eng.get_block_targets(),
         "timers": eng.timers(), "stats": eng.neighbour_stats(),
         "force_bpp": eng.pass_bytes(abi.TIMER_NAMES.index("force")),
         "pass_bytes": {k: eng.pass_bytes(abi.TIMER_NAMES.index(k)) for k in abi.TIMER_NAMES}}
    if grouped:
        gi = np.asarray(group_its, dtype=np.float64)
        sizes = np.bincount(case.owner).astype(np.float64)
        r["groups"] = {"drops": int(sizes.size), "mean_subiterations_per_drop_step": float(gi.mean()),
                       "particle_weighted_subiterations": float((gi * sizes).sum() / (sizes.sum() * gi.shape[0])),
                       "max_subiterations": float(gi.max()),
                       "time_steps": "per drop (own find_timestep, CFL control and Newmark-Beta convergence)"}
    if case.settings.aero_source in (abi.MESH_INFL, abi.SURROGATE):
        cid = eng.field("CELLID")
        aero = cid >= 0
        r["aero_particles"] = int(aero.sum())
        if case.settings.aero_source == abi.MESH_INFL:
            r["mesh_lost"] = eng.mesh_lost()
            r["aero_x"] = eng.field("X")[aero]
            r["aero_vals"] = np.concatenate([eng.field("CELLV")[aero], eng.field("CELLP")[aero, None],
                                             eng.field("CELLRHO")[aero, None]], axis=1)
    eng.close()
    print(f"[bench] {label}: {r['value'] / 1e6:.2f}M particle-steps/s", file=sys.stderr, flush=True)
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="jet4m")
    ap.add_argument("--particles", type=int, default=4_000_000)
    ap.add_argument("--aero", default="surrogate", choices=["surrogate", "mesh", "const"],
                    help="gas velocity source: MLP surrogate of the mesh (timed, with the mesh lookup timed "
                         "beside it), the mesh lookup alone, or the constant coflow velocity")
    ap.add_argument("--mesh-cells", type=int, default=20_000_000,
                    help="cells of the structured tetrahedral gas mesh (the coflow configuration's 20M)")
    ap.add_argument("--mesh-kind", default="kuhn", choices=["kuhn", "delaunay"],
                    help="structured tetrahedral mesh, or a seeded Delaunay mesh of --mesh-points points")
    ap.add_argument("--mesh-points", type=int, default=300_000)
    ap.add_argument("--train-samples", type=int, default=1_000_000)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--cpu-particles", type=int, default=60_000)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the constant-gas and FP64 runs and the error check (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the tile blocks as set (default 256 targets)")
    ap.add_argument("--block-targets", type=int, default=0,
                    help="targets per tile block before any autotuning (0: the engine's default, 256)")
    ap.add_argument("--reorder-every", type=int, default=4,
                    help="steps between the cell-order permutations of the particle arrays")
    ap.add_argument("--copy-probe", type=int, default=0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the path shards by "
              "spatial domain decomposition, which is described but not built.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    phases = {}
    t0 = time.perf_counter()
    case, desc = workload(args.workload, args.particles)
    phases["case_build_s"] = time.perf_counter() - t0
    if args.workload == "jicf" and args.aero != "const":
        # the configuration's mesh: a seeded Delaunay triangulation of 2M points in a 0.2 m box
        args.mesh_kind = "delaunay"
        args.mesh_points = args.mesh_points if args.mesh_points != 300_000 else 2_000_000
    aero_info = None
    if args.workload == "batched":
        args.aero = "const"  # each drop's crossflow is its own initial velocity in still gas
    const_case = None
    if args.aero == "const":
        desc += ", gas velocity: the constant coflow velocity"
        t0 = time.perf_counter()
        run = run_engine(case, args, "constant-velocity aero")
        phases["engine_const_s"] = time.perf_counter() - t0
        const_run = run
    else:
        from fjsph_amd import surrogate
        const_case = _copy_case(case)
        t0 = time.perf_counter()
        mesh_kw = dict(n_points=args.mesh_points, speed=float(case.settings.v_inf[0]), kind=args.mesh_kind,
                       n_cells=args.mesh_cells)
        if args.workload == "jicf":
            mesh_kw["box"] = ((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1))
        cases.with_mesh(case, **mesh_kw)
        t_mesh = time.perf_counter() - t0
        phases["mesh_build_s"] = t_mesh
        print(f"[bench] mesh: {case.mesh.ncells} cells ({t_mesh:.0f} s)", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        mesh_run = run_engine(case, args, "mesh aero")
        phases["engine_mesh_s"] = time.perf_counter() - t0

        def lookup_ms(r):
            ms, calls = r["timers"]["aero"]
            return ms / calls if calls else None

        aero_info = {"mesh_cells": int(case.mesh.ncells), "mesh_points": int(case.mesh.nverts),
                     "mesh_lookup_ms_per_call": lookup_ms(mesh_run),
                     "mesh_aero_value": mesh_run["value"], "mesh_aero_ms_per_step": mesh_run["elapsed"] / args.steps * 1e3,
                     "mesh_subiterations_per_step": mesh_run["its"],
                     "mesh_lost": mesh_run.get("mesh_lost"), "aero_particles": mesh_run.get("aero_particles")}
        run = mesh_run
        if args.aero == "surrogate":
            t0 = time.perf_counter()
            from fjsph_amd.engine import Engine
            lab = Engine(case.settings)
            lab.set_mesh(case.mesh)
            sg = surrogate.train(case.mesh, lab, n_samples=args.train_samples, steps=args.train_steps)
            lab.close()
            phases["surrogate_train_s"] = time.perf_counter() - t0
            aero_info["train_seconds"] = phases["surrogate_train_s"]
            case.settings.aero_source = abi.SURROGATE
            case.surrogate = sg
            t0 = time.perf_counter()
            run = run_engine(case, args, "surrogate aero")
            phases["engine_surrogate_s"] = time.perf_counter() - t0
            chk = Engine(case.settings)
            chk.set_surrogate(sg)
            pred = chk.surrogate_eval(mesh_run["aero_x"])
            chk.close()
            err = pred - mesh_run["aero_vals"]
            dim = case.settings.dim
            aero_info.update({
                "surrogate_value": run["value"],
                "surrogate_lookup_ms_per_call": lookup_ms(run),
                "surrogate_vs_mesh_speedup": run["value"] / mesh_run["value"],
                "surrogate_rms_error": {"vel_m_s": float(np.sqrt(np.mean(np.sum(err[:, :dim] ** 2, axis=1)))),
                                        "p_pa": float(np.sqrt(np.mean(err[:, dim] ** 2))),
                                        "rho": float(np.sqrt(np.mean(err[:, dim + 1] ** 2)))},
                "surrogate": f"MLP 3-64x{sg.n_hidden}-5, GELU, BF16 MFMA with FP32 accumulation, trained on "
                             f"{args.train_samples} mesh-lookup samples",
            })
        desc += (f", gas velocity from {'the MLP surrogate of ' if args.aero == 'surrogate' else ''}a "
                 f"{case.mesh.ncells}-cell {'structured tetrahedral' if args.mesh_kind == 'kuhn' else 'Delaunay'} "
                 f"mesh (potential flow past a cylinder, "
                 f"{case.settings.v_inf[0]:.0f} m/s)")
        const_run = None
        if not args.no_extra:
            # the same state with the constant gas velocity the CPU sample uses
            t0 = time.perf_counter()
            const_run = run_engine(const_case, args, "constant-velocity aero (CPU-comparable)", autotune=False,
                                   block_targets=run["block_targets"])
            phases["engine_const_s"] = time.perf_counter() - t0

    timers = run["timers"]
    pairs, max_count = run["stats"]
    passes = {}
    for name, (ms, launches) in timers.items():
        if launches == 0:
            continue
        avg_s = ms / 1e3 / launches
        bpp = run["pass_bytes"][name]
        passes[name] = {
            "ms_per_step": round(ms / args.steps, 3),
            "ms_per_launch": round(avg_s * 1e3, 3),
            "launches_per_step": round(launches / args.steps, 2),
            "alg_bytes_per_particle": bpp,
            "alg_GBs": round(bpp * case.n / avg_s / 1e9, 1),
        }

    # roofline of the dominant kernel: k_force_m (HIP events on the engine stream).  The pair loop is
    # bound by the FP32 vector pipe, not by HBM: 176 B per particle against ~240 pairs of ~100
    # operations.  Both sides are reported; `bound` names the larger fraction.
    ms, launches = timers["force"]
    avg_s = (ms / 1e3) / max(launches, 1)
    bpp = run["force_bpp"]
    achieved_gbs = bpp * case.n / avg_s / 1e9
    traffic, tmeta = load_traffic(case.n)
    fpp = force_flops_per_pair(case.settings.dim, bool(case.settings.ale))
    tflops = fpp * pairs / avg_s / 1e12
    hbm = {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
           "alg_bytes_per_particle": bpp, "traffic": traffic,
           "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if tmeta else None,
           "traffic_rate_GBs": traffic / avg_s / 1e9 if traffic else None}
    valu = {"achieved": tflops, "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / FP32_VECTOR_PEAK_TFLOPS,
            # the published FP32 peak counts packed (2-wide) FP32; the engine is built without packed
            # FP32 instructions, whose single-issue peak is half of it
            "peak_unpacked": FP32_VECTOR_PEAK_TFLOPS / 2, "frac_unpacked": tflops / (FP32_VECTOR_PEAK_TFLOPS / 2),
            "flop_per_pair": fpp, "pairs_per_launch": pairs}
    bound = "valu" if valu["frac"] >= hbm["frac"] else "hbm"
    side = valu if bound == "valu" else hbm
    roofline = {
        "bound": bound,
        "achieved": side["achieved"],
        "peak": side["peak"],
        "unit": side["unit"],
        "frac": side["frac"],
        "traffic": traffic,
        "kernel": "k_force_m<3,FR,ALE> (pair sums of get_acc_and_Rrho_on_i, mixed precision)",
        "avg_launch_ms": avg_s * 1e3,
        "launches": launches,
        "pairs_per_launch": pairs + case.n,
        "hbm": hbm,
        "valu": valu,
    }

    # FP64 GPU run and the error of the mixed step against it (the FP64 engine is the oracle-pinned path)
    fp64 = None
    if not args.no_extra:
        t0 = time.perf_counter()
        fp64 = fp64_check(const_case or case, args, run["block_targets"])
        phases["fp64_and_error_s"] = time.perf_counter() - t0

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        cpu = cpu_baseline(args)
        phases["cpu_s"] = time.perf_counter() - t0
        ref = const_run or run
        cpu["like_for_like"] = {
            "gpu_value_constant_gas": ref["value"],
            "gpu_subiterations_per_step": ref["its"],
            "cpu_subiterations_per_step": cpu["subiterations_per_step"],
            "ratio": ref["value"] / cpu["value"],
            "ratio_per_subiteration": (ref["value"] * ref["its"]) / (cpu["value"] * cpu["subiterations_per_step"]),
            "ratio_fp64": (fp64["value"] / cpu["value"]) if fp64 else None,
            "note": "both with the constant gas velocity; the GPU state is the 4M column, the CPU sample a "
                    "short column of the same physics; ratio_per_subiteration compares force passes per second",
        }

    value = run["value"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": run["elapsed"] / args.steps * 1e3,
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
            "subiterations_per_step": run["its"],
            "groups": run.get("groups"),
            "aero": aero_info,
            "tile_blocks": run.get("autotune") or {"block_targets": run["block_targets"], "autotuned": False},
            "reorder_every": args.reorder_every,
            "passes": passes,
            "constant_gas_value": const_run["value"] if const_run else None,
            "phases_s": {k: round(v, 2) for k, v in phases.items()},
        },
        "fp64_value": fp64["value"] if fp64 else None,
        "oracle_error": fp64["oracle_error"] if fp64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        # CLOCK_MONOTONIC bounds of the timed steps of the reported run: a kernel
        # trace of the same command gives the GPU-busy fraction inside them
        "timed_window_ns": run["window_ns"],
    }
    if cpu:
        out["config"]["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


def _copy_case(case):
    import copy
    c = copy.copy(case)
    c.settings = case.settings.copy()
    return c


def fp64_check(case, args, block_targets):
    """The state in FP64 on the GPU: its throughput (the like-for-like
    precision of the FP64 CPU restatement), and the error of the mixed-
    precision step against the FP64 one from the same state, normwise
    (max |mixed - fp64| / max |fp64|) per field after the start-up pass and
    one step.  The FP64 engine is the path pinned against the oracle.
    Particles are matched by id (inlets insert particles during the step);
    positions are compared against the step's displacement."""
    from fjsph_amd.engine import Engine
    grouped = getattr(case, "owner", None) is not None
    c64 = _copy_case(case)
    c64.settings.precision = abi.PREC_FP64
    cm = _copy_case(case)
    cm.settings.precision = abi.PREC_MIXED
    res, infos = {}, {}
    for label, c in (("fp64", c64), ("mixed", cm)):
        eng = Engine(c)
        if grouped:
            eng.set_groups(c.owner)
        eng.set_block_targets(block_targets)
        # positions the step starts from (the start-up pass leaves a predicted
        # state in place, which the step recomputes from these)
        x_before = dict(zip(eng.field("PID").tolist(), eng.field("X")))
        eng.integrate_no_update()
        infos[label] = eng.integrate_step()
        r = {f: eng.field(f) for f in ("X", "V", "RHO", "ACC", "PID")}
        r["X0"] = np.array([x_before.get(int(q), np.full(c.dim, np.nan)) for q in r["PID"]])
        res[label] = r
        eng.close()
    pid64, pidm = res["fp64"]["PID"], res["mixed"]["PID"]
    common, i64, im = np.intersect1d(pid64, pidm, return_indices=True)
    err = {}
    for f in ("X", "V", "RHO", "ACC"):
        a, b = res["mixed"][f][im], res["fp64"][f][i64]
        if f == "X":
            disp = np.abs(b - res["fp64"]["X0"][i64])
            err[f] = float(np.max(np.abs(a - b)) / max(np.nanmax(disp), 1e-300))
        else:
            err[f] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    run = run_engine(c64, args, "FP64", autotune=False, block_targets=block_targets, steps=max(2, args.steps // 2))
    return {"value": run["value"], "its": run["its"],
            "oracle_error": {"fields": err, "per": "one step from the same state, normwise over the particles both "
                             "runs hold; X against the step's displacement",
                             "particles": {"fp64": int(pid64.size), "mixed": int(pidm.size), "common": int(common.size)},
                             "subiterations": {"fp64": infos["fp64"]["iterations"],
                                               "mixed": infos["mixed"]["iterations"]}}}


if __name__ == "__main__":
    main()
