"""Benchmark of the VBD time step on one MI355X.

Default workload: the 256x32x32-voxel cantilever (279,873 vertices, 1,310,720
tets, x = 0 face pinned), stable Neo-Hookean, E = 1e5 Pa, nu = 0.4,
rho = 1000 kg/m^3, gravity -z, dt = 1e-3 s, tolerance 1e-7 m, at most 1000
iterations per step, FP32.  A "step" is one implicit-Euler time step (all
colour-sweep iterations to the tolerance).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--material snh|stvk]
                  [--workload cantilever256|cantilever128|cantilever10] [--no-cpu]

value = simulated seconds per wall second over the K timed steps.  The
roofline entry is for the colour-sweep kernel: algorithmic bytes of the swept
colours divided by their summed kernel time, from HIP events recorded on the
solver's stream around every sweep launch of a colour with >= 4096 vertices
inside the timed region.  cpu_baseline times the FP64 CPU oracle (a C
restatement of the reference's algorithm) on a bounded sample of the same
workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sim-s/wall-s at tol 1e-7 (warm-start+batched vs plain VBD); HBM GB/s"
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "cantilever256": (256, 32, 32),
    "cantilever128": (128, 16, 16),
    "cantilever10": (10, 2, 2),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=300)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--workload", default="cantilever256", choices=sorted(WORKLOADS))
    p.add_argument("--material", default="snh", choices=["snh", "stvk"])
    p.add_argument("--tol", type=float, default=1e-7)
    p.add_argument("--max-iter", type=int, default=1000)
    p.add_argument("--dt", type=float, default=1e-3)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    p.add_argument("--lanes", type=int, default=0, help="lanes per vertex (0 = library default)")
    return p.parse_args()


def cpu_baseline(scene, args, gpu_iters_per_step):
    """Time the FP64 oracle (reference formula, 1 thread) on a bounded sample:
    one time step of the same workload, capped at a number of iterations that
    fits the time budget; scaled to sim-s/wall-s with the GPU run's mean
    iterations per step (both solve the same system to the same tolerance)."""
    import oracle

    oracle.build()
    kw = scene.solver_kwargs()
    t0 = time.perf_counter()
    s = oracle.OracleSolver(**kw, dx_tol=0.0, max_iter=1, material=args.material, expanded=True, threads=1)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.step(args.dt, dx_tol=0.0, max_iter=1)
    t_one = time.perf_counter() - t0
    n_it = max(1, min(50, int(args.cpu_seconds / max(t_one, 1e-9))))
    s.reset()
    t0 = time.perf_counter()
    s.step(args.dt, dx_tol=0.0, max_iter=n_it)
    t_it = (time.perf_counter() - t0) / n_it
    sim_per_wall = args.dt / (t_it * gpu_iters_per_step)
    return {
        "value": sim_per_wall,
        "unit": "sim-s/wall-s",
        "cores": 1,
        "kind": "port",
        "sample": (f"FP64 oracle (reference formula, oracle/vbd_oracle.c, 1 thread) on the same mesh: "
                   f"{n_it} colour-sweep iterations of one time step, {t_it * 1e3:.1f} ms/iteration, "
                   f"scaled by the GPU's {gpu_iters_per_step:.2f} iterations/step; setup {setup:.1f} s excluded"),
        "ms_per_iteration": t_it * 1e3,
        "host_cpus": len(os.sched_getaffinity(0)),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)
    from warpvbd_amd import VBDSolver
    from warpvbd_amd.scenes import cantilever

    nx = WORKLOADS[args.workload]
    sc = cantilever(nx)
    solver = VBDSolver(**sc.solver_kwargs(), dx_tol=args.tol, max_iter=args.max_iter, material=args.material,
                       verbose=False)
    if args.lanes:
        solver.set_launch(args.lanes, 256, 2048, 0)
    for _ in range(args.warmup):
        solver.step(solver.old_positions, args.dt)
    solver.profile(True, 4096)
    solver.profile_read()
    iters = []
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        solver.step(solver.old_positions, args.dt)
        step_ms.append((time.perf_counter() - a) * 1e3)
        iters.append(solver.last_iterations)
    wall = time.perf_counter() - t0
    prof = solver.profile_read()
    n_active = int(sc.active_mask.sum())
    achieved = prof["bytes"] / (prof["ms"] * 1e-3) / 1e9 if prof["ms"] > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"sweep_traffic_{args.workload}_{args.material}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    sim_per_wall = args.steps * args.dt / wall
    out = {
        "metric": METRIC,
        "value": sim_per_wall,
        "unit": "sim-s/wall-s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (voxel cantilever built in-process; no dataset)",
        "config": {
            "workload": f"plain VBD cantilever {nx[0]}x{nx[1]}x{nx[2]} voxels of 1 cm "
                        f"({sc.n_vertices} vertices, {sc.n_elements} tets), {args.material.upper()}",
            "E": 1e5, "nu": 0.4, "rho": 1000.0, "dt": args.dt, "tol": args.tol, "max_iter": args.max_iter,
            "parallelism": "single GPU",
        },
        "iterations_per_step": float(np.mean(iters)),
        "vertex_updates_per_s": float(np.sum(iters)) * n_active / wall,
        "step_ms_p50_p90_p99": [float(np.percentile(step_ms, q)) for q in (50, 90, 99)],
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_sweep (one colour, >= 4096 vertices)",
            "launches": prof["launches"],
            "avg_launch_us": prof["ms"] * 1e3 / max(1, prof["launches"]),
            "algorithmic_bytes_per_launch": prof["bytes"] / max(1, prof["launches"]),
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args, float(np.mean(iters)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
