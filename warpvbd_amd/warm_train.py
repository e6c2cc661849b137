"""Training and held-out evaluation of the learned warm start on one mesh.

    python -m warpvbd_amd.warm_train --mesh 128 16 16 --trajectories 200 --steps 300 \\
        --out warpvbd_amd/weights/warm_start_cantilever128x16x16.npy --report report.json

Trajectory t (t = 0 .. trajectories-1) is drawn from seed t: E log-uniform in
[1e4, 1e6] Pa, nu uniform in [0.3, 0.45], gravity of 9.81 m/s^2 in a random
direction within 30 degrees of -z, initial velocity N(0, 0.05 m/s) per
component (WarmStart.generate).  Every `sample_every`-th step of a trajectory
is stored as a training sample on the device.  Training is Adam over the
stored samples, in stages of decreasing learning rate.

The held-out evaluation draws its scenes from seeds the training never saw
(1000 .. 1019 by default) and runs, from the same start, a plain solver and a
solver per warm-start policy for `eval_steps` steps each, timing every step
(the cyclic garbage collector off) and recording the iteration counts, the
guess each step started from and the largest position difference from the
plain trajectory.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

from .scenes import cantilever
from .vbd_solver import VBDSolver
from .warm_start import WarmStart, gravity_in_cone

POLICIES = {"history": WarmStart.HISTORY, "choose": WarmStart.CHOOSE, "adaptive": WarmStart.ADAPTIVE,
            "network": WarmStart.NETWORK}


def scene_of_seed(n_vertices, active_mask, seed):
    """The trajectory parameters WarmStart.generate draws for `seed` (one
    trajectory): E, nu, gravity, initial velocity."""
    rng = np.random.default_rng(seed)
    E = float(10 ** rng.uniform(4.0, 6.0))
    nu = float(rng.uniform(0.3, 0.45))
    g = gravity_in_cone(rng, 30.0)
    v0 = rng.normal(0.0, 0.05, size=(n_vertices, 3)) * active_mask[:, None]
    return E, nu, g, v0


def train(sc, trajectories, steps, sample_every=2, stages=((8, 2e-3), (6, 7e-4), (4, 2e-4)), batch=8, dt=1e-3,
          tol=1e-7, max_iter=1000, log=print):
    """Generate seeded trajectories 0 .. trajectories-1 and train; returns
    (parameters, report)."""
    s = VBDSolver(**sc.solver_kwargs(), dx_tol=tol, max_iter=max_iter, verbose=False)
    per_traj = (steps + sample_every - 1) // sample_every
    cap = trajectories * per_traj
    ws = WarmStart(s, capacity=cap, seed=0)
    t0 = time.perf_counter()
    iters = []
    for t in range(trajectories):
        iters.append(ws.generate(trajectories=1, steps=steps, dt=dt, tol=tol, max_iter=max_iter, seed=t,
                                 sample_every=sample_every))
        if (t + 1) % 20 == 0 or t + 1 == trajectories:
            log(f"  trajectories {t + 1}/{trajectories}: {ws.n_samples} samples, {time.perf_counter() - t0:.1f} s")
    t_gen = time.perf_counter() - t0
    losses = []
    t0 = time.perf_counter()
    for k, (epochs, lr) in enumerate(stages):
        for e in range(epochs):  # one call per epoch (Adam's state carries over), for a progress line each
            h = ws.train(epochs=1, batch=batch, lr=lr, seed=1000 * k + e)
            losses.append(float(h[0]))
            log(f"  stage {k} epoch {e}: lr {lr:g}, loss {h[0]:.4g} ({time.perf_counter() - t0:.1f} s)")
    t_train = time.perf_counter() - t0
    params = ws.get_params()
    report = {"trajectories": trajectories, "seeds": [0, trajectories - 1], "steps": steps,
              "sample_every": sample_every, "samples": ws.n_samples, "n_vertices": sc.n_vertices,
              "stages": [list(x) for x in stages], "batch": batch, "loss_per_epoch": losses,
              "plain_iterations_per_step": float(np.mean(np.concatenate(iters))),
              "generate_s": t_gen, "train_s": t_train}
    ws.close()
    s.close()
    return params, report


def evaluate(sc, params, seeds, steps, policies=("history", "choose", "adaptive", "network"), dt=1e-3, tol=1e-7,
             max_iter=1000, log=print):
    """Held-out scenes: plain and warm-started solvers from the same start.
    Returns per policy and for plain: iterations per step, ms per step, and
    per warm policy the guess counts and max |x - x_plain|."""
    plain = VBDSolver(**sc.solver_kwargs(), dx_tol=tol, max_iter=max_iter, verbose=False)
    warm = {}
    for name in policies:
        sol = VBDSolver(**sc.solver_kwargs(), dx_tol=tol, max_iter=max_iter, verbose=False)
        ws = WarmStart(sol, capacity=1)
        ws.set_params(params)
        ws.set_policy(POLICIES[name])
        warm[name] = (sol, ws)
    rec = {k: {"iters": [], "ms": [], "guess": {"plain": 0, "network": 0, "history": 0}, "max_diff_m": 0.0}
           for k in ("plain",) + tuple(policies)}
    per_seed = []
    for seed in seeds:
        E, nu, g, v0 = scene_of_seed(sc.n_vertices, sc.active_mask, seed)
        sols = [("plain", plain, None)] + [(k, sol, ws) for k, (sol, ws) in warm.items()]
        for _, sol, ws in sols:
            sol.reset()
            sol.set_material(E, nu)
            sol.set_gravity(g)
            sol.old_velocities = v0
            if ws is not None:
                ws.set_material(E, nu)
        gc.collect()
        gc.disable()
        try:
            for _ in range(steps):
                for name, sol, ws in sols:
                    a = time.perf_counter()
                    sol.step(sol.old_positions, dt, warm_start=ws)
                    rec[name]["ms"].append((time.perf_counter() - a) * 1e3)
                    rec[name]["iters"].append(sol.last_iterations)
                    if ws is not None:
                        for k, v in ws.last_guess_counts().items():
                            rec[name]["guess"][k] += v
        finally:
            gc.enable()
        xp = plain.old_positions.numpy()
        for name, (sol, _) in warm.items():
            rec[name]["max_diff_m"] = max(rec[name]["max_diff_m"], float(np.abs(sol.old_positions.numpy() - xp).max()))
        per_seed.append({"seed": int(seed), "E": E, "nu": nu, "gravity": [float(x) for x in g],
                         **{k: {"iterations_per_step": float(np.mean(v["iters"][-steps:])),
                                "ms_per_step": float(np.mean(v["ms"][-steps:]))} for k, v in rec.items()}})
        log(f"  seed {seed}: E {E:.3g} nu {nu:.3f}: iterations/step " +
            ", ".join(f"{k} {np.mean(v['iters'][-steps:]):.2f}" for k, v in rec.items()))
    out = {}
    for name, r in rec.items():
        ms = np.array(r["ms"])
        out[name] = {"iterations_per_step": float(np.mean(r["iters"])), "ms_per_step": float(ms.mean()),
                     "sim_s_per_wall_s": dt / (ms.mean() * 1e-3),
                     "step_ms_p50_p90_p99": [float(np.percentile(ms, q)) for q in (50, 90, 99)]}
        if name != "plain":
            out[name]["guess_fraction"] = {k: v / (len(seeds) * steps) for k, v in r["guess"].items()}
            out[name]["max_abs_diff_vs_plain_m"] = r["max_diff_m"]
            out[name]["vs_plain"] = out["plain"]["ms_per_step"] / out[name]["ms_per_step"]
    if "history" in out:
        for name in policies:
            if name != "history":
                out[name]["vs_history"] = out["history"]["ms_per_step"] / out[name]["ms_per_step"]
                out[name]["iterations_vs_history"] = (out[name]["iterations_per_step"]
                                                      / out["history"]["iterations_per_step"])
    for _, (sol, ws) in warm.items():
        ws.close()
        sol.close()
    plain.close()
    return {"seeds": [int(seeds[0]), int(seeds[-1])], "steps_per_seed": steps, "policies": out, "per_seed": per_seed}


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--mesh", type=int, nargs=3, default=(128, 16, 16))
    p.add_argument("--trajectories", type=int, default=200)
    p.add_argument("--steps", type=int, default=300)
    p.add_argument("--sample-every", type=int, default=2)
    p.add_argument("--stages", default="8:2e-3,6:7e-4,4:2e-4", help="epochs:lr,... of the training stages")
    p.add_argument("--batch", type=int, default=8)
    p.add_argument("--eval-seeds", default="1000-1019")
    p.add_argument("--eval-steps", type=int, default=300)
    p.add_argument("--params", help="evaluate these parameters (.npy) instead of training")
    p.add_argument("--out", help="write the trained parameters (.npy)")
    p.add_argument("--report", help="write the report (.json)")
    a = p.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    sc = cantilever(tuple(a.mesh))
    report = {"mesh": list(a.mesh), "n_vertices": sc.n_vertices}
    if a.params:
        params = np.load(a.params)
        report["params"] = os.path.basename(a.params)
    else:
        stages = tuple((int(e), float(lr)) for e, lr in (x.split(":") for x in a.stages.split(",")))
        log(f"training on {a.trajectories} trajectories x {a.steps} steps of the {a.mesh} cantilever")
        params, report["training"] = train(sc, a.trajectories, a.steps, a.sample_every, stages, a.batch, log=log)
        if a.out:
            os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
            np.save(a.out, params)
    s0, s1 = (int(x) for x in a.eval_seeds.split("-"))
    seeds = list(range(s0, s1 + 1))
    log(f"held-out seeds {s0}..{s1}, {a.eval_steps} steps each")
    report["heldout"] = evaluate(sc, params, seeds, a.eval_steps, log=log)
    for k, v in report["heldout"]["policies"].items():
        log(f"{k:9s} {v['iterations_per_step']:6.2f} it/step {v['ms_per_step']:7.3f} ms/step "
            f"{v['sim_s_per_wall_s']:6.3f} sim-s/wall-s" + (f"  vs history {v['vs_history']:.3f}"
                                                            if "vs_history" in v else ""))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
