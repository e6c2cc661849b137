"""Write the FP64 oracle trajectories committed next to this script.

The reference (WarpVBD) ships no test vectors and cannot run here (NVIDIA Warp
is not installed), so these fixtures are produced by the project's FP64
restatement (oracle/vbd_oracle.c) of src/vbd_solver.py, after that restatement
has been pinned by the closed-form, finite-difference and brute-force checks in
tests/test_oracle.py.  They freeze the oracle's behaviour (a regression
anchor) and are the trajectories the GPU solver is compared with.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from warpvbd_amd.scenes import cantilever, mass_spring  # noqa: E402


def trajectory(scene, steps, dt, tol, max_iter, material="snh", damping=0.0, expanded=True):
    kw = scene.solver_kwargs()
    kw["damping_coefficient"] = damping
    s = oracle.OracleSolver(**kw, dx_tol=tol, max_iter=max_iter, material=material, expanded=expanded)
    iters, tipz, metric = [], [], []
    for _ in range(steps):
        x = s.step(dt)
        iters.append(s.last_iters)
        metric.append(s.last_metric)
        tipz.append(x[scene.tip_idx, 2].mean())
    return dict(iters=np.array(iters, np.int32), tipz=np.array(tipz), metric=np.array(metric), x_final=x,
                colors=s.colors())


def main():
    cases = {
        # the 10x2x2 cantilever workload: SNH, E=1e5, nu=0.4, rho=1000, dt=1e-3, 300 steps, tol 1e-7
        "cantilever10x2x2_snh_tol1e-7": (cantilever((10, 2, 2)), 300, 1e-3, 1e-7, 1000, "snh", 0.0),
        # same beam, StVK and the reference's default damping, fixed 10 iterations per step
        "cantilever10x2x2_stvk_fixed10": (cantilever((10, 2, 2)), 50, 1e-3, 0.0, 10, "stvk", 0.0),
        "cantilever10x2x2_snh_damp1_fixed10": (cantilever((10, 2, 2)), 50, 1e-3, 0.0, 10, "snh", 1.0),
        # a small mass-spring body (cube 6, column 3), tol 1e-7
        "massspring6+3_snh_tol1e-7": (mass_spring(6, 3), 100, 1e-3, 1e-7, 1000, "snh", 0.0),
    }
    for name, (sc, steps, dt, tol, mi, mat, damp) in cases.items():
        d = trajectory(sc, steps, dt, tol, mi, mat, damp)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, dt=dt, tol=tol, max_iter=mi, steps=steps, damping=damp, material=mat, **d)
        print(f"{name}: iterations {int(d['iters'].sum())}, final tip z {d['tipz'][-1]:.9f} -> {path}")


if __name__ == "__main__":
    main()
