"""Scene drivers, tip-trace outputs and timing sweeps of WarpVBD's examples.

The reference drives its solver from example scripts: the scene classes
(`CantileverSim`, examples/run_cantilever.py:28-88; `MassSpringSim`,
examples/run_mass_spring.py:27-90; `MassSpringParallelSim`,
examples/run_mass_spring_parallel.py:27-127), the run loops that write the tip
trace (a CSV of the tip's mean position per step, run_cantilever.py:150-177 and
run_mass_spring.py; an NPZ with one trace per environment,
run_mass_spring_parallel.py:193-221), and the timing sweeps that produced the
README figures (benchmark_tolerance.py:29-66, benchmark_vbd_solver.py:29-70,
benchmark_vbd_parallel.py:26-60, benchmark_graph_coloring.py:24-45, all timed
with `benchmark_functions`, src/_utils.py:6-32).

This module restates them on this package's solvers with the same arguments,
defaults and file formats.  Rendering (PBRT) and plotting are left out; each
sweep returns its metrics dict and can save it as NPZ.
`MassSpringParallelSim(batched=True)` runs the environments as separate
bodies of a `BatchedVBDSolver` (per-environment convergence) instead of one
mesh with a global convergence test.

Command line: ``python -m warpvbd_amd.harness <run_cantilever | run_mass_spring
| run_mass_spring_parallel | tolerance | mesh_size | parallel | coloring> [--out DIR] [--quick]``.
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from ._utils import benchmark_functions, hex2tets, voxel2hex


# ----------------------------------------------------------------- scenes
class CantileverSim:
    """Voxel beam along x, x = 0 face pinned, tip = the x = L face
    (examples/run_cantilever.py:28-88)."""

    def __init__(self, nx=(10, 3, 3), dx=(0.01, 0.01, 0.01), density=1070, youngs_modulus=250e3,
                 poissons_ratio=0.45, dx_tol=1e-9, max_iter=100, device="cuda", material="snh", verbose=True):
        from .vbd_solver import VBDSolver
        vertices, hexes = voxel2hex(np.ones(nx, dtype=bool), *dx)
        elements = hex2tets(hexes)
        self.initial_positions, self.elements = vertices, elements
        active = ~(vertices[:, 0] < 1e-3)
        self.tip_idx = list(np.nonzero(active & (vertices[:, 0] > nx[0] * dx[0] - 1e-3))[0])
        self.active_mask = active
        m = len(elements)
        self.solver = VBDSolver(vertices, elements, density * np.ones(m), youngs_modulus * np.ones(m),
                                poissons_ratio * np.ones(m), damping_coefficient=0.0, active_mask=active,
                                gravity=(0.0, 0.0, -9.81), dx_tol=dx_tol, max_iter=max_iter, device=device,
                                material=material, verbose=verbose)

    def step(self, dt):
        return self.solver.step(self.solver.old_positions, float(dt))

    def reset(self):
        self.solver.reset()


class MassSpringSim:
    """A 1 m cube of nx^3 voxels hanging from a one-voxel column of nx voxels
    whose top face is pinned; tip = the cube's bottom face
    (examples/run_mass_spring.py:27-90)."""

    def __init__(self, nx=11, density=1, youngs_modulus=9e3, poissons_ratio=0.45, dx_tol=1e-9, max_iter=1000,
                 device="cuda", material="snh", verbose=True):
        from .vbd_solver import VBDSolver
        h = 1.0 / nx
        vox = np.ones((nx, nx, 2 * nx), dtype=bool)
        vox[:, :, nx:] = False
        vox[nx // 2:nx // 2 + 1, nx // 2:nx // 2 + 1, nx:] = True
        vertices, hexes = voxel2hex(vox, h, h, h)
        vertices[:, :2] -= vertices.mean(axis=0)[:2]
        elements = hex2tets(hexes)
        self.initial_positions, self.elements = vertices, elements
        active = ~(vertices[:, 2] > 2 * nx * h - 1e-3)
        self.tip_idx = list(np.nonzero(active & (vertices[:, 2] < 1e-3))[0])
        self.active_mask = active
        m = len(elements)
        self.solver = VBDSolver(vertices, elements, density * np.ones(m), youngs_modulus * np.ones(m),
                                poissons_ratio * np.ones(m), damping_coefficient=0.0, active_mask=active,
                                gravity=(0.0, 0.0, -9.81), dx_tol=dx_tol, max_iter=max_iter, device=device,
                                material=material, verbose=verbose)

    def step(self, dt):
        return self.solver.step(self.solver.old_positions, float(dt))

    def reset(self):
        self.solver.reset()


def _mass_spring_env(nx, x0, y0):
    """One environment of the parallel scene: cube at voxels [x0, x0+nx) x
    [y0, y0+nx) x [0, nx), column at (x0+nx//2, y0+nx//2) up to z = 2 nx."""
    h = 1.0 / nx
    vox = np.zeros((nx, nx, 2 * nx), dtype=bool)
    vox[:, :, :nx] = True
    vox[nx // 2, nx // 2, nx:] = True
    vertices, hexes = voxel2hex(vox, h, h, h)
    vertices[:, 0] += x0 * h
    vertices[:, 1] += y0 * h
    return vertices, hex2tets(hexes)


def _mass_spring_grid(nsim, nx):
    """The parallel scene as the reference builds it
    (examples/run_mass_spring_parallel.py:35-58,66-95): one voxel grid with
    nx voxels of padding between the environments' cubes and columns,
    converted and split as a whole (so vertex numbering runs across the
    environments), each element assigned to the environment its centroid
    lies in.  Returns vertices, tets and each element's environment."""
    h = 1.0 / nx
    nsimy = int(np.ceil(np.sqrt(nsim)))
    nsimx = int(np.ceil(nsim / nsimy))
    pad = int(1.0 * nx)
    vox = np.zeros((nsimx * (nx + pad) - pad, nsimy * (nx + pad) - pad, 2 * nx), dtype=bool)
    for k in range(nsim):
        x0, y0 = (k // nsimy) * (nx + pad), (k % nsimy) * (nx + pad)
        vox[x0:x0 + nx, y0:y0 + nx, :nx] = True
        vox[x0 + nx // 2, y0 + nx // 2, nx:] = True
    vertices, hexes = voxel2hex(vox, h, h, h)
    tets = hex2tets(hexes)
    com = vertices[tets].mean(axis=1)
    cell = (nx + pad) * h
    env = (com[:, 0] // cell).astype(np.int64) * nsimy + (com[:, 1] // cell).astype(np.int64)
    assert env.max() < nsim, "element outside every environment"
    return vertices, tets, env


class MassSpringParallelSim:
    """nsim mass-spring environments on an xy grid with nx voxels of padding
    between them, Young's modulus E (1 + U(-0.5, 2.5)) per environment
    (examples/run_mass_spring_parallel.py:27-127).  tip_idx maps each
    environment to its probe vertices (bottom face, z < 1e-3).

    batched=False builds the reference's scene: one padded voxel grid, one
    mesh, one solver with a single convergence test (same vertex numbering,
    so the same colouring and sweep order as the reference).  batched=True
    gives each environment its own mesh and convergence test in a
    BatchedVBDSolver.  seed fixes the per-environment moduli (the reference
    draws them unseeded)."""

    def __init__(self, nsim=4, nx=7, density=1, youngs_modulus=2e3, poissons_ratio=0.45, dx_tol=1e-9,
                 max_iter=1000, device="cuda", batched=False, seed=0, material="snh", verbose=True):
        from .batched import BatchedVBDSolver, concat_bodies
        from .vbd_solver import VBDSolver
        self.nsim = nsim
        nsimy = int(np.ceil(np.sqrt(nsim)))
        pad = int(1.0 * nx)
        rng = np.random.default_rng(seed)
        random_ym = rng.uniform(-0.5, 2.5, size=(nsim,))
        h = 1.0 / nx
        self.batched = batched
        if batched:
            bodies = []
            for k in range(nsim):
                i, j = k // nsimy, k % nsimy
                V, T = _mass_spring_env(nx, i * (nx + pad), j * (nx + pad))
                m = len(T)
                active = ~(V[:, 2] > 2 * nx * h - 1e-3)
                bodies.append(dict(initial_positions=V, elements=T, densities=density * np.ones(m),
                                   youngs_modulus=youngs_modulus * (1.0 + random_ym[k]) * np.ones(m),
                                   poissons_ratio=poissons_ratio * np.ones(m), active_mask=active,
                                   gravity=(0.0, 0.0, -9.81)))
            cb = concat_bodies(bodies)
            self.initial_positions, self.elements = cb["initial_positions"], cb["elements"]
            self.youngs_moduli = cb["youngs_modulus"]
            voff = cb["body_vertex_offsets"]
            self.vertex_sim_env = np.repeat(np.arange(nsim), np.diff(voff))
            act = cb["active_mask"]
            self.solver = BatchedVBDSolver(bodies, damping_coefficient=0.0, dx_tol=dx_tol, max_iter=max_iter,
                                           device=device, material=material, verbose=verbose)
        else:
            V, T, env = _mass_spring_grid(nsim, nx)
            m = len(T)
            self.initial_positions, self.elements = V, T
            self.element_sim_env = env
            self.vertex_sim_env = -np.ones(len(V), dtype=np.int64)
            self.vertex_sim_env[T.reshape(-1)] = np.repeat(env, 4)
            self.youngs_moduli = youngs_modulus * (1.0 + random_ym[env])
            act = ~(V[:, 2] > 2 * nx * h - 1e-3)
            self.solver = VBDSolver(V, T, density * np.ones(m), self.youngs_moduli, poissons_ratio * np.ones(m),
                                    damping_coefficient=0.0, active_mask=act, gravity=(0.0, 0.0, -9.81),
                                    dx_tol=dx_tol, max_iter=max_iter, device=device, material=material,
                                    verbose=verbose)
        Z = self.initial_positions[:, 2]
        self.tip_idx = {k: list(np.nonzero(act & (Z < 1e-3) & (self.vertex_sim_env == k))[0]) for k in range(nsim)}

    def step(self, dt):
        return self.solver.step(self.solver.old_positions, float(dt))

    def reset(self):
        self.solver.reset()


# ----------------------------------------------------------------- run loops (tip traces)
def run_cantilever(out_dir="outputs/cantilever", nx=(10, 3, 3), dx=(0.01, 0.01, 0.01), n_seconds=1.5, fps=100,
                   n_substeps=10, dx_tol=1e-9, max_iter=100, verbose=False):
    """The cantilever run of examples/run_cantilever.py:129-177: E=150 kPa,
    nu=0.45, rho=1070; dt = 1/fps/n_substeps.  Writes
    out_dir/displacement_cantilever.csv (the tip's mean x, y, z per step, comma
    separated) and returns the [steps, 3] array."""
    sim = CantileverSim(nx=nx, dx=dx, density=1070, youngs_modulus=150e3, poissons_ratio=0.45, dx_tol=dx_tol,
                        max_iter=max_iter, verbose=verbose)
    dt = 1.0 / fps / n_substeps
    tips = []
    for _ in range(int(n_seconds * fps) * n_substeps):
        x = sim.step(dt).numpy()
        tips.append(x[sim.tip_idx].mean(axis=0))
    tips = np.array(tips)
    os.makedirs(out_dir, exist_ok=True)
    np.savetxt(os.path.join(out_dir, "displacement_cantilever.csv"), tips, delimiter=",")
    return tips


def run_mass_spring(out_dir="outputs/mass_spring", nx=11, n_seconds=5.0, fps=50, n_substeps=5, dx_tol=1e-5,
                    max_iter=1000, verbose=False):
    """The mass-spring run of examples/run_mass_spring.py:133-180 (tol 1e-5,
    dt = 1/fps/n_substeps): writes out_dir/displacement_mass_spring.csv (tip
    mean x, y, z per step)."""
    sim = MassSpringSim(nx=nx, dx_tol=dx_tol, max_iter=max_iter, verbose=verbose)
    dt = 1.0 / fps / n_substeps
    tips = []
    for _ in range(int(n_seconds * fps) * n_substeps):
        x = sim.step(dt).numpy()
        tips.append(x[sim.tip_idx].mean(axis=0))
    tips = np.array(tips)
    os.makedirs(out_dir, exist_ok=True)
    np.savetxt(os.path.join(out_dir, "displacement_mass_spring.csv"), tips, delimiter=",")
    return tips


def run_mass_spring_parallel(out_dir="outputs/mass_spring_parallel", nsim=6, nx=7, n_seconds=5.0, fps=50,
                             n_substeps=5, dx_tol=1e-6, max_iter=1000, batched=False, seed=0, verbose=False):
    """The parallel run of examples/run_mass_spring_parallel.py:176-221:
    writes out_dir/tip_positions_mass_spring_parallel.npz with one [steps, 3]
    array per environment, keyed "0", "1", ..."""
    sim = MassSpringParallelSim(nsim=nsim, nx=nx, dx_tol=dx_tol, max_iter=max_iter, batched=batched, seed=seed,
                                verbose=verbose)
    dt = 1.0 / fps / n_substeps
    tips = {str(i): [] for i in range(nsim)}
    for _ in range(int(n_seconds * fps) * n_substeps):
        x = sim.step(dt).numpy()
        for i in range(nsim):
            tips[str(i)].append(x[sim.tip_idx[i]].mean(axis=0))
    tips = {k: np.array(v) for k, v in tips.items()}
    os.makedirs(out_dir, exist_ok=True)
    np.savez_compressed(os.path.join(out_dir, "tip_positions_mass_spring_parallel.npz"), **tips)
    return tips


# ----------------------------------------------------------------- timing sweeps
def _solve_loop(trace):
    def vbd_solve(simulation, nt, dt):
        simulation.reset()
        if trace is not None:
            trace.clear()
        for _ in range(nt):
            x = simulation.step(dt)
            if trace is not None:
                trace.append(x.numpy()[simulation.tip_idx].mean(axis=0))
    return vbd_solve


def benchmark_tolerance(tolerances=(1e-6, 5e-7, 1e-7, 5e-8, 1e-8), num_samples=3, n_timesteps=1000, timestep=1e-3,
                        max_iter=2000, out=None):
    """Wall seconds per simulated second and the tip trace against dx_tol on a
    10x3x3 cantilever, E=250 kPa, nu=0.45, rho=1070
    (examples/benchmark_tolerance.py:29-66)."""
    metrics = {"tolerance": [], "times": {"vbd_solve": []}, "tip_positions": []}
    for tol in tolerances:
        sim = CantileverSim(density=1070, youngs_modulus=250e3, poissons_ratio=0.45, dx_tol=tol, max_iter=max_iter,
                            verbose=False)
        trace = []
        times = benchmark_functions([_solve_loop(trace)], sim, n_timesteps, timestep, num_samples=num_samples)
        metrics["tolerance"].append(tol)
        metrics["times"]["vbd_solve"].append(times["vbd_solve"])
        metrics["tip_positions"].append(np.array(trace))
    metrics["time_per_sim_second"] = [float(np.mean(t)) / (timestep * n_timesteps)
                                      for t in metrics["times"]["vbd_solve"]]
    if out:
        np.savez_compressed(out, tolerance=np.array(metrics["tolerance"]),
                            times=np.array(metrics["times"]["vbd_solve"]),
                            tip_positions=np.array(metrics["tip_positions"]),
                            time_per_sim_second=np.array(metrics["time_per_sim_second"]))
    return metrics


def benchmark_mesh_size(reduction_factors=tuple(range(1, 11)), num_samples=1, n_timesteps=300, timestep=1e-3,
                        dx_tol=1e-6, max_iter=3000, out=None):
    """Wall seconds per simulated second against mesh size: cantilevers of
    (10 rf, 3 rf, 3 rf) voxels of 1/rf cm, E=150 kPa, nu=0.45, rho=1070,
    tol 1e-6 (examples/benchmark_vbd_solver.py:29-70)."""
    metrics = {"num_vertices": [], "num_elements": [], "times": {"vbd_solve": []}, "tip_positions": []}
    for rf in reduction_factors:
        h = 0.01 / rf
        sim = CantileverSim(nx=(10 * rf, 3 * rf, 3 * rf), dx=(h, h, h), density=1070, youngs_modulus=150e3,
                            poissons_ratio=0.45, dx_tol=dx_tol, max_iter=max_iter, verbose=False)
        trace = []
        times = benchmark_functions([_solve_loop(trace)], sim, n_timesteps, timestep, num_samples=num_samples)
        metrics["num_vertices"].append(len(sim.initial_positions))
        metrics["num_elements"].append(len(sim.elements))
        metrics["times"]["vbd_solve"].append(times["vbd_solve"])
        metrics["tip_positions"].append(np.array(trace))
        print(f"  rf {rf}: {len(sim.initial_positions)} vertices, "
              f"{float(np.mean(times['vbd_solve'])) / (timestep * n_timesteps):.3f} wall s per sim s", flush=True)
    metrics["time_per_sim_second"] = [float(np.mean(t)) / (timestep * n_timesteps)
                                      for t in metrics["times"]["vbd_solve"]]
    if out:
        np.savez_compressed(out, num_vertices=np.array(metrics["num_vertices"]),
                            num_elements=np.array(metrics["num_elements"]),
                            times=np.array(metrics["times"]["vbd_solve"]),
                            time_per_sim_second=np.array(metrics["time_per_sim_second"]))
    return metrics


def benchmark_parallel(nsims=tuple(2 ** i for i in range(13)), num_samples=3, n_timesteps=125, timestep=4e-3,
                       nx=7, dx_tol=1e-5, batched=False, out=None):
    """Wall seconds per simulated second and the per-environment realtime
    factor against the number of parallel mass-spring environments (nx=7,
    tol 1e-5, dt 4 ms, 125 steps; examples/benchmark_vbd_parallel.py:26-60)."""
    metrics = {"n_vertices": [], "n_simulations": [], "times": {"vbd_solve": []}}
    for nsim in nsims:
        sim = MassSpringParallelSim(nsim=nsim, nx=nx, dx_tol=dx_tol, batched=batched, verbose=False)
        times = benchmark_functions([_solve_loop(None)], sim, n_timesteps, timestep, num_samples=num_samples)
        metrics["n_vertices"].append(len(sim.initial_positions))
        metrics["n_simulations"].append(nsim)
        metrics["times"]["vbd_solve"].append(times["vbd_solve"])
        print(f"  {nsim} environments ({len(sim.initial_positions)} vertices): "
              f"{float(np.mean(times['vbd_solve'])) / (timestep * n_timesteps):.3f} wall s per sim s", flush=True)
        del sim
    sim_s = timestep * n_timesteps
    metrics["time_per_sim_second"] = [float(np.mean(t)) / sim_s for t in metrics["times"]["vbd_solve"]]
    metrics["realtime_factor_per_env"] = [n * sim_s / float(np.mean(t))
                                          for n, t in zip(metrics["n_simulations"], metrics["times"]["vbd_solve"])]
    if out:
        np.savez_compressed(out, n_vertices=np.array(metrics["n_vertices"]),
                            n_simulations=np.array(metrics["n_simulations"]),
                            times=np.array(metrics["times"]["vbd_solve"]),
                            time_per_sim_second=np.array(metrics["time_per_sim_second"]),
                            realtime_factor_per_env=np.array(metrics["realtime_factor_per_env"]))
    return metrics


def benchmark_coloring(num_voxels=(5, 8, 12, 16, 20), num_samples=5, out=None):
    """Time of the reference's two colouring interfaces on nx^3-voxel cubes,
    both from the element array (examples/benchmark_graph_coloring.py:24-45):
    "numpy" is graph_coloring on the adjacency dict compute_adjacency_dict
    builds (graph_coloring takes that dict, src/coloring.py:40; the dict is
    built and flattened on the host, the first-fit rounds run on the GPU),
    "parallel" is graph_coloring_wp on the elements (adjacency and colouring
    on the GPU).  The two colourings are checked equal at every size."""
    from .coloring import compute_adjacency_dict, graph_coloring, graph_coloring_wp

    def graph_coloring_host(elements):
        return graph_coloring(compute_adjacency_dict(elements)[0], verbose=False)[0]

    metrics = {"num_vertices": [], "num_elements": [], "times": {"numpy": [], "parallel": []}}
    for nx in num_voxels:
        V, H = voxel2hex(np.ones((nx, nx, nx), dtype=bool), 1.0, 1.0, 1.0)
        T = hex2tets(H)
        if not np.array_equal(graph_coloring_host(T), graph_coloring_wp(T)):
            raise AssertionError(f"host and GPU colourings differ on the {nx}^3 cube")
        times = benchmark_functions([graph_coloring_host, graph_coloring_wp], T, num_samples=num_samples)
        metrics["num_vertices"].append(len(V))
        metrics["num_elements"].append(len(T))
        metrics["times"]["numpy"].append(times["graph_coloring_host"])
        metrics["times"]["parallel"].append(times["graph_coloring_wp"])
    if out:
        np.savez_compressed(out, num_vertices=np.array(metrics["num_vertices"]),
                            num_elements=np.array(metrics["num_elements"]),
                            numpy=np.array(metrics["times"]["numpy"]), parallel=np.array(metrics["times"]["parallel"]))
    return metrics


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("what", choices=["run_cantilever", "run_mass_spring", "run_mass_spring_parallel", "tolerance",
                                    "mesh_size", "parallel", "coloring"])
    p.add_argument("--out", default="outputs")
    p.add_argument("--quick", action="store_true", help="fewer samples and steps")
    p.add_argument("--batched", action="store_true", help="parallel scenes as a BatchedVBDSolver")
    a = p.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    t0 = time.perf_counter()
    q = a.quick
    if a.what == "run_cantilever":
        r = run_cantilever(os.path.join(a.out, "cantilever"), n_seconds=0.3 if q else 1.5)
        print(f"tip z min {r[:, 2].min():.4e} at step {int(r[:, 2].argmin())}")
    elif a.what == "run_mass_spring":
        r = run_mass_spring(os.path.join(a.out, "mass_spring"), n_seconds=0.5 if q else 3.0)
        print(f"tip z min {r[:, 2].min():.4e}")
    elif a.what == "run_mass_spring_parallel":
        r = run_mass_spring_parallel(os.path.join(a.out, "mass_spring_parallel"), n_seconds=0.5 if q else 5.0,
                                     batched=a.batched)
        print({k: float(v[:, 2].min()) for k, v in r.items()})
    elif a.what == "tolerance":
        m = benchmark_tolerance(num_samples=1 if q else 3, n_timesteps=200 if q else 1000,
                                out=os.path.join(a.out, "benchmark_tolerance_metrics.npz"))
        print(dict(zip(m["tolerance"], m["time_per_sim_second"])))
    elif a.what == "mesh_size":
        m = benchmark_mesh_size(reduction_factors=(1, 2, 4) if q else tuple(range(1, 11)),
                                n_timesteps=50 if q else 300, out=os.path.join(a.out, "benchmark_mesh_size_metrics.npz"))
        print(dict(zip(m["num_vertices"], m["time_per_sim_second"])))
    elif a.what == "parallel":
        m = benchmark_parallel(nsims=(1, 4, 16, 64) if q else tuple(2 ** i for i in range(13)),
                               num_samples=1 if q else 3, n_timesteps=25 if q else 125, batched=a.batched,
                               out=os.path.join(a.out, "benchmark_parallel_metrics.npz"))
        print(dict(zip(m["n_simulations"], m["realtime_factor_per_env"])))
    else:
        m = benchmark_coloring(num_voxels=(5, 8) if q else (5, 8, 12, 16, 20), num_samples=2 if q else 5,
                               out=os.path.join(a.out, "benchmark_coloring_metrics.npz"))
        print({n: (float(np.mean(a_)), float(np.mean(b_))) for n, a_, b_ in
               zip(m["num_vertices"], m["times"]["numpy"], m["times"]["parallel"])})
    print(f"{a.what} done in {time.perf_counter() - t0:.1f} s")


if __name__ == "__main__":
    main()
