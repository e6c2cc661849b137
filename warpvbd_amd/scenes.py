"""Scene builders for the benchmark and test workloads.

They restate the mesh, pinning and probe rules of the reference's demo scenes
(examples/run_cantilever.py:28-47, examples/run_mass_spring.py:31-54) with the
sizes and materials of the workloads this package is measured on.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._utils import hex2tets, voxel2hex


@dataclass
class Scene:
    name: str
    vertices: np.ndarray        # [N, 3] float64
    elements: np.ndarray        # [M, 4] int64
    active_mask: np.ndarray     # [N] bool (False = pinned)
    tip_idx: np.ndarray         # probe vertices whose mean z is traced
    density: np.ndarray         # [M]
    youngs_modulus: np.ndarray  # [M]
    poissons_ratio: np.ndarray  # [M]
    gravity: tuple = (0.0, 0.0, -9.81)
    meta: dict = field(default_factory=dict)

    @property
    def n_vertices(self):
        return len(self.vertices)

    @property
    def n_elements(self):
        return len(self.elements)

    def solver_kwargs(self):
        return dict(initial_positions=self.vertices, elements=self.elements, densities=self.density,
                    youngs_modulus=self.youngs_modulus, poissons_ratio=self.poissons_ratio,
                    active_mask=self.active_mask, gravity=self.gravity, damping_coefficient=0.0)


def cantilever(nx=(10, 2, 2), h=0.01, youngs_modulus=1e5, poissons_ratio=0.4, density=1000.0,
               gravity=(0.0, 0.0, -9.81)) -> Scene:
    """Voxel beam along x, the x = 0 face pinned, the x = L face probed
    (examples/run_cantilever.py:28-46)."""
    vox = np.ones(nx, dtype=bool)
    V, H = voxel2hex(vox, h, h, h)
    T = hex2tets(H)
    active = ~(V[:, 0] < 1e-3)
    tip = np.nonzero(active & (V[:, 0] > nx[0] * h - 1e-3))[0]
    m = len(T)
    return Scene(f"cantilever{nx[0]}x{nx[1]}x{nx[2]}", V, T, active, tip, np.full(m, float(density)),
                 np.full(m, float(youngs_modulus)), np.full(m, float(poissons_ratio)), tuple(gravity),
                 {"nx": tuple(nx), "h": h})


def mass_spring(cube=6, column=3, h=0.01, youngs_modulus=1e5, poissons_ratio=0.4, density=1000.0,
                gravity=(0.0, 0.0, -9.81)) -> Scene:
    """A cube of cube^3 voxels hanging from a 1x1-voxel column of `column`
    voxels whose top face is pinned (layout of examples/run_mass_spring.py:32-53:
    column at voxel (cube//2, cube//2), pin z > top - 1e-3, probe z < 1e-3)."""
    vox = np.zeros((cube, cube, cube + column), dtype=bool)
    vox[:, :, :cube] = True
    vox[cube // 2, cube // 2, cube:] = True
    V, H = voxel2hex(vox, h, h, h)
    T = hex2tets(H)
    top = (cube + column) * h
    active = ~(V[:, 2] > top - 1e-3)
    tip = np.nonzero(active & (V[:, 2] < 1e-3))[0]
    m = len(T)
    return Scene(f"massspring{cube}+{column}", V, T, active, tip, np.full(m, float(density)),
                 np.full(m, float(youngs_modulus)), np.full(m, float(poissons_ratio)), tuple(gravity),
                 {"cube": cube, "column": column, "h": h})


def mass_spring_batch(n_bodies=512, seed=3, youngs_modulus=1e5, poissons_ratio=0.4, density=1000.0):
    """Bodies of the batched mass-spring workload: cube edge uniform in 6..14
    voxels, column length uniform in 3..12 voxels, drawn from `seed`."""
    rng = np.random.default_rng(seed)
    cubes = rng.integers(6, 15, size=n_bodies)
    cols = rng.integers(3, 13, size=n_bodies)
    return [mass_spring(int(c), int(s), youngs_modulus=youngs_modulus, poissons_ratio=poissons_ratio,
                        density=density) for c, s in zip(cubes, cols)]


def mixed_batch(n_bodies=1024, seed=5, density=1000.0):
    """Bodies of the mixed workload: half cantilevers (length uniform in 16..64
    voxels, square cross-section 4..8 voxels), half mass-springs as in
    mass_spring_batch; E log-uniform in [1e4, 1e6] Pa, nu uniform in
    [0.3, 0.45], drawn from `seed`."""
    rng = np.random.default_rng(seed)
    half = n_bodies // 2
    bodies = []
    for i in range(n_bodies):
        E = float(10 ** rng.uniform(4.0, 6.0))
        nu = float(rng.uniform(0.3, 0.45))
        if i < half:
            L, c = int(rng.integers(16, 65)), int(rng.integers(4, 9))
            bodies.append(cantilever((L, c, c), youngs_modulus=E, poissons_ratio=nu, density=density))
        else:
            m, s = int(rng.integers(6, 15)), int(rng.integers(3, 13))
            bodies.append(mass_spring(m, s, youngs_modulus=E, poissons_ratio=nu, density=density))
    return bodies
