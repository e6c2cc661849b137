"""Many soft bodies of uneven size in one solver.

The reference simulates parallel environments by concatenating them into one
mesh handled by one ``VBDSolver`` with one global convergence test
(examples/run_mass_spring_parallel.py:35-119).  ``BatchedVBDSolver`` keeps that
input form (a list of bodies) but gives every body its own convergence test,
gravity and material, and runs all bodies in one launch per colour (replayed
from a captured hipGraph) or one workgroup per body, whichever the autotuner
measures faster.  Each body's trajectory equals its own single-body solve.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os

import numpy as np

from . import _lib
from ._lib import check, f64, i32, ptr, to_numpy
from .vbd_solver import VBDSolver, VertexArray

MODES = {"launches": 1, "resident": 2, "launches_full": 3}
_MODE_NAMES = {v: k for k, v in MODES.items()}


def _cache_path():
    return os.environ.get("WARPVBD_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "warpvbd_amd", "autotune.json"))


def concat_bodies(bodies):
    """Concatenate bodies into one mesh (the reference's parallel-environment
    form, examples/run_mass_spring_parallel.py:35-119).  Returns a dict with
    initial_positions, elements, densities, youngs_modulus, poissons_ratio,
    active_mask, gravity (per body, [B, 3]) and body_vertex_offsets."""
    Xs, Ts, rho, E, nu, act, grav, voff = [], [], [], [], [], [], [], [0]
    for b in bodies:
        d = b if isinstance(b, dict) else b.solver_kwargs()
        X = f64(d["initial_positions"]).reshape(-1, 3)
        T = np.asarray(to_numpy(d["elements"]), dtype=np.int64).reshape(-1, 4)
        m = len(T)
        Xs.append(X)
        Ts.append(T + voff[-1])
        rho.append(np.broadcast_to(f64(d["densities"]), (m,)))
        E.append(np.broadcast_to(f64(d["youngs_modulus"]), (m,)))
        nu.append(np.broadcast_to(f64(d["poissons_ratio"]), (m,)))
        a = d.get("active_mask")
        act.append(np.ones(len(X), bool) if a is None else np.asarray(to_numpy(a), bool).reshape(len(X)))
        grav.append(f64(d.get("gravity", (0.0, 0.0, -9.81))).reshape(3))
        voff.append(voff[-1] + len(X))
    return dict(initial_positions=np.ascontiguousarray(np.concatenate(Xs)), elements=i32(np.concatenate(Ts)),
                densities=np.ascontiguousarray(np.concatenate(rho)),
                youngs_modulus=np.ascontiguousarray(np.concatenate(E)),
                poissons_ratio=np.ascontiguousarray(np.concatenate(nu)), active_mask=np.concatenate(act),
                gravity=np.ascontiguousarray(np.stack(grav)), body_vertex_offsets=np.array(voff, dtype=np.int32))


class BatchedVBDSolver(VBDSolver):
    """Bodies are objects with ``vertices``/``initial_positions``, ``elements``,
    per-element ``density``/``densities``, ``youngs_modulus``, ``poissons_ratio``,
    ``active_mask`` and ``gravity`` (e.g. :class:`warpvbd_amd.scenes.Scene`) or
    dicts with the ``VBDSolver`` keyword names."""

    def __init__(self, bodies, damping_coefficient: float = 0.0, dx_tol: float = 1e-9, max_iter: int = 100,
                 device="cuda", material: str = "snh", verbose: bool = True):
        L = _lib.load()
        self.device = device
        self.device_index = _lib.use_device(device)
        cb = concat_bodies(bodies)
        X, T = cb["initial_positions"], cb["elements"]
        rho, E, nu = cb["densities"], cb["youngs_modulus"], cb["poissons_ratio"]
        actu8 = np.ascontiguousarray(cb["active_mask"], dtype=np.uint8)
        G = cb["gravity"]
        voff = cb["body_vertex_offsets"]
        self.body_vertex_offsets = np.array(voff, dtype=np.int32)
        self.n_bodies = len(bodies)
        n, m = len(X), len(T)
        self.initial_positions, self.elements = X, T
        self.densities, self.youngs_modulus, self.poissons_ratio = rho, E, nu
        self.damping_coefficient = float(damping_coefficient)
        self.gravity = G
        self.dx_tol, self.max_iter, self.material = dx_tol, max_iter, material
        self.active_mask = actu8.astype(bool)
        self.n_vertices, self.n_elements = n, m
        self.last_iterations, self.last_metric = 0, 0.0
        self.body_iterations = np.zeros(self.n_bodies, np.int32)
        self.body_metrics = np.zeros(self.n_bodies)
        self._version = 0
        import weakref
        self._views = weakref.WeakSet()
        self._free_slots = list(range(_lib.VBD_MAX_SNAPSHOTS))
        h = ctypes.c_void_p()
        check(L.vbd_batch_create(ptr(X, ctypes.c_double), n, ptr(T, ctypes.c_int32), m,
                                 ptr(self.body_vertex_offsets, ctypes.c_int32), self.n_bodies,
                                 ptr(rho, ctypes.c_double), ptr(E, ctypes.c_double), ptr(nu, ctypes.c_double),
                                 ptr(actu8, ctypes.c_uint8), ptr(G, ctypes.c_double), self.damping_coefficient,
                                 _lib.MATERIALS[material], ctypes.byref(h)), "vbd_batch_create")
        self._h = h
        info = self.info()
        self.n_colors = info["n_colors"]
        self.max_incident_elements = info["max_incident"]
        self.vertex_valence = info["max_degree"]
        self._color_groups = None
        self._adj_v2e = None
        if verbose:
            print(f"Batched solver: {self.n_bodies} bodies, {n} vertices, {m} elements, {self.n_colors} colors")

    # ------------------------------------------------------------------ stepping
    def step(self, positions=None, dt: float = 1e-3, dx_tol: float = None, max_iter: int = None,
             warm_start=None) -> VertexArray:
        """Advance every body by dt.  ``positions`` (optional, all bodies
        concatenated) overrides x_prev as in ``VBDSolver.step``.  With
        ``warm_start`` each body starts from the network guess when that wins
        the first-sweep comparison; ``last_used_network`` is the number of
        bodies that did."""
        if positions is not None and not (isinstance(positions, VertexArray) and positions._is_current(self)):
            self.old_positions = positions
        tol = self.dx_tol if dx_tol is None else dx_tol
        mi = self.max_iter if max_iter is None else max_iter
        self._before_mutation()
        its = np.empty(self.n_bodies, np.int32)
        mets = np.empty(self.n_bodies)
        if warm_start is not None:
            it, met, used = ctypes.c_int32(), ctypes.c_double(), ctypes.c_int32()
            check(_lib.load().vbd_step_warm(self._h, warm_start._h, float(dt), float(tol), int(mi), ctypes.byref(it),
                                            ctypes.byref(met), ctypes.byref(used)), "vbd_step_warm")
            self.last_used_network = int(used.value)
        check(_lib.load().vbd_batch_step(self._h, float(dt), float(tol), int(mi), ptr(its, ctypes.c_int32),
                                         ptr(mets, ctypes.c_double)) if warm_start is None else
              _lib.load().vbd_batch_results(self._h, ptr(its, ctypes.c_int32), ptr(mets, ctypes.c_double)),
              "vbd_batch_step")
        self.body_iterations, self.body_metrics = its, mets
        self.last_iterations = int(its.max()) if len(its) else 0
        self.last_metric = float(np.nanmax(mets)) if len(mets) else 0.0
        return self._view()

    def body_slice(self, b: int) -> slice:
        return slice(int(self.body_vertex_offsets[b]), int(self.body_vertex_offsets[b + 1]))

    # ------------------------------------------------------------------ execution form
    def set_mode(self, mode="launches", lanes_per_vertex=4, block_size=256, use_graph=True):
        check(_lib.load().vbd_batch_set_mode(self._h, MODES[mode], lanes_per_vertex, block_size, int(use_graph)),
              "vbd_batch_set_mode")

    def get_mode(self) -> dict:
        m, l, b, g = (ctypes.c_int32() for _ in range(4))
        check(_lib.load().vbd_batch_get_mode(self._h, *(ctypes.byref(x) for x in (m, l, b, g))), "vbd_batch_get_mode")
        return {"mode": _MODE_NAMES[m.value], "lanes_per_vertex": l.value, "block_size": b.value,
                "use_graph": bool(g.value)}

    def _signature(self, dt, tol, max_iter):
        h = hashlib.sha1()
        h.update(np.asarray(self.body_vertex_offsets).tobytes())
        h.update(self.color_sizes().tobytes())
        h.update(f"{self.n_elements}|{self.material}|{dt}|{tol}|{max_iter}".encode())
        return h.hexdigest()

    def autotune(self, dt: float, dx_tol: float = None, max_iter: int = None, trial_steps: int = 3,
                 cache: bool = True) -> dict:
        """Pick the execution form, lanes per vertex and block size by timing
        every candidate from the current state (which is restored).  Results are
        cached on disk by mesh signature."""
        tol = self.dx_tol if dx_tol is None else dx_tol
        mi = self.max_iter if max_iter is None else max_iter
        key = self._signature(dt, tol, mi)
        path = _cache_path()
        if cache and os.path.exists(path):
            try:
                with open(path) as f:
                    entry = json.load(f).get(key)
                if entry:
                    self.set_mode(entry["mode"], entry["lanes_per_vertex"], entry["block_size"], entry["use_graph"])
                    return dict(entry, cached=True)
            except (OSError, ValueError, KeyError):
                pass
        best = ctypes.c_double()
        check(_lib.load().vbd_batch_autotune(self._h, float(dt), float(tol), int(mi), int(trial_steps),
                                             ctypes.byref(best)), "vbd_batch_autotune")
        entry = dict(self.get_mode(), ms_per_step=best.value)
        if cache:
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                data = {}
                if os.path.exists(path):
                    with open(path) as f:
                        data = json.load(f)
                data[key] = entry
                with open(path, "w") as f:
                    json.dump(data, f, indent=1)
            except (OSError, ValueError):
                pass
        return dict(entry, cached=False)
