"""``VBDSolver`` with the constructor, attributes and ``step`` of WarpVBD's
``src/vbd_solver.py:366-612``, running on an MI355X through libwarpvbd.so.

Differences a caller can see (all additive):

* ``material='snh'`` (default, the reference's live energy) or ``'stvk'``
  (the reference's commented-out energy, src/vbd_solver.py:127-144).
* ``step`` returns a :class:`VertexArray` whose ``.numpy()`` gives float64
  positions, like the ``wp.array`` the reference returns.  The array keeps its
  values after later steps (a device snapshot is taken when needed).
* ``last_iterations`` / ``last_metric`` report what the reference only prints.
* Positions are stored on the device as displacements from the rest shape:
  the iterate packed to 32 significant bits (16 bytes per vertex), the step's
  start and inertia target in FP64.  The per-element arithmetic is FP32 (the
  reference computes in FP64).
"""
from __future__ import annotations

import ctypes
import time
import weakref

import numpy as np

from . import _lib
from ._lib import VBDError, check, f64, i32, ptr, to_numpy

EPS = 1e-12                      # src/vbd_solver.py:11
MAX_ELEMENTS_PER_VERTEX = 32     # src/vbd_solver.py:12


def _is_gpu_tensor(x) -> bool:
    return type(x).__module__.startswith("torch") and bool(getattr(x, "is_cuda", False))


class VertexArray:
    """An [N, 3] float64 vertex array that lives on the solver's device.

    ``.numpy()`` copies it to the host.  Indexing and ``np.asarray`` work on the
    host copy."""

    __slots__ = ("_solver", "_version", "_slot", "_host", "__weakref__")

    def __init__(self, solver=None, version=None, host=None):
        self._solver = solver
        self._version = version
        self._slot = None
        self._host = host

    # -- state transitions driven by the solver
    def _preserve(self):
        s = self._solver
        if self._host is not None or self._slot is not None or s is None:
            return
        slot = s._take_snapshot_slot()
        if slot is None:
            self._host = s._read_current()
        else:
            check(_lib.load().vbd_snapshot(s._h, slot), "vbd_snapshot")
            self._slot = slot

    def _is_current(self, solver):
        return self._solver is solver and self._host is None and self._slot is None and \
            self._version == solver._version

    def numpy(self) -> np.ndarray:
        if self._host is None:
            s = self._solver
            if self._slot is not None:
                out = np.empty((s.n_vertices, 3))
                check(_lib.load().vbd_read_snapshot(s._h, self._slot, ptr(out, ctypes.c_double)), "vbd_read_snapshot")
                s._release_snapshot_slot(self._slot)
                self._slot = None
                self._host = out
            else:
                self._host = s._read_current()
        return self._host

    def __del__(self):
        s, slot = self._solver, self._slot
        if s is not None and slot is not None:
            try:
                s._release_snapshot_slot(slot)
            except Exception:
                pass

    @property
    def shape(self):
        return (self._solver.n_vertices, 3) if self._host is None else self._host.shape

    def __len__(self):
        return self.shape[0]

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)

    def __getitem__(self, idx):
        return self.numpy()[idx]


class VBDSolver:
    """Vertex Block Descent solver for tetrahedral meshes (src/vbd_solver.py:366-612).

    Args mirror the reference: per-element ``densities``, ``youngs_modulus`` and
    ``poissons_ratio`` (scalars are broadcast), ``damping_coefficient`` (default
    1.0 as in the reference; its scenes pass 0.0), optional per-vertex
    ``active_mask`` (False = pinned), ``gravity``, the convergence tolerance
    ``dx_tol`` on max|dx_x+dx_y+dx_z| and ``max_iter``.
    """

    def __init__(self, initial_positions, elements, densities, youngs_modulus, poissons_ratio,
                 damping_coefficient: float = 1.0, active_mask=None, gravity=(0.0, 0.0, -9.81),
                 dx_tol: float = 1e-9, max_iter: int = 100, device="cuda", material: str = "snh",
                 verbose: bool = True):
        L = _lib.load()
        self.device = device
        self.device_index = _lib.use_device(device)
        X = f64(initial_positions).reshape(-1, 3)
        T = i32(elements).reshape(-1, 4)
        n, m = len(X), len(T)
        rho = f64(np.broadcast_to(to_numpy(densities), (m,)))
        E = f64(np.broadcast_to(to_numpy(youngs_modulus), (m,)))
        nu = f64(np.broadcast_to(to_numpy(poissons_ratio), (m,)))
        act = None if active_mask is None else np.ascontiguousarray(to_numpy(active_mask), dtype=np.uint8)
        if act is not None and act.shape != (n,):
            raise ValueError(f"active_mask must have shape ({n},)")
        g = f64(gravity).reshape(3)
        if material not in _lib.MATERIALS:
            raise ValueError(f"material must be one of {sorted(_lib.MATERIALS)}")

        self.initial_positions = X
        self.elements = T
        self.densities, self.youngs_modulus, self.poissons_ratio = rho, E, nu
        self.damping_coefficient = float(damping_coefficient)
        self.gravity = tuple(float(x) for x in g)
        self.dx_tol = dx_tol
        self.max_iter = max_iter
        self.material = material
        self.active_mask = np.ones(n, dtype=bool) if act is None else act.astype(bool)
        self.n_vertices, self.n_elements = n, m
        self.last_iterations, self.last_metric = 0, 0.0
        self.last_dx_max = float("nan")  # set by a step that ran its last allowed iteration
        self.last_used_network = False
        self._version = 0
        self._views = weakref.WeakSet()
        self._free_slots = list(range(_lib.VBD_MAX_SNAPSHOTS))

        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        check(L.vbd_create(ptr(X, ctypes.c_double), n, ptr(T, ctypes.c_int32), m, ptr(rho, ctypes.c_double),
                           ptr(E, ctypes.c_double), ptr(nu, ctypes.c_double),
                           None if act is None else ptr(act, ctypes.c_uint8), ptr(g, ctypes.c_double),
                           self.damping_coefficient, _lib.MATERIALS[material], ctypes.byref(h)), "vbd_create")
        self._h = h
        t1 = time.perf_counter()
        info = self.info()
        self.n_colors = info["n_colors"]
        self.max_incident_elements = info["max_incident"]
        self.vertex_valence = info["max_degree"]
        if verbose:
            sizes = self.color_sizes()
            print(f"Assigned {self.n_colors} colors in {(t1 - t0) * 1000:.4f}ms. Vertex valence: "
                  f"{self.vertex_valence}")
            print(f"Max incident elements per vertex: {self.max_incident_elements}")
            print(f"Active vertices per color: {sizes.tolist()}")
        self._color_groups = None
        self._adj_v2e = None

    # ------------------------------------------------------------------ lifetime
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            for v in list(self._views):
                if v._host is None:
                    try:
                        v.numpy()
                    except Exception:
                        pass
            _lib.load().vbd_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ snapshots
    def _take_snapshot_slot(self):
        return self._free_slots.pop() if self._free_slots else None

    def _release_snapshot_slot(self, slot):
        if slot not in self._free_slots:
            self._free_slots.append(slot)

    def _read_current(self) -> np.ndarray:
        out = np.empty((self.n_vertices, 3))
        check(_lib.load().vbd_get_state(self._h, ptr(out, ctypes.c_double), None), "vbd_get_state")
        return out

    def _before_mutation(self):
        for v in list(self._views):
            if v._is_current(self):
                v._preserve()
        self._version += 1

    def _view(self) -> VertexArray:
        v = VertexArray(self, self._version)
        self._views.add(v)
        return v

    # ------------------------------------------------------------------ state (src/vbd_solver.py:402-417)
    @property
    def old_positions(self) -> VertexArray:
        return self._view()

    @old_positions.setter
    def old_positions(self, x):
        if isinstance(x, VertexArray) and x._is_current(self):
            return
        xs = f64(x).reshape(self.n_vertices, 3)
        self._before_mutation()
        check(_lib.load().vbd_set_state(self._h, ptr(xs, ctypes.c_double), None), "vbd_set_state")

    @property
    def old_velocities(self) -> VertexArray:
        v = np.empty((self.n_vertices, 3))
        check(_lib.load().vbd_get_state(self._h, None, ptr(v, ctypes.c_double)), "vbd_get_state")
        return VertexArray(host=v)

    @old_velocities.setter
    def old_velocities(self, v):
        vs = f64(v).reshape(self.n_vertices, 3)
        check(_lib.load().vbd_set_state(self._h, None, ptr(vs, ctypes.c_double)), "vbd_set_state")

    def reset(self) -> None:
        """x_prev <- initial positions, v_prev <- 0 (src/vbd_solver.py:501-506)."""
        self._before_mutation()
        check(_lib.load().vbd_reset(self._h), "vbd_reset")

    # ------------------------------------------------------------------ time step (src/vbd_solver.py:509-612)
    def step(self, positions, dt: float, dx_tol: float = None, max_iter: int = None,
             warm_start=None) -> VertexArray:
        """One implicit-Euler step from ``positions`` (normally ``self.old_positions``).

        Returns the new positions; the input is not overwritten.  ``warm_start``
        (a :class:`warpvbd_amd.warm_start.WarmStart` built on this solver)
        starts the iterations from the initial guess its policy picks (by
        default the one of the plain, network and history guesses with the
        smallest first-sweep residual); ``last_used_network`` records whether
        that was the network's."""
        if dx_tol is None:
            dx_tol = self.dx_tol
        if max_iter is None:
            max_iter = self.max_iter
        if _is_gpu_tensor(positions) and warm_start is None:
            return self._step_tensor(positions, dt, dx_tol, max_iter)
        if _is_gpu_tensor(positions):
            positions = positions.detach().cpu().numpy()
        L = _lib.load()
        pos_ptr = None
        if isinstance(positions, VertexArray) and positions._solver is self and positions._host is None:
            # an older array is loaded from its device snapshot; the current
            # one (no snapshot yet) is x_prev already
            load_slot = positions._slot
            if load_slot is None and positions._version != self._version:
                raise VBDError("stale position array")
            # the arrays that still show the current positions keep them
            # before a snapshot load overwrites x_prev
            self._before_mutation()
            if load_slot is not None:
                check(L.vbd_load_snapshot(self._h, load_slot), "vbd_load_snapshot")
        else:
            xs = f64(positions).reshape(self.n_vertices, 3)
            pos_ptr = ptr(xs, ctypes.c_double)
            self._before_mutation()
        it, met = ctypes.c_int32(0), ctypes.c_double(0.0)
        if warm_start is not None:
            if pos_ptr is not None:
                check(L.vbd_set_state(self._h, pos_ptr, None), "vbd_set_state")
            used = ctypes.c_int32(0)
            check(L.vbd_step_warm(self._h, warm_start._h, float(dt), float(dx_tol), int(max_iter), ctypes.byref(it),
                                  ctypes.byref(met), ctypes.byref(used)), "vbd_step_warm")
            self.last_used_network = bool(used.value)
        else:
            check(L.vbd_step(self._h, pos_ptr, float(dt), float(dx_tol), int(max_iter), None, ctypes.byref(it),
                             ctypes.byref(met)), "vbd_step")
            self.last_used_network = False
        self.last_iterations, self.last_metric = it.value, met.value
        if it.value == max_iter:
            self._warn_not_converged()
        else:
            self.last_dx_max = float("nan")
        return self._view()

    def _warn_not_converged(self):
        """The reference's warning when the last allowed iteration ran, with
        the largest |dx| component of that iteration (src/vbd_solver.py:581-582)."""
        d = ctypes.c_double()
        check(_lib.load().vbd_last_dx_max(self._h, ctypes.byref(d)), "vbd_last_dx_max")
        self.last_dx_max = d.value
        print(f"Warning: VBD solver did not converge within the maximum number of iterations. "
              f"Final dx max: {d.value}")

    def _step_tensor(self, positions, dt, dx_tol, max_iter):
        """step() for an [N, 3] tensor on the solver's GPU: positions in and the
        new positions out stay in device memory (a new float64 tensor), ordered
        on torch's current stream."""
        import torch
        x = positions.detach().to(torch.float64).reshape(self.n_vertices, 3).contiguous()
        if x.device.index != self.device_index:
            raise VBDError(f"positions are on {x.device}, the solver on cuda:{self.device_index}")
        out = torch.empty_like(x)
        stream = torch.cuda.current_stream(x.device).cuda_stream
        self._before_mutation()
        it, met = ctypes.c_int32(0), ctypes.c_double(0.0)
        check(_lib.load().vbd_step_dev(self._h, ctypes.c_void_p(x.data_ptr()), float(dt), float(dx_tol),
                                       int(max_iter), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream),
                                       ctypes.byref(it), ctypes.byref(met)), "vbd_step_dev")
        self.last_iterations, self.last_metric = it.value, met.value
        self.last_used_network = False
        if it.value == max_iter:
            self._warn_not_converged()
        else:
            self.last_dx_max = float("nan")
        return out

    # ------------------------------------------------------------------ material / gravity
    def set_material(self, youngs_modulus: float, poissons_ratio: float):
        """Give every element a new (E, nu); the elements must share one
        material already."""
        check(_lib.load().vbd_set_uniform_material(self._h, float(youngs_modulus), float(poissons_ratio)),
              "vbd_set_uniform_material")
        self.youngs_modulus = np.full(self.n_elements, float(youngs_modulus))
        self.poissons_ratio = np.full(self.n_elements, float(poissons_ratio))

    def set_gravity(self, gravity):
        g = f64(gravity).reshape(3)
        check(_lib.load().vbd_set_gravity(self._h, ptr(g, ctypes.c_double)), "vbd_set_gravity")
        self.gravity = tuple(float(x) for x in g)

    # ------------------------------------------------------------------ mesh facts
    def info(self) -> dict:
        n, m, c, w, d = (ctypes.c_int32() for _ in range(5))
        s = ctypes.c_int64()
        check(_lib.load().vbd_info(self._h, *(ctypes.byref(x) for x in (n, m, c, w, d, s))), "vbd_info")
        return {"n_vertices": n.value, "n_elements": m.value, "n_colors": c.value, "max_incident": w.value,
                "max_degree": d.value, "n_slots": s.value}

    def colors(self) -> np.ndarray:
        out = np.empty(self.n_vertices, np.int32)
        check(_lib.load().vbd_get_colors(self._h, ptr(out, ctypes.c_int32)), "vbd_get_colors")
        return out

    def color_sizes(self) -> np.ndarray:
        out = np.empty(max(1, self.n_colors), np.int32)
        check(_lib.load().vbd_get_color_sizes(self._h, ptr(out, ctypes.c_int32)), "vbd_get_color_sizes")
        return out[: self.n_colors]

    def staged_colors(self) -> int:
        """Colours whose sweep reads its neighbours and constants through
        LDS staging (4 lanes per vertex, 256 threads; same results)."""
        n = ctypes.c_int32()
        check(_lib.load().vbd_staged_colors(self._h, ctypes.byref(n)), "vbd_staged_colors")
        return n.value

    def tail_sweep(self) -> tuple:
        """(colours, clusters) of the tail sweep: the trailing small colours
        swept in one launch, one block per cluster of their vertices ((0, 0):
        a launch per colour; same results either way)."""
        k, c = ctypes.c_int32(), ctypes.c_int32()
        check(_lib.load().vbd_tail_sweep(self._h, ctypes.byref(k), ctypes.byref(c)), "vbd_tail_sweep")
        return k.value, c.value

    def record_layout(self) -> dict:
        """Slot records of this solver: record dictionary entries (as many as
        slots when no constants are shared) and HBM bytes per slot read by a
        sweep (16 with shared entries, else 64).  edge_form: whether the
        entries also hold their rest edges, from which the SNH sweeps form
        cof(F) b and J - 1."""
        n, b, e = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(_lib.load().vbd_records(self._h, ctypes.byref(n), ctypes.byref(b)), "vbd_records")
        check(_lib.load().vbd_edge_form(self._h, ctypes.byref(e)), "vbd_edge_form")
        return {"dictionary_entries": n.value, "bytes_per_slot": b.value, "edge_form": bool(e.value)}

    def color_layout_bytes(self) -> np.ndarray:
        """HBM bytes of one sweep of each colour with this solver's layout:
        68 per swept vertex, 16 or 64 per slot, 16 per gathered neighbour
        (color_bytes: the algorithmic model with the reference's 68 bytes per
        slot)."""
        out = np.empty(max(1, self.n_colors))
        check(_lib.load().vbd_color_layout_bytes(self._h, ptr(out, ctypes.c_double)), "vbd_color_layout_bytes")
        return out[: self.n_colors]

    def color_bytes(self) -> np.ndarray:
        out = np.empty(max(1, self.n_colors))
        check(_lib.load().vbd_color_bytes(self._h, ptr(out, ctypes.c_double)), "vbd_color_bytes")
        return out[: self.n_colors]

    @property
    def color_groups(self) -> np.ndarray:
        """[n_colors, max group] int32, vertices ascending, padded with -1
        (src/vbd_solver.py:476-483)."""
        if self._color_groups is None:
            col = self.colors()
            groups = [np.nonzero(col == c)[0] for c in range(self.n_colors)]
            gmax = max((len(g) for g in groups), default=0)
            out = np.full((self.n_colors, gmax), -1, dtype=np.int32)
            for c, g in enumerate(groups):
                out[c, : len(g)] = g
            self._color_groups = out
        return self._color_groups

    @property
    def adj_v2e(self) -> np.ndarray:
        """[n_vertices, max incident] int32, ascending element ids padded with -1
        (src/vbd_solver.py:485-498)."""
        if self._adj_v2e is None:
            out = np.empty((self.n_vertices, max(1, self.max_incident_elements)), np.int32)
            check(_lib.load().vbd_get_incidence(self._h, ptr(out, ctypes.c_int32)), "vbd_get_incidence")
            self._adj_v2e = out[:, : self.max_incident_elements]
        return self._adj_v2e

    # ------------------------------------------------------------------ execution knobs / profiling
    def set_launch(self, lanes_per_vertex=4, block_size=256, small_color=2048, chunk=0):
        check(_lib.load().vbd_set_launch(self._h, lanes_per_vertex, block_size, small_color, chunk), "vbd_set_launch")

    def set_graph(self, use_graph=True):
        """Replay each step's iterations from captured hipGraphs (default)."""
        check(_lib.load().vbd_set_graph(self._h, int(bool(use_graph))), "vbd_set_graph")

    def profile(self, enable=True, min_vertices=0):
        check(_lib.load().vbd_profile_enable(self._h, int(enable), int(min_vertices)), "vbd_profile_enable")

    def profile_read(self):
        ms, n, b = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        check(_lib.load().vbd_profile_read(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)),
              "vbd_profile_read")
        return {"ms": ms.value, "launches": n.value, "bytes": b.value}
