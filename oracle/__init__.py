"""FP64 CPU oracle for the VBD time step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this package, and only as the checker.  The product package
``warpvbd_amd`` never imports it and has no CPU code path.

Contents
--------
* ``liboracle.so`` (built from ``vbd_oracle.c`` by ``Makefile``): op-for-op FP64
  restatement of ``src/vbd_solver.py`` (kernels :11-362, ``VBDSolver`` :366-612)
  and ``src/coloring.py:24-72`` of WarpVBD.  ``OracleSolver`` wraps it.
* ``mesh_loops``: literal loop restatement of ``src/_utils.py:35-108``
  (``voxel2hex`` / ``hex2tets``), used to check the product's vectorised mesh
  generator.
* ``color_first_fit``: literal dict/set restatement of ``src/coloring.py:24-72``.

How each function is pinned (tests/test_oracle.py):
* Lame, invDm, V, m: closed forms on voxel tets.
* Elastic slot (SNH, StVK): central finite differences of the energy
  (``orc_energy``), symmetry, zero gradient at rest, rest Hessian eigenvalues
  against linear elasticity, expanded form == compact form.
* 3x3 solve: numpy.linalg.solve.
* Coloring: brute-force dict restatement, properness, first-fit mex property,
  and the known colour histograms of the voxel beams.
* Step: exact free fall without pins (dx == 0 at iteration 0), pinned vertices
  unchanged, decrease of the incremental potential per vertex update, and the
  FP64 trajectory of the 10x2x2 cantilever committed under tests/golden/.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

SNH, STVK = 0, 1


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        d, i32, u8 = ctypes.c_double, ctypes.c_int32, ctypes.c_uint8
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, P(d), ctypes.c_int, P(i32), P(d), P(d), P(d), P(u8), P(d),
                                 d, ctypes.c_int, ctypes.c_int, P(ctypes.c_int)]
        L.orc_step.restype = ctypes.c_int
        L.orc_step.argtypes = [ctypes.c_void_p, P(d), d, d, ctypes.c_int, P(d), P(ctypes.c_int), P(d)]
        L.orc_step_guess.restype = ctypes.c_int
        L.orc_step_guess.argtypes = [ctypes.c_void_p, P(d), P(d), d, d, ctypes.c_int, P(d), P(ctypes.c_int), P(d)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_reset.argtypes = [ctypes.c_void_p]
        L.orc_get_state.argtypes = [ctypes.c_void_p, P(d), P(d)]
        L.orc_set_state.argtypes = [ctypes.c_void_p, P(d), P(d)]
        L.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_num_colors.argtypes = [ctypes.c_void_p]
        L.orc_max_incident.argtypes = [ctypes.c_void_p]
        L.orc_get_colors.argtypes = [ctypes.c_void_p, P(i32)]
        L.orc_lame.argtypes = [ctypes.c_int, P(d), P(d), P(d), P(d)]
        L.orc_rest_shape.restype = ctypes.c_int
        L.orc_rest_shape.argtypes = [P(d), ctypes.c_int, P(i32), P(d), P(d), P(d), P(d)]
        L.orc_slot_grad_hess.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(d), P(d), d, d, d, P(d), P(d)]
        L.orc_energy.restype = d
        L.orc_energy.argtypes = [ctypes.c_int, P(d), P(d), d, d, d]
        L.orc_solve3.restype = ctypes.c_int
        L.orc_solve3.argtypes = [P(d), P(d), P(d)]
        L.orc_color_mesh.restype = ctypes.c_int
        L.orc_color_mesh.argtypes = [ctypes.c_int, ctypes.c_int, P(i32), P(i32), P(ctypes.c_int)]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ----------------------------------------------------------------- pieces

def lame(E, nu):
    E, nu = _f64(E), _f64(nu)
    mu, lam = np.empty_like(E), np.empty_like(E)
    lib().orc_lame(len(E), _p(E, ctypes.c_double), _p(nu, ctypes.c_double),
                   _p(mu, ctypes.c_double), _p(lam, ctypes.c_double))
    return mu, lam


def rest_shape(X, tets, rho):
    X, tets, rho = _f64(X), _i32(tets), _f64(rho)
    m = len(tets)
    invDm = np.empty((m, 3, 3))
    mass, vol = np.empty(m), np.empty(m)
    bad = lib().orc_rest_shape(_p(X, ctypes.c_double), m, _p(tets, ctypes.c_int32), _p(rho, ctypes.c_double),
                               _p(invDm, ctypes.c_double), _p(mass, ctypes.c_double), _p(vol, ctypes.c_double))
    return invDm, mass, vol, bad


def slot_grad_hess(material, vloc, xs, invDm, V, mu, lam, expanded=True):
    xs, invDm = _f64(xs).reshape(12), _f64(invDm).reshape(9)
    g, H = np.empty(3), np.empty(9)
    lib().orc_slot_grad_hess(int(material), int(bool(expanded)), int(vloc), _p(xs, ctypes.c_double),
                             _p(invDm, ctypes.c_double), float(V), float(mu), float(lam),
                             _p(g, ctypes.c_double), _p(H, ctypes.c_double))
    return g, H.reshape(3, 3)


def energy(material, xs, invDm, V, mu, lam):
    xs, invDm = _f64(xs).reshape(12), _f64(invDm).reshape(9)
    return lib().orc_energy(int(material), _p(xs, ctypes.c_double), _p(invDm, ctypes.c_double),
                            float(V), float(mu), float(lam))


def solve3(H, G):
    H, G = _f64(H).reshape(9), _f64(G).reshape(3)
    dx = np.empty(3)
    kind = lib().orc_solve3(_p(H, ctypes.c_double), _p(G, ctypes.c_double), _p(dx, ctypes.c_double))
    return dx, kind


def color_mesh(tets, n=None):
    """Sequential first-fit coloring (src/coloring.py:24-72).  Returns
    (colors[n] int32, ncolors, max_degree)."""
    tets = _i32(tets)
    if n is None:
        n = int(tets.max()) + 1
    colors = np.empty(n, np.int32)
    md = ctypes.c_int(0)
    nc = lib().orc_color_mesh(int(n), len(tets), _p(tets, ctypes.c_int32), _p(colors, ctypes.c_int32),
                              ctypes.byref(md))
    return colors, nc, md.value


# ----------------------------------------------------------------- solver

class OracleSolver:
    """FP64 restatement of ``VBDSolver`` (src/vbd_solver.py:366-612)."""

    ERRORS = {-1: "element index out of range", -2: "more than 256 colors",
              -3: "more than 32 elements at a vertex", -4: "singular rest-shape tet"}

    def __init__(self, initial_positions, elements, densities, youngs_modulus, poissons_ratio,
                 damping_coefficient=1.0, active_mask=None, gravity=(0.0, 0.0, -9.81),
                 dx_tol=1e-9, max_iter=100, material="snh", expanded=True, threads=1):
        X = _f64(initial_positions).reshape(-1, 3)
        tets = _i32(elements).reshape(-1, 4)
        self.n, self.m = len(X), len(tets)
        rho, E, nu = (_f64(np.broadcast_to(a, (self.m,))) for a in (densities, youngs_modulus, poissons_ratio))
        act = None if active_mask is None else np.ascontiguousarray(active_mask, dtype=np.uint8)
        g = _f64(gravity)
        err = ctypes.c_int(0)
        mat = {"snh": SNH, "stvk": STVK}[material] if isinstance(material, str) else int(material)
        self._keep = (X, tets, rho, E, nu, act, g)
        h = lib().orc_create(self.n, _p(X, ctypes.c_double), self.m, _p(tets, ctypes.c_int32),
                             _p(rho, ctypes.c_double), _p(E, ctypes.c_double), _p(nu, ctypes.c_double),
                             None if act is None else _p(act, ctypes.c_uint8), _p(g, ctypes.c_double),
                             float(damping_coefficient), mat, int(bool(expanded)), ctypes.byref(err))
        if not h:
            raise ValueError(f"oracle create failed: {self.ERRORS.get(err.value, err.value)}")
        self._h = h
        self.dx_tol, self.max_iter = dx_tol, max_iter
        self.last_iters, self.last_metric = 0, 0.0
        if threads != 1:
            lib().orc_set_threads(h, int(threads))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    @property
    def num_colors(self):
        return lib().orc_num_colors(self._h)

    @property
    def max_incident(self):
        return lib().orc_max_incident(self._h)

    def colors(self):
        c = np.empty(self.n, np.int32)
        lib().orc_get_colors(self._h, _p(c, ctypes.c_int32))
        return c

    def set_threads(self, t):
        lib().orc_set_threads(self._h, int(t))

    def reset(self):
        lib().orc_reset(self._h)

    @property
    def old_positions(self):
        x = np.empty((self.n, 3))
        lib().orc_get_state(self._h, _p(x, ctypes.c_double), None)
        return x

    @property
    def old_velocities(self):
        v = np.empty((self.n, 3))
        lib().orc_get_state(self._h, None, _p(v, ctypes.c_double))
        return v

    def set_state(self, x=None, v=None):
        x = None if x is None else _f64(x).reshape(self.n, 3)
        v = None if v is None else _f64(v).reshape(self.n, 3)
        lib().orc_set_state(self._h, None if x is None else _p(x, ctypes.c_double),
                            None if v is None else _p(v, ctypes.c_double))

    def step(self, dt, dx_tol=None, max_iter=None, positions=None, guess=None):
        """One step; guess (optional [n,3]) replaces the inertia prediction as
        the starting positions of the active vertices."""
        tol = self.dx_tol if dx_tol is None else dx_tol
        mi = self.max_iter if max_iter is None else max_iter
        out = np.empty((self.n, 3))
        it, met = ctypes.c_int(0), ctypes.c_double(0.0)
        pos = None if positions is None else _f64(positions).reshape(self.n, 3)
        gs = None if guess is None else _f64(guess).reshape(self.n, 3)
        rc = lib().orc_step_guess(self._h, None if pos is None else _p(pos, ctypes.c_double),
                                  None if gs is None else _p(gs, ctypes.c_double), float(dt), float(tol),
                                  int(mi), _p(out, ctypes.c_double), ctypes.byref(it), ctypes.byref(met))
        if rc != 0:
            raise ValueError("oracle step: invalid dt or max_iter")
        self.last_iters, self.last_metric = it.value, met.value
        return out


# ----------------------------------------------------------------- literal restatements

def voxel2hex_loops(voxels, dx, dy, dz):
    """Loop-for-loop restatement of src/_utils.py:35-89."""
    nx, ny, nz = voxels.shape
    flag = np.full((nx + 1, ny + 1, nz + 1), -1, dtype=int)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if voxels[i][j][k]:
                    for ii in range(2):
                        for jj in range(2):
                            for kk in range(2):
                                flag[i + ii, j + jj, k + kk] = 0
    cnt = 0
    verts = []
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                if flag[i, j, k] == 0:
                    flag[i, j, k] = cnt
                    verts.append((dx * i, dy * j, dz * k))
                    cnt += 1
    hexes = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if voxels[i, j, k]:
                    hexes.append([flag[i, j, k], flag[i + 1, j, k], flag[i + 1, j + 1, k], flag[i, j + 1, k],
                                  flag[i, j, k + 1], flag[i + 1, j, k + 1], flag[i + 1, j + 1, k + 1],
                                  flag[i, j + 1, k + 1]])
    return np.stack(verts, axis=0), np.stack(hexes, axis=0)


def hex2tets_loops(hexes):
    """Loop-for-loop restatement of src/_utils.py:92-108."""
    n = len(hexes)
    t = np.zeros([5 * n, 4], dtype=int)
    for i in range(n):
        h = hexes[i]
        t[5 * i] = [h[0], h[1], h[3], h[4]]
        t[5 * i + 1] = [h[1], h[2], h[3], h[6]]
        t[5 * i + 2] = [h[1], h[4], h[5], h[6]]
        t[5 * i + 3] = [h[3], h[4], h[6], h[7]]
        t[5 * i + 4] = [h[1], h[3], h[4], h[6]]
    return t


def color_first_fit(elements):
    """Dict/set restatement of compute_adjacency_dict + graph_coloring
    (src/coloring.py:24-72).  Small meshes only."""
    elements = np.asarray(elements)
    nv = int(elements.max()) + 1
    adj = {i: set() for i in range(nv)}
    md = 0
    for ele in elements:
        for i, a in enumerate(ele):
            for j in range(i + 1, len(ele)):
                adj[int(a)].add(int(ele[j]))
                adj[int(ele[j])].add(int(a))
            md = max(md, len(adj[int(a)]))
    coloring = -np.ones(nv, dtype=int)
    groups = {}
    for v in range(nv):
        used = {coloring[u] for u in adj[v] if coloring[u] != -1}
        for c in range(256):
            if c not in used:
                coloring[v] = c
                groups.setdefault(c, []).append(v)
                break
    return coloring, groups, md
