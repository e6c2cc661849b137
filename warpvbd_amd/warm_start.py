"""Learned warm start for the VBD time step (no counterpart in WarpVBD).

A 4-layer message-passing network on the mesh's vertex graph (width 64, BF16
weights on MFMA with FP32 accumulation, hand-written forward and training
kernels in csrc/vbd_mpnn.hip) predicts each step's converged correction
x_{n+1} - x_inertia from the state at the start of the step.  The guess is
x_inertia + h + dt^2 |g| out.  The history part h is c_prev, the previous
step's converged correction, or its linear extrapolation 2 c_prev - c_prev2
once the two steps before were consecutive (each scaled by (dt / dt of its
step)^2; h = 0 after a reset).  It carries the part of the correction that
evolves smoothly from step to step, and the network predicts how the
correction departs from it.  The solver then iterates from that guess to the same
tolerance, so its accuracy is still decided by VBD; a body whose warm guess
has a larger first-sweep residual than the plain inertia guess starts from
the plain guess instead.

Typical use::

    ws = WarmStart(solver, capacity=2000)
    ws.generate(trajectories=..., steps=..., dt=1e-3, tol=1e-7)   # solver-made samples
    ws.train(epochs=20)
    solver.step(solver.old_positions, 1e-3, warm_start=ws)
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, ptr

F_IN, WIDTH, LAYERS, F_OUT = 32, 64, 4, 3


def param_layout():
    """[(name, shape)] of the flat FP32 parameter vector."""
    out = [("W_in", (WIDTH, F_IN)), ("b_in", (WIDTH,))]
    for l in range(LAYERS):
        out += [(f"W_{l}", (WIDTH, 2 * WIDTH)), (f"b_{l}", (WIDTH,))]
    out += [("W_out", (F_OUT, WIDTH)), ("b_out", (F_OUT,))]
    return out


def flatten(params: dict) -> np.ndarray:
    return np.concatenate([np.asarray(params[k], np.float32).reshape(-1) for k, _ in param_layout()])


def unflatten(vec: np.ndarray) -> dict:
    out, o = {}, 0
    for k, shp in param_layout():
        n = int(np.prod(shp))
        out[k] = np.asarray(vec[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def init_params(seed: int = 0) -> dict:
    """He-style initialisation with a small output layer."""
    rng = np.random.default_rng(seed)
    p = {"W_in": rng.normal(0, math.sqrt(2.0 / F_IN), (WIDTH, F_IN)), "b_in": np.zeros(WIDTH)}
    for l in range(LAYERS):
        p[f"W_{l}"] = rng.normal(0, math.sqrt(2.0 / (2 * WIDTH)) * 0.5, (WIDTH, 2 * WIDTH))
        p[f"b_{l}"] = np.zeros(WIDTH)
    p["W_out"] = rng.normal(0, math.sqrt(1.0 / WIDTH) * 0.1, (F_OUT, WIDTH))
    p["b_out"] = np.zeros(F_OUT)
    return p


class WarmStart:
    """The network bound to one ``VBDSolver`` or ``BatchedVBDSolver`` (its mesh
    and graph).  Parameters can be moved between solvers with
    get_params/set_params: the network itself is mesh independent."""

    def __init__(self, solver, capacity: int = 256, seed: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.vbd_mpnn_create(solver._h, int(capacity), ctypes.byref(h)), "vbd_mpnn_create")
        self._h = h
        self.solver = solver
        self.capacity = int(capacity)
        self.n_samples = 0
        self.n_params = L.vbd_mpnn_num_params()
        self.policy = self.CHOOSE
        self.set_params(flatten(init_params(seed)))
        self.refresh_materials()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().vbd_mpnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ parameters
    def set_params(self, vec):
        v = np.ascontiguousarray(vec, dtype=np.float32)
        assert v.size == self.n_params
        check(_lib.load().vbd_mpnn_set_params(self._h, ptr(v, ctypes.c_float)), "vbd_mpnn_set_params")

    def get_params(self) -> np.ndarray:
        v = np.empty(self.n_params, np.float32)
        check(_lib.load().vbd_mpnn_get_params(self._h, ptr(v, ctypes.c_float)), "vbd_mpnn_get_params")
        return v

    def set_material(self, youngs_modulus: float, poissons_ratio: float):
        check(_lib.load().vbd_mpnn_set_material(self._h, float(youngs_modulus), float(poissons_ratio)),
              "vbd_mpnn_set_material")

    def refresh_materials(self):
        """Material features from the solver's element arrays (first element of
        each body)."""
        s = self.solver
        E = np.asarray(s.youngs_modulus, np.float64).reshape(-1)
        nu = np.asarray(s.poissons_ratio, np.float64).reshape(-1)
        if len(E) == 0:
            return
        offs = getattr(s, "body_vertex_offsets", None)
        if offs is None:
            self.set_material(float(E[0]), float(nu[0]))
            return
        T = np.asarray(s.elements)
        offs = np.asarray(offs)
        nb = len(offs) - 1
        body = np.searchsorted(offs, T[:, 0], side="right") - 1
        bodies, first = np.unique(body, return_index=True)
        Eb = np.full(nb, E[0])
        nub = np.full(nb, nu[0])
        Eb[bodies] = E[first]
        nub[bodies] = nu[first]
        Eb = np.ascontiguousarray(Eb)
        nub = np.ascontiguousarray(nub)
        check(_lib.load().vbd_mpnn_set_body_materials(self._h, ptr(Eb, ctypes.c_double), ptr(nub, ctypes.c_double)),
              "vbd_mpnn_set_body_materials")

    # initial-guess policy of warm steps
    CHOOSE, NETWORK, PLAIN, ADAPTIVE, HISTORY = 0, 1, 2, 3, 4

    def set_policy(self, policy: int):
        """CHOOSE (default): per body the guess with the smallest first-sweep
        residual among the plain inertia prediction, the network guess and the
        history guess x_inertia + h; NETWORK: always the network's;
        PLAIN: always the inertia prediction (the network is not evaluated);
        ADAPTIVE: CHOOSE, but after a step where fewer than 2% of the bodies
        took the network guess the network is skipped for 1, 2, 4, ... (at
        most 16) steps, which choose between the plain and history guesses
        only; HISTORY: that choice on every step, without the network.  On
        steps with a history guess the plain guess is compared only on check
        steps: after a check where no body took it, the next 1, 2, 4, ... (at
        most 16) such steps leave it out; setting the state, material or
        gravity makes the next step a check step again."""
        check(_lib.load().vbd_mpnn_set_policy(self._h, int(policy)), "vbd_mpnn_set_policy")
        self.policy = int(policy)

    def last_guess_counts(self) -> dict:
        """Bodies that started the last warm step from each guess."""
        c = (ctypes.c_int32 * 3)()
        check(_lib.load().vbd_mpnn_guess_counts(self._h, c), "vbd_mpnn_guess_counts")
        return {"plain": int(c[0]), "network": int(c[1]), "history": int(c[2])}

    def save(self, path):
        np.save(path, self.get_params())

    def load(self, path):
        self.set_params(np.load(path))

    # ------------------------------------------------------------------ data
    def plain_step(self, dt, tol, max_iter=1000) -> int:
        """A step from the plain guess that still passes through the warm
        start, so that the next sample's features see its correction as the
        step right before (a solver step without the warm start breaks that
        chain)."""
        keep = self.policy
        self.set_policy(self.PLAIN)
        try:
            self.solver.step(self.solver.old_positions, dt, dx_tol=tol, max_iter=max_iter, warm_start=self)
        finally:
            self.set_policy(keep)
        return self.solver.last_iterations

    def collect(self, dt, tol, max_iter=1000) -> int:
        """One plain solver step that also stores a training sample."""
        if self.n_samples >= self.capacity:
            raise ValueError("sample store is full")
        it = ctypes.c_int32()
        self.solver._before_mutation()
        check(_lib.load().vbd_mpnn_collect(self._h, self.n_samples, float(dt), float(tol), int(max_iter),
                                           ctypes.byref(it)), "vbd_mpnn_collect")
        self.n_samples += 1
        return it.value

    def generate(self, trajectories: int, steps: int, dt: float = 1e-3, tol: float = 1e-7, max_iter: int = 1000,
                 seed: int = 0, sample_every: int = 1, E_range=(1e4, 1e6), nu_range=(0.3, 0.45),
                 cone_deg: float = 30.0, v0_std: float = 0.05, g_norm: float = 9.81):
        """Run seeded solver trajectories and store samples from them.  Per
        trajectory: E log-uniform in E_range, nu uniform in nu_range, gravity
        direction uniform in a cone of cone_deg about -z, initial velocity
        N(0, v0_std) per component of every active vertex."""
        s = self.solver
        rng = np.random.default_rng(seed)
        iters = []
        for t in range(trajectories):
            E = float(10 ** rng.uniform(np.log10(E_range[0]), np.log10(E_range[1])))
            nu = float(rng.uniform(*nu_range))
            g = gravity_in_cone(rng, cone_deg, g_norm)
            v0 = rng.normal(0.0, v0_std, size=(s.n_vertices, 3)) * s.active_mask[:, None]
            s.reset()
            s.set_material(E, nu)
            s.set_gravity(g)
            s.old_velocities = v0
            self.set_material(E, nu)
            for k in range(steps):
                if k % sample_every == 0 and self.n_samples < self.capacity:
                    iters.append(self.collect(dt, tol, max_iter))
                else:
                    iters.append(self.plain_step(dt, tol, max_iter))
        return np.array(iters)

    def sample(self, slot):
        n = self.solver.n_vertices
        F = np.empty((n, F_IN), np.float32)
        T = np.empty((n, 3), np.float32)
        ne = ctypes.c_int32()
        off = np.empty(n + 1, np.int32)
        check(_lib.load().vbd_mpnn_read_sample(self._h, int(slot), None, None, None, None, ctypes.byref(ne)),
              "vbd_mpnn_read_sample")
        nb = np.empty(max(1, ne.value), np.int32)
        check(_lib.load().vbd_mpnn_read_sample(self._h, int(slot), ptr(F, ctypes.c_float), ptr(T, ctypes.c_float),
                                               ptr(off, ctypes.c_int32), ptr(nb, ctypes.c_int32), ctypes.byref(ne)),
              "vbd_mpnn_read_sample")
        return F, T, off, nb[: ne.value]

    # ------------------------------------------------------------------ training
    def loss_grad(self, slot):
        loss = ctypes.c_float()
        g = np.empty(self.n_params, np.float32)
        check(_lib.load().vbd_mpnn_loss_grad(self._h, int(slot), ctypes.byref(loss), ptr(g, ctypes.c_float)),
              "vbd_mpnn_loss_grad")
        return float(loss.value), g

    def train(self, epochs: int = 10, batch: int = 8, lr: float = 1e-3, seed: int = 0, n_samples: int = None):
        n = self.n_samples if n_samples is None else n_samples
        hist = np.empty(epochs, np.float32)
        check(_lib.load().vbd_mpnn_train(self._h, int(n), int(epochs), int(batch), float(lr), int(seed),
                                         ptr(hist, ctypes.c_float)), "vbd_mpnn_train")
        return hist

    def predict(self, dt, slot: int = -1) -> np.ndarray:
        """The network's output [n, 3] in units of dt^2 |g| (the warm guess
        adds the history part h) for stored sample `slot`, or for the solver's
        state."""
        out = np.empty((self.solver.n_vertices, 3), np.float32)
        check(_lib.load().vbd_mpnn_predict(self._h, int(slot), float(dt), ptr(out, ctypes.c_float)),
              "vbd_mpnn_predict")
        return out


def pretrained(name: str = "cantilever128x16x16") -> str:
    """Path of committed parameters: ``cantilever128x16x16`` was trained by
    ``python -m warpvbd_amd.warm_train`` on 200 trajectories x 300 steps of the
    128x16x16 cantilever (seeds 0..199); load with ``WarmStart.load``."""
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", f"warm_start_{name}.npy")
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    return path


def gravity_in_cone(rng, cone_deg, g_norm=9.81):
    """Gravity of magnitude g_norm with direction uniform (in solid angle) in a
    cone of half-angle cone_deg about -z."""
    cmin = math.cos(math.radians(cone_deg))
    c = rng.uniform(cmin, 1.0)
    phi = rng.uniform(0.0, 2.0 * math.pi)
    s = math.sqrt(max(0.0, 1.0 - c * c))
    return (g_norm * s * math.cos(phi), g_norm * s * math.sin(phi), -g_norm * c)
