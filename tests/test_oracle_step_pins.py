"""Pins of the oracle's time-step assembly that its per-slot checks leave open:
the lumped vertex mass m_e/4 in the inertia term, the damping term, and the
convergence quantity |dx_x + dx_y + dx_z| (src/vbd_solver.py:32,39-44,
222-227,578).  None of these reads the oracle's own formulas:

* One free vertex (the other seven vertices of a voxel pinned): one oracle
  iteration must be the Newton step -H^-1 G of the vertex's local incremental
  potential, with the mass taken from the geometry (the vertex lies in three
  corner tets of volume h^3/6 and the central tet of h^3/3, so a quarter of
  their mass is 5 rho h^3 / 24) and the elastic gradient and Hessian taken
  from the energy's definition by complex-step and central differences.
* A start whose Newton step has components that cancel: the solver stops
  after one iteration at a tolerance far below |dx|, and the metric it
  reports is |sum dx|, not a norm of dx.
* The README's tip curve (examples/run_cantilever.py:135-159; 19 crests and
  troughs digitised from asset/imgs/displacement_cantilever.png by
  tests/golden/digitize_readme_plot.py): the oscillation's period and
  amplitudes fix the balance of mass, gravity and stiffness, independently of
  any restatement.  A planted lumping error (m_e/3, as a density scaled by
  4/3) moves the first crest outside the gate.
"""
import json
import os

import numpy as np
import pytest

import oracle
from warpvbd_amd._utils import hex2tets, voxel2hex
from warpvbd_amd.scenes import cantilever

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _built(oracle_lib):
    return oracle_lib


def lame(E, nu):
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def psi(material, F, mu, lam):
    """Energy density from its definition; plain products only (no conjugate),
    so a complex step through it differentiates it."""
    if material == "snh":
        m, l = 4.0 * mu / 3.0, lam + 5.0 * mu / 6.0
        alpha = 1.0 + 0.75 * m / l
        Ic = np.sum(F * F)
        J = np.linalg.det(F)
        return 0.5 * m * (Ic - 3.0) - 0.5 * m * np.log(Ic + 1.0) + 0.5 * l * (J - alpha) ** 2
    E = 0.5 * (F.T @ F - np.eye(3))
    return mu * np.sum(E * E) + 0.5 * lam * np.trace(E) ** 2


class OneVertex:
    """A voxel of edge h (8 vertices, 5 tets by hex2tets); only vertex 1 (at
    x = h, y = z = 0) is free."""

    FREE = 1

    def __init__(self, material="snh", h=0.01, rho=1000.0, E=1e5, nu=0.4):
        self.X, H = voxel2hex(np.ones((1, 1, 1), bool), h, h, h)
        self.T = hex2tets(H)
        assert np.allclose(self.X[self.FREE], [h, 0.0, 0.0])
        self.material, self.h, self.rho, self.E, self.nu = material, h, rho, E, nu
        self.mu, self.lam = lame(E, nu)
        self.inc = [e for e, t in enumerate(self.T) if self.FREE in t]
        self.Dm = {e: np.stack([self.X[t[k]] - self.X[t[3]] for k in range(3)], axis=1)
                   for e, t in enumerate(self.T)}
        self.vol = {e: abs(np.linalg.det(D)) / 6.0 for e, D in self.Dm.items()}
        self.active = np.zeros(8, bool)
        self.active[self.FREE] = True

    def lumped_mass(self):
        return 5.0 * self.rho * self.h ** 3 / 24.0

    def energy(self, xv):
        tot = 0.0
        for e in self.inc:
            t = self.T[e]
            xs = [xv if k == self.FREE else self.X[k].astype(complex) for k in t]
            Ds = np.stack([xs[k] - xs[3] for k in range(3)], axis=1)
            tot = tot + self.vol[e] * psi(self.material, Ds @ np.linalg.inv(self.Dm[e]), self.mu, self.lam)
        return tot

    def grad(self, xv):
        step = 1e-30
        return np.array([self.energy(xv + 1j * step * np.eye(3)[k]).imag / step for k in range(3)])

    def hess(self, xv, d=1e-9):
        return np.stack([(self.grad(xv + d * np.eye(3)[k]) - self.grad(xv - d * np.eye(3)[k])) / (2 * d)
                         for k in range(3)], axis=1)

    def newton_step(self, x0, x_prev, v_prev, g, dt, damping):
        """The update that minimises the quadratic model of the vertex's local
        incremental potential: inertia M/(2 dt^2)|x - y|^2 (y = x_prev + dt
        v_prev), gravity -M g.x, elasticity, and the damping terms k_d/dt K
        (x - x_prev) in the gradient and k_d/dt K in the Hessian."""
        M = self.lumped_mass()
        K = self.hess(x0)
        y = x_prev + dt * v_prev
        G = M / dt ** 2 * (x0 - y) - M * g + self.grad(x0) + damping / dt * K @ (x0 - x_prev)
        H = M / dt ** 2 * np.eye(3) + (1.0 + damping / dt) * K
        return -np.linalg.solve(H, G)

    def oracle(self, damping=0.0, g=(0.0, 0.0, -9.81)):
        m = len(self.T)
        return oracle.OracleSolver(self.X, self.T, np.full(m, self.rho), np.full(m, self.E), np.full(m, self.nu),
                                   damping_coefficient=damping, active_mask=self.active, gravity=g,
                                   material=self.material, expanded=True)


def test_one_vertex_geometry_and_mass():
    ov = OneVertex()
    h = ov.h
    assert ov.inc == [0, 1, 2, 4]
    assert np.allclose(sorted(ov.vol[e] for e in ov.inc), [h ** 3 / 6] * 3 + [h ** 3 / 3], rtol=1e-12)
    assert np.isclose(ov.lumped_mass(), ov.rho * sum(ov.vol[e] for e in ov.inc) / 4.0, rtol=1e-12)
    # the differentiated energy: zero gradient and the linear-elastic stiffness at rest
    assert np.abs(ov.grad(ov.X[1].astype(float))).max() < 1e-9
    K = ov.hess(ov.X[1].astype(float))
    assert np.allclose(K, K.T, atol=1e-6) and np.all(np.linalg.eigvalsh(K) > 0)


@pytest.mark.parametrize("material", ["snh", "stvk"])
@pytest.mark.parametrize("damping", [0.0, 1.0])
def test_one_iteration_is_the_newton_step_of_the_local_potential(material, damping):
    ov = OneVertex(material)
    dt = 1e-3
    g = np.array([0.3, -0.5, -9.81])
    x_prev = ov.X.copy()
    x_prev[1] += [1e-4, -2e-4, 3e-4]
    v_prev = np.zeros((8, 3))
    v_prev[1] = [0.05, 0.02, -0.1]
    x0 = x_prev[1] + dt * v_prev[1] + np.array([2e-5, -1e-5, 1.5e-5])
    guess = x_prev.copy()
    guess[1] = x0
    r = ov.oracle(damping, g)
    r.set_state(x=x_prev, v=v_prev)
    x1 = r.step(dt, dx_tol=0.0, max_iter=1, guess=guess)
    dx = ov.newton_step(x0, x_prev[1], v_prev[1], g, dt, damping)
    err = np.abs((x1[1] - x0) - dx).max()
    print(f"{material} damping={damping}: |dx| = {np.abs(dx).max():.3e} m, oracle - closed form {err:.2e} m")
    assert err < 1e-8 * np.abs(dx).max()
    pinned = np.arange(8) != 1
    assert np.array_equal(x1[pinned], x_prev[pinned])
    # the same step with a quarter of the tet mass replaced by a third is far off
    ov3 = OneVertex(material, rho=ov.rho * 4.0 / 3.0)
    assert np.abs(ov3.newton_step(x0, x_prev[1], v_prev[1], g, dt, damping) - dx).max() > 1e3 * err


def _cancelling_start(ov, dt, eps=1e-4):
    """A start x0 = X + eps (cos s, sin s, 0) whose Newton step has
    dx_x + dx_y + dx_z = 0 (bisection on s), and a generic one (s = 0.3)."""
    zero = np.zeros(3)

    def step(s):
        x0 = ov.X[1] + eps * np.array([np.cos(s), np.sin(s), 0.0])
        return x0, ov.newton_step(x0, ov.X[1].astype(float), zero, zero, dt, 0.0)

    f = lambda s: step(s)[1].sum()
    lo, hi = 0.0, np.pi
    assert f(lo) * f(hi) < 0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return step(0.5 * (lo + hi)), step(0.3)


def test_convergence_metric_is_the_absolute_component_sum():
    """src/vbd_solver.py:578 stops on max_v |dx_x + dx_y + dx_z| < tol: a
    start whose update components cancel converges after one iteration at a
    tolerance a thousand times below the update's size, and the reported
    metric is |sum dx| (not max |dx_i| or the norm)."""
    ov = OneVertex()
    dt = 1e-3
    (xc, dxc), (xg, dxg) = _cancelling_start(ov, dt)
    assert abs(dxc.sum()) < 1e-9 * np.abs(dxc).max()
    tol = 1e-3 * np.abs(dxc).max()
    r = ov.oracle(0.0, g=(0.0, 0.0, 0.0))
    for x0, dx in ((xc, dxc), (xg, dxg)):
        guess = ov.X.copy()
        guess[1] = x0
        r.reset()
        r.step(dt, dx_tol=0.0, max_iter=1, guess=guess)
        assert abs(r.last_metric - abs(dx.sum())) < 1e-8 * np.abs(dx).max()
    assert abs(dxg.sum()) > 0.1 * np.linalg.norm(dxg)
    guess = ov.X.copy()
    guess[1] = xc
    r.reset()
    r.step(dt, dx_tol=tol, max_iter=50, guess=guess)
    assert r.last_iters == 1 and r.last_metric < tol < np.linalg.norm(dxc)
    guess[1] = xg
    r.reset()
    r.step(dt, dx_tol=tol, max_iter=50, guess=guess)
    assert r.last_iters >= 2


def _extrema(z, dt):
    """Local extrema of a sampled curve, with the plot's time axis (index x dt)."""
    out = []
    for i in range(1, len(z) - 1):
        if z[i] > z[i - 1] and z[i] >= z[i + 1]:
            out.append(("crest", i * dt, z[i]))
        elif z[i] < z[i - 1] and z[i] <= z[i + 1]:
            out.append(("trough", i * dt, z[i]))
    return out


def readme_anchor_errors(z, dt, anchors):
    """Per digitised extremum: (|t - t_sim|, |z - z_sim|) against the nearest
    simulated extremum of the same kind (inf when there is none)."""
    sim = _extrema(np.asarray(z), dt)
    errs = []
    for a in anchors:
        cand = [(abs(t - a["t"]), abs(v - a["z"])) for k, t, v in sim if k == a["kind"]]
        errs.append(min(cand) if cand else (np.inf, np.inf))
    return np.array(errs)


# Gates: 5 ms (2.3 pixels, 3% of the 0.152 s period) and 0.3 mm (8 pixels,
# 2% of the first crest) per extremum.
README_T_GATE, README_Z_GATE = 5e-3, 3e-4


def readme_scene(density_scale=1.0):
    with open(os.path.join(GOLDEN, "readme_displacement_cantilever_extrema.json")) as f:
        doc = json.load(f)
    s = doc["scene"]
    sc = cantilever(tuple(s["voxels"]), s["h"], s["youngs_modulus"], s["poissons_ratio"],
                    s["density"] * density_scale)
    return sc, s, doc["extrema"]


def test_readme_tip_curve_anchors():
    sc, s, anchors = readme_scene()
    r = oracle.OracleSolver(**sc.solver_kwargs(), dx_tol=s["dx_tol"], max_iter=s["max_iter"], expanded=False)
    z = [r.step(s["dt"])[sc.tip_idx, 2].mean() for _ in range(s["steps"])]
    errs = readme_anchor_errors(z, s["dt"], anchors)
    print(f"README anchors: max |dt| {errs[:, 0].max() * 1e3:.2f} ms, max |dz| {errs[:, 1].max():.2e} m "
          f"over {len(anchors)} extrema")
    assert len(anchors) == 19
    assert errs[:, 0].max() <= README_T_GATE and errs[:, 1].max() <= README_Z_GATE


def test_readme_anchors_catch_a_lumping_error():
    """The same scene with each tet's mass split in thirds instead of quarters
    (a density scaled by 4/3 gives the same inertia and gravity terms)."""
    sc, s, anchors = readme_scene(4.0 / 3.0)
    r = oracle.OracleSolver(**sc.solver_kwargs(), dx_tol=s["dx_tol"], max_iter=s["max_iter"], expanded=False)
    z = [r.step(s["dt"])[sc.tip_idx, 2].mean() for _ in range(240)]
    errs = readme_anchor_errors(z, s["dt"], anchors[:2])
    assert errs[1, 0] > 3 * README_T_GATE
