"""GPU tests of the learned warm start: the MFMA forward and the hand-written
backward against the FP64 reference network (oracle/mpnn_ref.py), training,
the first-sweep-residual fallback, and the gain over the plain initial guess."""
import numpy as np
import pytest

from oracle import mpnn_ref as R
from warpvbd_amd import VBDSolver
from warpvbd_amd.scenes import cantilever
from warpvbd_amd.warm_start import WarmStart, flatten, init_params, unflatten

pytestmark = pytest.mark.gpu


def make_solver(nx=(12, 3, 3), **kw):
    sc = cantilever(nx)
    s = VBDSolver(**sc.solver_kwargs(), dx_tol=1e-7, max_iter=1000, verbose=False, **kw)
    return sc, s


def test_forward_and_gradients_match_reference():
    sc, s = make_solver()
    ws = WarmStart(s, capacity=4, seed=1)
    rng = np.random.default_rng(0)
    p = init_params(1)
    for k in p:
        if k.startswith("b_"):
            p[k] = rng.normal(0, 0.05, p[k].shape)
    p["W_out"] *= 10.0
    ws.set_params(flatten(p))
    for _ in range(3):
        ws.collect(1e-3, 1e-7)
    for slot in range(3):
        F, T, off, nb = ws.sample(slot)
        pr = unflatten(flatten(p))  # the FP32-rounded parameters the GPU holds
        loss_ref, g_ref, out_ref = R.loss_and_grads(pr, F.astype(np.float64), off, nb, T.astype(np.float64),
                                                    sc.active_mask, rounding=True)
        loss, g = ws.loss_grad(slot)
        out = ws.predict(1e-3, slot)
        scale = np.abs(out_ref).max()
        assert np.abs(out - out_ref).max() <= 2e-3 * scale + 1e-6, slot
        assert abs(loss - loss_ref) <= 1e-3 * abs(loss_ref) + 1e-9, (loss, loss_ref)
        gg = unflatten(g)
        for name in g_ref:
            a, b = gg[name].reshape(-1), g_ref[name].reshape(-1)
            nb_ = np.linalg.norm(b)
            if nb_ == 0:
                assert np.abs(a).max() < 1e-6
                continue
            rel = np.linalg.norm(a - b) / nb_
            cos = float(a @ b) / (np.linalg.norm(a) * nb_)
            assert rel < 2e-2 and cos > 0.9995, (name, rel, cos)


def test_training_reduces_loss():
    sc, s = make_solver()
    ws = WarmStart(s, capacity=40, seed=2)
    ws.generate(trajectories=4, steps=10, seed=3)
    assert ws.n_samples == 40
    l0 = np.mean([ws.loss_grad(i)[0] for i in range(ws.n_samples)])
    hist = ws.train(epochs=40, batch=4, lr=2e-3, seed=0)
    l1 = np.mean([ws.loss_grad(i)[0] for i in range(ws.n_samples)])
    print(f"loss {l0:.4g} -> {l1:.4g}; epochs {hist[0]:.4g} .. {hist[-1]:.4g}")
    assert l1 < 0.5 * l0 and hist[-1] < hist[0]


def test_fallback_keeps_plain_guess_bit_exact():
    """A network whose guess is far off must lose the first-sweep comparison,
    and the step must then equal the plain step exactly."""
    sc, s = make_solver()
    sc2, s2 = make_solver()
    for sol in (s, s2):
        sol.set_launch(8, 256, 0, 0)
    ws = WarmStart(s, capacity=1, seed=4)
    p = init_params(4)
    p["b_out"] = np.array([50.0, -80.0, 120.0])  # a guess 100s of dt^2 g away
    ws.set_params(flatten(p))
    for k in range(10):
        x = s.step(s.old_positions, 1e-3, warm_start=ws).numpy()
        assert not s.last_used_network
        x2 = s2.step(s2.old_positions, 1e-3).numpy()
        assert np.array_equal(x, x2) and s.last_iterations == s2.last_iterations


def test_trained_network_gain():
    """Train on solver-made trajectories of a 32x4x4 cantilever, then run
    held-out seeds with and without the warm start to tol 1e-7: fewer
    iterations, same converged result to within the tolerance's reach."""
    nx = (32, 4, 4)
    sc, s = make_solver(nx)
    ws = WarmStart(s, capacity=600, seed=5)
    ws.generate(trajectories=12, steps=50, seed=10)
    hist = ws.train(epochs=30, batch=8, lr=2e-3, seed=1)
    print("train loss", hist[0], "->", hist[-1])
    rng = np.random.default_rng(1000)
    from warpvbd_amd.warm_start import gravity_in_cone
    it_plain, it_warm, used, err = [], [], [], []
    sc_b, sb = make_solver(nx)
    for seed in range(3):
        E = float(10 ** rng.uniform(4, 6))
        nu = float(rng.uniform(0.3, 0.45))
        g = gravity_in_cone(rng, 30.0)
        v0 = rng.normal(0, 0.05, size=(sc.n_vertices, 3)) * sc.active_mask[:, None]
        for sol in (s, sb):
            sol.reset()
            sol.set_material(E, nu)
            sol.set_gravity(g)
            sol.old_velocities = v0
        ws.set_material(E, nu)
        for k in range(40):
            xw = s.step(s.old_positions, 1e-3, warm_start=ws).numpy()
            it_warm.append(s.last_iterations)
            used.append(s.last_used_network)
            xp = sb.step(sb.old_positions, 1e-3).numpy()
            it_plain.append(sb.last_iterations)
            err.append(np.abs(xw - xp).max())
    it_plain, it_warm = np.array(it_plain), np.array(it_warm)
    print(f"iterations plain {it_plain.mean():.2f} warm {it_warm.mean():.2f}; network used {np.mean(used):.2f}; "
          f"max |x_warm - x_plain| {max(err):.3e}")
    assert it_warm.mean() < it_plain.mean()
    diag = float(np.linalg.norm(sc.vertices.max(0) - sc.vertices.min(0)))
    assert max(err) < 1e-4 * diag
