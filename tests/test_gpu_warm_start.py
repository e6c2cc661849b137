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
    """A network whose guess is far off must lose the first-sweep comparison:
    the first step (no previous correction yet) equals the plain step exactly,
    and every step equals the step of the HISTORY policy (the same choice
    between the plain and history guesses, without the network)."""
    sc, s = make_solver()
    sc2, s2 = make_solver()
    sc3, s3 = make_solver()
    for sol in (s, s2, s3):
        sol.set_launch(8, 256, 0, 0)
    ws = WarmStart(s, capacity=1, seed=4)
    p = init_params(4)
    p["b_out"] = np.array([50.0, -80.0, 120.0])  # a guess 100s of dt^2 g away
    ws.set_params(flatten(p))
    wh = WarmStart(s3, capacity=1)
    wh.set_policy(wh.HISTORY)
    for k in range(10):
        x = s.step(s.old_positions, 1e-3, warm_start=ws).numpy()
        assert not s.last_used_network
        x3 = s3.step(s3.old_positions, 1e-3, warm_start=wh).numpy()
        assert np.array_equal(x, x3) and s.last_iterations == s3.last_iterations
        if k == 0:
            x2 = s2.step(s2.old_positions, 1e-3).numpy()
            assert np.array_equal(x, x2) and s.last_iterations == s2.last_iterations


def resync_errors(ref, solvers, steps, dt=1e-3, warm=None):
    """Per-step resync against a converged oracle trajectory: from each oracle
    state, one step of every solver (warm: the warm start of the first solver,
    or a list with one entry, possibly None, per solver); returns per-solver
    arrays of iterations and max |x - x_oracle| per step."""
    warms = list(warm) if isinstance(warm, (list, tuple)) else [warm] + [None] * (len(solvers) - 1)
    its = [[] for _ in solvers]
    errs = [[] for _ in solvers]
    for k in range(steps):
        x0, v0 = ref.old_positions, ref.old_velocities
        xr = ref.step(dt)
        for i, sol in enumerate(solvers):
            sol.old_positions = x0
            sol.old_velocities = v0
            x = sol.step(sol.old_positions, dt, warm_start=warms[i]).numpy()
            its[i].append(sol.last_iterations)
            errs[i].append(np.abs(x - xr).max())
    return [np.array(a) for a in its], [np.array(e) for e in errs]


def test_plain_policy_equals_plain_step():
    """PLAIN never evaluates the network: the warm step equals the plain step
    bit for bit (same launch shapes on both solvers: a warm step does not use
    the one-workgroup whole-step kernel)."""
    sc, s = make_solver()
    sc2, s2 = make_solver()
    for sol in (s, s2):
        sol.set_launch(8, 256, 0, 0)
    ws = WarmStart(s, capacity=1, seed=4)
    p = init_params(4)
    p["b_out"] = np.array([50.0, -80.0, 120.0])
    ws.set_params(flatten(p))
    ws.set_policy(ws.PLAIN)
    for k in range(12):
        x = s.step(s.old_positions, 1e-3, warm_start=ws).numpy()
        x2 = s2.step(s2.old_positions, 1e-3).numpy()
        assert not s.last_used_network and ws.last_guess_counts()["plain"] == 1
        assert np.array_equal(x, x2) and s.last_iterations == s2.last_iterations


def test_adaptive_policy_with_a_losing_network_equals_history():
    """ADAPTIVE with a network that always loses the first-sweep comparison:
    no body takes the network guess, the network is skipped with growing
    back-off, and every step equals the HISTORY policy's step bit for bit
    (both choose between the same plain and history guesses)."""
    sc, s = make_solver()
    sc2, s2 = make_solver()
    for sol in (s, s2):
        sol.set_launch(8, 256, 0, 0)
    p = init_params(4)
    p["b_out"] = np.array([50.0, -80.0, 120.0])
    wa = WarmStart(s, capacity=1, seed=4)
    wa.set_params(flatten(p))
    wa.set_policy(wa.ADAPTIVE)
    wh = WarmStart(s2, capacity=1, seed=4)
    wh.set_policy(wh.HISTORY)
    hist_taken = 0
    for k in range(12):
        x = s.step(s.old_positions, 1e-3, warm_start=wa).numpy()
        x2 = s2.step(s2.old_positions, 1e-3, warm_start=wh).numpy()
        ca, ch = wa.last_guess_counts(), wh.last_guess_counts()
        assert not s.last_used_network and ca == ch
        hist_taken += ca["history"]
        assert np.array_equal(x, x2) and s.last_iterations == s2.last_iterations
    assert hist_taken > 0


def test_network_policy_forces_the_guess():
    """NETWORK takes the network guess even when it is far off; the step still
    converges to the tolerance from it."""
    sc, s = make_solver()
    ws = WarmStart(s, capacity=1, seed=4)
    p = init_params(4)
    p["b_out"] = np.array([5.0, -8.0, 12.0])
    ws.set_params(flatten(p))
    ws.set_policy(ws.NETWORK)
    x = s.step(s.old_positions, 1e-3, warm_start=ws).numpy()
    assert s.last_used_network and s.last_metric < 1e-7 and np.all(np.isfinite(x))


def test_trained_network_gain():
    """Train on solver-made trajectories of a 32x4x4 cantilever, then step
    held-out seeds with and without the warm start to tol 1e-7.  Each step
    starts from the same state of a converged FP64 oracle trajectory (tol
    1e-11), so the test measures one step's error: the warm-started step takes
    fewer iterations and its error against the converged step stays that of
    the plain step (within 1.5x in mean and max: both stop once the update
    falls below the tolerance, and from the network's guess the error left at
    that point can be somewhat larger or smaller than from the plain guess)."""
    import oracle
    from warpvbd_amd.warm_start import gravity_in_cone
    nx = (32, 4, 4)
    sc, s = make_solver(nx)
    ws = WarmStart(s, capacity=600, seed=5)
    ws.generate(trajectories=12, steps=50, seed=10)
    hist = ws.train(epochs=30, batch=8, lr=2e-3, seed=1)
    print("train loss", hist[0], "->", hist[-1])
    rng = np.random.default_rng(1000)
    sc_b, sb = make_solver(nx)
    sc_h, sh = make_solver(nx)
    wh = WarmStart(sh, capacity=1)
    wh.set_policy(wh.HISTORY)
    it_w, it_p, it_h, e_w, e_p = [], [], [], [], []
    for seed in range(3):
        E = float(10 ** rng.uniform(4, 6))
        nu = float(rng.uniform(0.3, 0.45))
        g = gravity_in_cone(rng, 30.0)
        v0 = rng.normal(0, 0.05, size=(sc.n_vertices, 3)) * sc.active_mask[:, None]
        for sol in (s, sb, sh):
            sol.set_material(E, nu)
            sol.set_gravity(g)
        ws.set_material(E, nu)
        ref = oracle.OracleSolver(**dict(sc.solver_kwargs(), youngs_modulus=E, poissons_ratio=nu, gravity=g),
                                  dx_tol=1e-11, max_iter=100000, expanded=False, threads=8)
        ref.set_state(v=v0)
        (iw, ip, ih), (ew, ep, _) = resync_errors(ref, [s, sb, sh], 25, warm=[ws, None, wh])
        it_w.append(iw); it_p.append(ip); it_h.append(ih); e_w.append(ew); e_p.append(ep)
        print(f"E {E:.3g} nu {nu:.3f}: iterations warm {iw.mean():.2f} plain {ip.mean():.2f} "
              f"history only {ih.mean():.2f}; step error warm {ew.mean():.3e} (max {ew.max():.3e}) "
              f"plain {ep.mean():.3e} (max {ep.max():.3e})")
    it_w, it_p, it_h, e_w, e_p = (np.concatenate(a) for a in (it_w, it_p, it_h, e_w, e_p))
    print(f"all seeds: iterations warm {it_w.mean():.2f} plain {it_p.mean():.2f} history only {it_h.mean():.2f}")
    assert it_w.mean() < it_p.mean()
    assert e_w.mean() <= 1.5 * e_p.mean() and e_w.max() <= 1.5 * e_p.max()


def test_pretrained_network_beats_history_in_iterations():
    """The committed full-budget parameters on held-out seeds 1000..1002 of
    the 128x16x16 cantilever (100 steps each, same start): CHOOSE (plain,
    history or network guess, whichever has the smallest first-sweep residual)
    needs fewer iterations than the history guess alone, and takes the network
    guess on some steps.  (In wall time the history guess wins on this mesh:
    the network costs more per step than the iterations it can save.)"""
    from warpvbd_amd.warm_start import pretrained
    from warpvbd_amd.warm_train import scene_of_seed
    sc = cantilever((128, 16, 16))
    params = np.load(pretrained("cantilever128x16x16"))
    sols = {}
    for name, pol in (("history", WarmStart.HISTORY), ("choose", WarmStart.CHOOSE)):
        s = VBDSolver(**sc.solver_kwargs(), dx_tol=1e-7, max_iter=1000, verbose=False)
        ws = WarmStart(s, capacity=1)
        ws.set_params(params)
        ws.set_policy(pol)
        sols[name] = (s, ws)
    its = {k: [] for k in sols}
    took_net = 0
    for seed in (1000, 1001, 1002):
        E, nu, g, v0 = scene_of_seed(sc.n_vertices, sc.active_mask, seed)
        for s, ws in sols.values():
            s.reset()
            s.set_material(E, nu)
            s.set_gravity(g)
            s.old_velocities = v0
            ws.set_material(E, nu)
        for _ in range(100):
            for k, (s, ws) in sols.items():
                s.step(s.old_positions, 1e-3, warm_start=ws)
                its[k].append(s.last_iterations)
            took_net += sols["choose"][1].last_guess_counts()["network"]
        print(f"seed {seed}: iterations/step history {np.mean(its['history'][-100:]):.2f} "
              f"choose {np.mean(its['choose'][-100:]):.2f}")
    print(f"network guess taken on {took_net} of 300 steps")
    assert np.mean(its["choose"]) < np.mean(its["history"])
    assert took_net > 0


def test_history_guess_matches_oracle():
    """With the network's output layer zeroed the warm guess is x_inertia plus
    the history part: c_prev, the previous step's converged correction, after
    one step, and the linear extrapolation 2 c_prev - c_prev2 once the two
    steps before were consecutive.  Forced (NETWORK policy) at a fixed
    iteration count, each step from the FP64 oracle's state matches the oracle
    started from that same guess to 1e-6 of the bounding-box diagonal; the
    first step after a reset has no c_prev and matches the oracle's plain
    step."""
    import oracle
    sc, s = make_solver((16, 3, 3))
    ws = WarmStart(s, capacity=1, seed=3)
    p = init_params(3)
    p["W_out"][:] = 0.0
    p["b_out"][:] = 0.0
    ws.set_params(flatten(p))
    ws.set_policy(ws.NETWORK)
    ref = oracle.OracleSolver(**sc.solver_kwargs(), dx_tol=0.0, max_iter=10, expanded=False)
    dt, K = 1e-3, 6
    act = sc.active_mask[:, None].astype(np.float64)
    g = np.asarray(sc.gravity, np.float64)
    diag = float(np.linalg.norm(sc.vertices.max(0) - sc.vertices.min(0)))
    c = np.zeros((sc.n_vertices, 3))
    c2 = None
    for k in range(8):
        x0, v0 = ref.old_positions, ref.old_velocities
        y = x0 + (dt * v0 + dt * dt * g) * act
        xr = ref.step(dt, dx_tol=0.0, max_iter=K, guess=y + (c if c2 is None else 2.0 * c - c2))
        s.old_positions = x0
        s.old_velocities = v0
        x = s.step(s.old_positions, dt, dx_tol=0.0, max_iter=K, warm_start=ws).numpy()
        err = np.abs(x - xr).max()
        print(f"step {k}: |c_prev| max {np.abs(c).max():.3e}, error {err:.3e} (bar {1e-6 * diag:.3e})")
        assert s.last_used_network and s.last_iterations == K
        assert err <= 1e-6 * diag
        if k > 0:
            c2 = c
        c = (xr - y) * act


def test_history_features_and_targets():
    """Stored samples carry c_prev (features 6..8, flag 24), its change over
    the last two steps (features 3..5, flag 31: only when the two steps were
    consecutive) and the target (x_{n+1} - x_inertia - h) / (dt^2 |g|), with
    the history part h = c_prev, or 2 c_prev - c_prev2 after two consecutive
    steps, all checked against corrections computed on the host from the
    exported states; a plain step in between makes the next sample's change
    feature absent (and h first order), and a reset clears both."""
    sc, s = make_solver((12, 3, 3))
    dt = 1e-3
    g = np.asarray(sc.gravity, np.float64)
    S = dt * dt * np.linalg.norm(g)
    act = sc.active_mask[:, None].astype(np.float64)
    ws = WarmStart(s, capacity=8)
    rng = np.random.default_rng(3)
    s.old_velocities = rng.normal(0, 0.05, size=(sc.n_vertices, 3)) * act
    plan = ["collect", "collect", "collect", "plain", "collect", "collect", "reset", "collect", "collect"]
    corr = []  # host corrections of finished steps, None across a reset
    slot = 0
    for what in plan:
        if what == "reset":
            s.reset()
            s.old_velocities = rng.normal(0, 0.05, size=(sc.n_vertices, 3)) * act
            corr.append(None)
            continue
        x0, v0 = s.old_positions.numpy(), s.old_velocities.numpy()
        y = x0 + (dt * v0 + dt * dt * g) * act
        c1 = corr[-1] if corr else None
        c2 = corr[-2] if len(corr) > 1 and c1 is not None else None
        if what == "collect":
            ws.collect(dt, 1e-7)
            F, T, _, _ = ws.sample(slot)
            slot += 1
        else:
            s.step(s.old_positions, dt)
            F = None
        x1 = s.old_positions.numpy()
        if F is not None:
            tol = lambda e: np.abs(e).max() * 2.0 ** -7 + 1e-3
            if c1 is None:
                assert np.all(F[:, 6:9] == 0) and np.all(F[:, 24] == 0) and np.all(F[:, 31] == 0)
                t_exp = (x1 - y) * act / S
            else:
                assert np.all(F[:, 24] == 1)
                assert np.abs(F[:, 6:9] - c1 / S).max() <= tol(c1 / S)
                t_exp = (x1 - y - (c1 if c2 is None else 2.0 * c1 - c2)) * act / S
                if c2 is None:
                    assert np.all(F[:, 31] == 0) and np.all(F[:, 3:6] == 0)
                else:
                    assert np.all(F[:, 31] == 1)
                    d = (c1 - c2) / S
                    assert np.abs(F[:, 3:6] - d).max() <= tol(d)
            # FP32 corrections on the device against FP64 on the host
            assert np.abs(T - t_exp).max() <= 2e-3 * np.abs(t_exp).max() + 2e-3
        corr.append((x1 - y) * act)
        if what == "plain" and len(corr) > 1:
            corr[-2] = None  # the network did not see that step: no change feature across it
    assert slot == 7


def test_history_policy_cuts_iterations():
    """HISTORY (per body the better of the plain guess and x_inertia + the
    history part, no network) on a stiff 32x4x4 cantilever, each step from the state of a
    converged FP64 oracle trajectory (tol 1e-11): fewer iterations than the
    plain step, and the step error against the converged step within 1.5x of
    the plain step's in mean and max."""
    import oracle
    nx = (32, 4, 4)
    sc, s = make_solver(nx)
    sc_b, sb = make_solver(nx)
    E = 1e6
    for sol in (s, sb):
        sol.set_material(E, 0.4)
    ws = WarmStart(s, capacity=1)
    ws.set_policy(ws.HISTORY)
    ref = oracle.OracleSolver(**dict(sc.solver_kwargs(), youngs_modulus=E), dx_tol=1e-11, max_iter=100000,
                              expanded=False, threads=8)
    (ih, ip), (eh, ep) = resync_errors(ref, [s, sb], 25, warm=ws)
    print(f"iterations history {ih.mean():.2f} plain {ip.mean():.2f}; step error history {eh.mean():.3e} "
          f"(max {eh.max():.3e}) plain {ep.mean():.3e} (max {ep.max():.3e})")
    assert ih[1:].mean() < 0.7 * ip[1:].mean()
    assert eh.mean() <= 1.5 * ep.mean() and eh.max() <= 1.5 * ep.max()


@pytest.mark.parametrize("E", [1e5, 1e6])
def test_history_trajectory_error(E):
    """Free-running trajectories (no resync) of a 24x4x4 cantilever over 100
    steps at tol 1e-7, plain and with the HISTORY policy, against a converged
    FP64 oracle trajectory (tol 1e-11): the history trajectory's tip error
    stays within 1.5x of the plain trajectory's (the FP64 oracle run the same
    way over 200 steps gives 2.3e-4 vs 1.0e-3 of the tip range at E = 1e5 and
    1.4e-3 vs 1.6e-1 at E = 1e6 from the extrapolated start: stopping on the
    update size leaves an error that grows with stiffness, and a start nearer
    the solution leaves less)."""
    import oracle
    nx = (24, 4, 4)
    sc, s = make_solver(nx)
    sc_b, sb = make_solver(nx)
    for sol in (s, sb):
        sol.set_material(E, 0.4)
    wh = WarmStart(s, capacity=1)
    wh.set_policy(wh.HISTORY)
    ref = oracle.OracleSolver(**dict(sc.solver_kwargs(), youngs_modulus=E), dx_tol=1e-11, max_iter=100000,
                              expanded=False, threads=8)
    tip = sc.vertices[:, 0] >= sc.vertices[:, 0].max() - 1e-9
    th, tp, tr, ih, ip = [], [], [], [], []
    for k in range(100):
        xr = ref.step(1e-3)
        xh = s.step(s.old_positions, 1e-3, warm_start=wh).numpy()
        xp = sb.step(sb.old_positions, 1e-3).numpy()
        ih.append(s.last_iterations)
        ip.append(sb.last_iterations)
        for t, x in ((th, xh), (tp, xp), (tr, xr)):
            t.append((x[tip] - sc.vertices[tip]).mean(0))
    th, tp, tr = (np.array(a) for a in (th, tp, tr))
    rng = np.abs(tr).max()
    eh, ep = np.abs(th - tr).max(), np.abs(tp - tr).max()
    print(f"E {E:.0e}: iterations history {np.mean(ih):.2f} plain {np.mean(ip):.2f}; tip error history "
          f"{eh:.3e} plain {ep:.3e} (tip range {rng:.3e})")
    assert np.mean(ih) < np.mean(ip)
    assert eh <= 1.5 * ep + 1e-6 * rng


@pytest.mark.parametrize("mode", ["launches", "resident"])
def test_batched_fallback_bit_exact(mode):
    """Per-body fallback: with a far-off network no body takes the network
    guess; the first batched step equals the plain batched step exactly and
    every step equals the HISTORY policy's."""
    from warpvbd_amd import BatchedVBDSolver
    from warpvbd_amd.scenes import mass_spring
    bodies = [cantilever((10, 2, 2)), mass_spring(6, 3), cantilever((14, 3, 2), youngs_modulus=4e4)]
    a = BatchedVBDSolver(bodies, dx_tol=1e-7, max_iter=1000, verbose=False)
    b = BatchedVBDSolver(bodies, dx_tol=1e-7, max_iter=1000, verbose=False)
    c = BatchedVBDSolver(bodies, dx_tol=1e-7, max_iter=1000, verbose=False)
    for sol in (a, b, c):
        sol.set_mode(mode, 8, 256 if mode == "launches" else 1024, True)
    ws = WarmStart(a, capacity=1, seed=4)
    p = init_params(4)
    p["b_out"] = np.array([50.0, -80.0, 120.0])
    ws.set_params(flatten(p))
    wh = WarmStart(c, capacity=1)
    wh.set_policy(wh.HISTORY)
    for k in range(8):
        xa = a.step(dt=1e-3, warm_start=ws).numpy()
        assert a.last_used_network == 0
        xc = c.step(dt=1e-3, warm_start=wh).numpy()
        assert np.array_equal(xa, xc) and np.array_equal(a.body_iterations, c.body_iterations)
        if k == 0:
            xb = b.step(dt=1e-3).numpy()
            assert np.array_equal(xa, xb) and np.array_equal(a.body_iterations, b.body_iterations)


def test_batched_warm_start_per_body_choice():
    """A network trained on one body, applied to a batch: 10 steps from the
    network guess (every body), then 20 where the step chooses per body; each
    body's one-step error against its converged FP64 oracle step (tol 1e-11,
    per-step resync) stays that of its plain step (within 1.5x in mean and
    max, as in the single-body gain test: both stop once the update falls
    below the tolerance, with errors of the tolerance's order)."""
    import oracle
    from warpvbd_amd import BatchedVBDSolver
    sc, s = make_solver((24, 4, 4))
    ws1 = WarmStart(s, capacity=200, seed=6)
    ws1.generate(trajectories=5, steps=40, seed=11)
    ws1.train(epochs=25, batch=8, lr=2e-3, seed=2)
    bodies = [cantilever((24, 4, 4)), cantilever((20, 4, 4), youngs_modulus=3e5), cantilever((28, 3, 3))]
    a = BatchedVBDSolver(bodies, dx_tol=1e-7, max_iter=1000, verbose=False)
    b = BatchedVBDSolver(bodies, dx_tol=1e-7, max_iter=1000, verbose=False)
    refs = [oracle.OracleSolver(**dict(body.solver_kwargs(), damping_coefficient=a.damping_coefficient),
                                dx_tol=1e-11, max_iter=100000, expanded=False, threads=8) for body in bodies]
    ws = WarmStart(a, capacity=1)
    ws.set_params(ws1.get_params())
    used = forced = 0
    nb = len(bodies)
    it_w, it_p = np.zeros((30, nb)), np.zeros((30, nb))
    e_w, e_p = np.zeros((30, nb)), np.zeros((30, nb))
    for k in range(30):
        # the first 10 steps start every body from the network guess (the
        # history guess usually wins the choice), then each body chooses
        ws.set_policy(ws.NETWORK if k < 10 else ws.CHOOSE)
        x0 = np.concatenate([r.old_positions for r in refs])
        v0 = np.concatenate([r.old_velocities for r in refs])
        xr = [r.step(1e-3) for r in refs]
        for sol in (a, b):
            sol.old_positions = x0
            sol.old_velocities = v0
        xa = a.step(dt=1e-3, warm_start=ws).numpy()
        if k < 10:
            forced += a.last_used_network
        else:
            used += a.last_used_network
        xb = b.step(dt=1e-3).numpy()
        it_w[k], it_p[k] = a.body_iterations, b.body_iterations
        for i in range(nb):
            e_w[k, i] = np.abs(xa[a.body_slice(i)] - xr[i]).max()
            e_p[k, i] = np.abs(xb[b.body_slice(i)] - xr[i]).max()
    print("bodies x steps that chose the network:", used, "of", 20 * nb)
    print("iterations warm", it_w.mean(0), "plain", it_p.mean(0))
    print("mean step error warm", e_w.mean(0), "plain", e_p.mean(0))
    assert forced == 10 * nb
    for i in range(nb):
        assert e_w[:, i].mean() <= 1.5 * e_p[:, i].mean() and e_w[:, i].max() <= 1.5 * e_p[:, i].max()
