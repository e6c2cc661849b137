"""GPU tests of the batched solver: every body of a batch against its own
single-body FP64 oracle solve (colour-major offset table, per-body
convergence, both execution forms, hipGraph replay, autotuner)."""
import numpy as np
import pytest

import oracle
from warpvbd_amd import BatchedVBDSolver
from warpvbd_amd.scenes import cantilever, mass_spring

pytestmark = pytest.mark.gpu


def bodies():
    out = [mass_spring(6, 3), mass_spring(7, 5, youngs_modulus=3e4, poissons_ratio=0.35),
           cantilever((8, 2, 2), youngs_modulus=2e5), cantilever((12, 3, 2), poissons_ratio=0.3),
           mass_spring(6, 4, youngs_modulus=5e5)]
    out[3].gravity = (0.0, 1.5, -9.7)
    return out


def diag(sc):
    return float(np.linalg.norm(sc.vertices.max(0) - sc.vertices.min(0)))


@pytest.mark.parametrize("mode,lanes,block,graph", [("launches", 4, 256, True), ("launches", 2, 128, False),
                                                    ("resident", 4, 1024, False), ("resident", 8, 512, False)])
def test_batch_fixed_iterations_match_single_body_oracle(mode, lanes, block, graph):
    bs = bodies()
    g = BatchedVBDSolver(bs, dx_tol=0.0, max_iter=8, verbose=False)
    g.set_mode(mode, lanes, block, graph)
    refs = [oracle.OracleSolver(**b.solver_kwargs(), dx_tol=0.0, max_iter=8) for b in bs]
    for k in range(6):
        x = g.step(dt=1e-3).numpy()
        assert np.all(g.body_iterations == 8)
        for i, (b, r) in enumerate(zip(bs, refs)):
            xr = r.step(1e-3)
            err = np.abs(x[g.body_slice(i)] - xr).max()
            assert err < 1e-6 * diag(b), (mode, k, i, err)


@pytest.mark.parametrize("mode", ["launches", "resident"])
def test_batch_tolerance_per_body_convergence(mode):
    """Tolerance termination, body by body, against two FP64 runs: the oracle
    at the same tolerance (iteration counts within one, 90% equal; positions
    within 1e-5 x diagonal; the tip curve within 1e-4 relative for a body
    whose counts all matched), and the oracle replaying the GPU's iteration
    count of each step (the tip curve within 1e-4 relative for every body).
    A count one apart moves a tolerance-terminated trajectory by as much as
    the tolerance allows (the oracle at tol 1e-7 against itself at 1e-10
    differs by 6e-3 to 5e-2 relative on these bodies), so the replay is what
    separates the solver's arithmetic from the termination decision."""
    bs = bodies()
    g = BatchedVBDSolver(bs, dx_tol=1e-7, max_iter=1000, verbose=False)
    g.set_mode(mode, 4, 256 if mode == "launches" else 1024, True)
    refs = [oracle.OracleSolver(**b.solver_kwargs(), dx_tol=1e-7, max_iter=1000) for b in bs]
    reps = [oracle.OracleSolver(**b.solver_kwargs(), dx_tol=0.0, max_iter=1000) for b in bs]
    diffs = []
    flips = [0] * len(bs)
    tip_g, tip_r, tip_p = ([[] for _ in bs] for _ in range(3))
    for k in range(30):
        x = g.step(dt=1e-3).numpy()
        for i, (b, r, q) in enumerate(zip(bs, refs, reps)):
            xr = r.step(1e-3)
            xq = q.step(1e-3, max_iter=int(g.body_iterations[i]))
            d = abs(int(g.body_iterations[i]) - r.last_iters)
            diffs.append(d)
            flips[i] += d != 0
            assert d <= 1
            xb = x[g.body_slice(i)]
            # tolerance termination: each step may differ by O(tol) per iteration-count flip
            assert np.abs(xb - xr).max() < 1e-5 * diag(b)
            tip_g[i].append(xb[b.tip_idx, 2].mean())
            tip_r[i].append(xr[b.tip_idx, 2].mean())
            tip_p[i].append(xq[b.tip_idx, 2].mean())
    assert np.mean(np.array(diffs) == 0) > 0.9
    for i, b in enumerate(bs):
        z0 = b.vertices[b.tip_idx, 2].mean()
        zg, zr, zp = np.array(tip_g[i]) - z0, np.array(tip_r[i]) - z0, np.array(tip_p[i]) - z0
        rel = np.abs(zg - zr).max() / np.abs(zr).max()
        rel_p = np.abs(zg - zp).max() / np.abs(zp).max()
        print(f"{mode} body {i}: {flips[i]} count flips; tip curve max rel err {rel:.2e} (same tolerance), "
              f"{rel_p:.2e} (same counts)")
        assert rel_p < 1e-4
        if flips[i] == 0:
            assert rel < 1e-4


def test_live_range_launches_equal_full_launches():
    """Colour launches restricted to the live bodies' ranges (after the first
    chunk of a step) do exactly the per-vertex work of full launches: same
    iteration counts and bit-identical positions at tol 1e-7."""
    bs = bodies() + [cantilever((20, 3, 3), youngs_modulus=8e5), mass_spring(5, 3, youngs_modulus=2e4)]
    a = BatchedVBDSolver(bs, dx_tol=1e-7, max_iter=1000, verbose=False)
    b = BatchedVBDSolver(bs, dx_tol=1e-7, max_iter=1000, verbose=False)
    a.set_mode("launches", 8, 256, True)
    b.set_mode("launches_full", 8, 256, True)
    assert b.get_mode()["mode"] == "launches_full"
    spread = 0
    for k in range(25):
        xa = a.step(dt=1e-3).numpy()
        xb = b.step(dt=1e-3).numpy()
        assert np.array_equal(a.body_iterations, b.body_iterations)
        assert np.array_equal(xa, xb)
        spread = max(spread, int(a.body_iterations.max() - a.body_iterations.min()))
    assert spread > 2  # bodies converge at different counts, so the live lists were used


def test_batch_autotune_restores_state_and_keeps_parity(tmp_path, monkeypatch):
    monkeypatch.setenv("WARPVBD_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    bs = bodies()
    g = BatchedVBDSolver(bs, dx_tol=1e-7, max_iter=1000, verbose=False)
    g.step(dt=1e-3)
    x_before = g.old_positions.numpy().copy()
    res = g.autotune(1e-3, trial_steps=2)
    assert res["mode"] in ("launches", "launches_full", "resident") and not res["cached"]
    assert np.array_equal(g.old_positions.numpy(), x_before)
    res2 = g.autotune(1e-3, trial_steps=2)
    assert res2["cached"] and res2["mode"] == res["mode"]
    refs = [oracle.OracleSolver(**b.solver_kwargs(), dx_tol=0.0, max_iter=6) for b in bs]
    g2 = BatchedVBDSolver(bs, dx_tol=0.0, max_iter=6, verbose=False)
    g2.set_mode(res["mode"], res["lanes_per_vertex"], res["block_size"], res["use_graph"])
    for k in range(3):
        x = g2.step(dt=1e-3).numpy()
        for i, (b, r) in enumerate(zip(bs, refs)):
            assert np.abs(x[g2.body_slice(i)] - r.step(1e-3)).max() < 1e-6 * diag(b)


def test_batch_rejects_element_across_bodies():
    from warpvbd_amd._lib import VBDError
    bs = [mass_spring(6, 3), mass_spring(6, 3)]
    g = BatchedVBDSolver(bs, verbose=False)
    assert g.n_bodies == 2
    T = g.elements.copy()
    T[0, 0] = g.body_vertex_offsets[1]  # points into the second body
    bad = [dict(initial_positions=g.initial_positions, elements=T, densities=1000.0, youngs_modulus=1e5,
                poissons_ratio=0.4, active_mask=None, gravity=(0, 0, -9.81))]
    # a single "body" containing both is fine; two bodies joined by an element is not
    BatchedVBDSolver(bad, verbose=False)
    import ctypes
    from warpvbd_amd import _lib
    h = ctypes.c_void_p()
    X = np.ascontiguousarray(g.initial_positions)
    Ti = np.ascontiguousarray(T, dtype=np.int32)
    off = np.ascontiguousarray(g.body_vertex_offsets, dtype=np.int32)
    m = len(Ti)
    one = np.ones(m)
    nu = np.full(m, 0.3)
    G = np.zeros(6)
    rc = _lib.load().vbd_batch_create(_lib.ptr(X, ctypes.c_double), len(X), _lib.ptr(Ti, ctypes.c_int32), m,
                                      _lib.ptr(off, ctypes.c_int32), 2, _lib.ptr(one, ctypes.c_double),
                                      _lib.ptr(one, ctypes.c_double), _lib.ptr(nu, ctypes.c_double), None,
                                      _lib.ptr(G, ctypes.c_double), 0.0, 0, ctypes.byref(h))
    assert rc == -4
