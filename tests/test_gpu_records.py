"""GPU tests of the slot record layouts: a record dictionary with shared
entries (16 bytes per slot from HBM) gives bit for bit the results of one
entry per slot, in single-body and batched solvers, before and after a
material change (both with the SNH terms formed from F's rows); meshes without
shared rest shapes get one entry per slot.  The rest-edge form that small
shared dictionaries switch on agrees with F's rows to FP32 rounding."""
import numpy as np
import pytest

from warpvbd_amd import BatchedVBDSolver, VBDSolver, _lib
from warpvbd_amd.scenes import cantilever, mass_spring

pytestmark = pytest.mark.gpu


def _pair(make):
    """The same solver built with shared entries and with one entry per slot,
    both forming the SNH terms from F's rows (one entry per slot has no rest
    edges)."""
    _lib.use_edge_form(False)
    try:
        _lib.use_record_dictionary(True)
        a = make()
        _lib.use_record_dictionary(False)
        b = make()
    finally:
        _lib.use_record_dictionary(True)
        _lib.use_edge_form(True)
    assert not a.record_layout()["edge_form"] and not b.record_layout()["edge_form"]
    return a, b


def _step_both(a, b, steps, **kw):
    for k in range(steps):
        xa = a.step(a.old_positions, 1e-3, **kw).numpy()
        xb = b.step(b.old_positions, 1e-3, **kw).numpy()
        assert np.array_equal(xa, xb), k
        assert a.last_iterations == b.last_iterations, k


@pytest.mark.parametrize("material", ["snh", "stvk"])
@pytest.mark.parametrize("scene", ["cantilever", "mass_spring"])
def test_dictionary_bit_exact(scene, material):
    sc = cantilever((24, 4, 4)) if scene == "cantilever" else mass_spring(8, 5)
    a, b = _pair(lambda: VBDSolver(**sc.solver_kwargs(), dx_tol=1e-7, max_iter=1000, material=material,
                                   verbose=False))
    la, lb = a.record_layout(), b.record_layout()
    print(scene, material, la, lb)
    # 5 rest shapes x 4 local indices; FP32 rounding of V mu' may split a few
    assert 20 <= la["dictionary_entries"] <= 80 and la["bytes_per_slot"] == 16
    assert lb == {"dictionary_entries": a.info()["n_slots"], "bytes_per_slot": 64, "edge_form": False}
    for s in (a, b):
        s.set_launch(8, 256, 0, 0)  # colour launches (the whole-step kernel is checked below)
    _step_both(a, b, 15)
    for s in (a, b):
        s.set_launch(8, 256, 2048, 0)
        s.set_material(3e5, 0.35)  # a uniform material change rescales the dictionary
    _step_both(a, b, 10)


def test_dictionary_bit_exact_batched():
    bodies = [cantilever((12, 3, 3)), mass_spring(6, 4), cantilever((16, 3, 2), youngs_modulus=4e4),
              mass_spring(7, 3, youngs_modulus=6e5)]
    a, b = _pair(lambda: BatchedVBDSolver(bodies, dx_tol=1e-7, max_iter=1000, verbose=False))
    print(a.record_layout(), b.record_layout())
    assert 60 <= a.record_layout()["dictionary_entries"] <= 240  # three materials
    assert b.record_layout() == {"dictionary_entries": b.info()["n_slots"], "bytes_per_slot": 64, "edge_form": False}
    for mode in ("launches", "resident"):
        for s in (a, b):
            s.set_mode(mode, 8, 256 if mode == "launches" else 1024, True)
        for k in range(8):
            xa = a.step(dt=1e-3).numpy()
            xb = b.step(dt=1e-3).numpy()
            assert np.array_equal(xa, xb) and np.array_equal(a.body_iterations, b.body_iterations), (mode, k)


def test_irregular_mesh_keeps_full_records():
    """Jittered rest positions make every tet's shape different: one
    dictionary entry per slot."""
    sc = cantilever((10, 3, 3))
    rng = np.random.default_rng(4)
    X = sc.vertices + rng.uniform(-1e-3, 1e-3, sc.vertices.shape)
    kw = dict(sc.solver_kwargs(), initial_positions=X)
    s = VBDSolver(**kw, dx_tol=1e-7, max_iter=1000, verbose=False)
    assert s.record_layout() == {"dictionary_entries": s.info()["n_slots"], "bytes_per_slot": 64, "edge_form": False}
    x = s.step(s.old_positions, 1e-3).numpy()
    assert np.all(np.isfinite(x)) and s.last_metric < 1e-7


@pytest.mark.parametrize("damping", [0.0, 1.0])
def test_edge_form_against_rows(damping):
    """cof(F) b and J - 1 from the rest edges (the default for a voxel mesh's
    shared dictionary) against the same terms from F's rows: positions within
    1e-7 of the diagonal after 12 steps of 6 iterations (undamped, a tip drop
    of 7.8 cm: a nonlinear deformation under gravity; the damped beam drops
    about 1 um); StVK uses neither."""
    sc = cantilever((24, 4, 4), youngs_modulus=3e4)
    kw = dict(sc.solver_kwargs(), dx_tol=0.0, max_iter=6, damping_coefficient=damping, verbose=False)
    a = VBDSolver(**kw)
    _lib.use_edge_form(False)
    try:
        b = VBDSolver(**kw)
    finally:
        _lib.use_edge_form(True)
    assert a.record_layout()["edge_form"] and not b.record_layout()["edge_form"]
    diag = float(np.linalg.norm(sc.vertices.max(0) - sc.vertices.min(0)))
    worst = 0.0
    for k in range(12):
        xa = a.step(a.old_positions, 1e-2).numpy()
        xb = b.step(b.old_positions, 1e-2).numpy()
        worst = max(worst, float(np.abs(xa - xb).max()))
    drop = float(sc.vertices[:, 2].min() - xa[:, 2].min())
    print(f"edge vs rows: {worst:.3e} m over a tip drop of {drop:.3e} m")
    assert drop > (1e-3 if damping == 0.0 else 1e-7) and worst < 1e-7 * diag, worst
