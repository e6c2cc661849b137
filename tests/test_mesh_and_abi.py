"""CPU tests of the host side: the mesh generator against a loop-for-loop
restatement of src/_utils.py:35-108, the scene rules, and the C ABI library
(loads, exports every entry point include/warpvbd.h declares)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from warpvbd_amd import _lib
from warpvbd_amd._utils import benchmark_functions, hex2tets, voxel2hex
from warpvbd_amd.scenes import cantilever, mass_spring

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("shape", [(1, 1, 1), (10, 2, 2), (3, 4, 5), (7, 1, 3)])
def test_voxel2hex_matches_loops(shape):
    rng = np.random.default_rng(sum(shape))
    vox = rng.random(shape) < 0.7
    vox.flat[0] = True
    for dx, dy, dz in [(0.01, 0.01, 0.01), (0.5, 0.25, 2.0)]:
        V, H = voxel2hex(vox, dx, dy, dz)
        V2, H2 = oracle.voxel2hex_loops(vox, dx, dy, dz)
        assert np.array_equal(V, V2) and V.dtype == V2.dtype
        assert np.array_equal(H, H2)
        assert np.array_equal(hex2tets(H), oracle.hex2tets_loops(H2))


@pytest.mark.parametrize("shape", [(10, 2, 2), (128, 16, 16), (256, 32, 32)])
def test_mesh_counts(shape):
    nx, ny, nz = shape
    V, H = voxel2hex(np.ones(shape, bool), 0.01, 0.01, 0.01)
    T = hex2tets(H)
    assert len(V) == (nx + 1) * (ny + 1) * (nz + 1)
    assert len(T) == 5 * nx * ny * nz
    # vertex id = i + (nx+1)(j + (ny+1)k)
    i, j, k = 3 % (nx + 1), 1, 1
    assert np.allclose(V[i + (nx + 1) * (j + (ny + 1) * k)], [0.01 * i, 0.01 * j, 0.01 * k])


def test_scene_rules():
    sc = cantilever((10, 2, 2))
    assert (~sc.active_mask).sum() == 9 and len(sc.tip_idx) == 9 and sc.n_vertices == 99
    big = cantilever((256, 32, 32))
    assert (~big.active_mask).sum() == 1089 and big.n_elements == 1310720
    ms = mass_spring(6, 3)
    assert (~ms.active_mask).sum() == 4


def test_benchmark_functions_signature():
    calls = []

    def f():
        calls.append(1)

    t = benchmark_functions([f], num_samples=3, skipOne=True)
    assert len(calls) == 4 and len(t["f"]) == 3


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "warpvbd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(vbd_\w+)\s*\(", src, flags=re.M)))


def test_c_abi_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)
    assert isinstance(L.vbd_last_error(), bytes)


def test_c_abi_rejects_bad_arguments_without_touching_a_device():
    L = _lib.load()
    h = ctypes.c_void_p()
    X = np.zeros((4, 3))
    T = np.array([[0, 1, 2, 7]], np.int32)
    one = np.ones(1)
    g = np.zeros(3)
    rc = L.vbd_create(_lib.ptr(X, ctypes.c_double), 4, _lib.ptr(T, ctypes.c_int32), 1, _lib.ptr(one, ctypes.c_double),
                      _lib.ptr(one, ctypes.c_double), _lib.ptr(one, ctypes.c_double), None,
                      _lib.ptr(g, ctypes.c_double), 0.0, 0, ctypes.byref(h))
    assert rc == -4 and b"outside" in L.vbd_last_error()
    rc = L.vbd_create(_lib.ptr(X, ctypes.c_double), 4, _lib.ptr(T, ctypes.c_int32), 1, _lib.ptr(one, ctypes.c_double),
                      _lib.ptr(one, ctypes.c_double), _lib.ptr(one, ctypes.c_double), None,
                      _lib.ptr(g, ctypes.c_double), 0.0, 7, ctypes.byref(h))
    assert rc == -1
    assert L.vbd_step(None, None, 1e-3, 1e-7, 10, None, None, None) == -1
    assert L.vbd_set_launch(None, 4, 256, 0, 0) == -1
