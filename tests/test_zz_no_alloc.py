"""vbd_color_noalloc, the colouring entry point that allocates nothing,
against vbd_color (same C ABI, host buffers, its own allocations): the
same colouring, colour count and largest degree, bit for bit, on two calls in
a row whose second mesh is smaller (the shared scratch array keeps the first
call's contents), and its errors for a mesh above the scratch's size and for
an element index out of range."""
import ctypes

import numpy as np
import pytest

from warpvbd_amd import _lib
from warpvbd_amd._utils import hex2tets, voxel2hex
from warpvbd_amd.scenes import cantilever, mass_spring

pytestmark = pytest.mark.gpu


def color_host(T, n):
    L = _lib.load()
    Ti = np.ascontiguousarray(T, dtype=np.int32)
    out = np.empty(n, np.int32)
    nc, md = ctypes.c_int32(), ctypes.c_int32()
    _lib.check(L.vbd_color(_lib.ptr(Ti, ctypes.c_int32), len(Ti), n, _lib.ptr(out, ctypes.c_int32), ctypes.byref(nc),
                           ctypes.byref(md)), "vbd_color")
    return out, nc.value, md.value


def color_noalloc(T, n, n_tets=None):
    import torch
    L = _lib.load()
    dev = torch.device("cuda:0")
    t = torch.as_tensor(np.ascontiguousarray(T, dtype=np.int32).reshape(-1), device=dev)
    c = torch.full((n,), -7, dtype=torch.int32, device=dev)
    nc, md = ctypes.c_int32(), ctypes.c_int32()
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = L.vbd_color_noalloc(ctypes.c_void_p(t.data_ptr()), len(T) if n_tets is None else n_tets, n,
                             ctypes.c_void_p(c.data_ptr()), ctypes.byref(nc), ctypes.byref(md), ctypes.c_void_p(stream))
    return rc, c.cpu().numpy(), nc.value, md.value


def meshes():
    big = cantilever((256, 32, 32))
    rng = np.random.default_rng(11)
    n = 300
    T = rng.integers(0, n, size=(700, 4))
    T = T[np.array([len(set(t)) == 4 for t in T])]
    T = rng.permutation(n)[T]
    return [("beam 256x32x32", big.elements, big.n_vertices),
            ("mass-spring 6+3", mass_spring(6, 3).elements, mass_spring(6, 3).n_vertices),
            ("shuffled random tets", T, n)]


def test_noalloc_colouring_equals_vbd_color_on_consecutive_calls():
    for name, T, n in meshes():
        print(name, flush=True)
        rc, c, nc, md = color_noalloc(T, n)
        assert rc == 0, _lib.load().vbd_last_error()
        cr, ncr, mdr = color_host(T, n)
        assert np.array_equal(c, cr) and nc == ncr and md == mdr, name


def test_noalloc_rejects_a_mesh_above_its_scratch():
    T = hex2tets(voxel2hex(np.ones((2, 1, 1), bool), 0.01, 0.01, 0.01)[1])
    # 5 M tets claimed for a 10-tet buffer: refused on the host before any read
    rc, _, _, _ = color_noalloc(T, 12, n_tets=5_000_000)
    assert rc == -1
    assert "scratch" in _lib.load().vbd_last_error().decode()
    # the entry point still works afterwards
    rc, c, nc, _ = color_noalloc(T, 12)
    assert rc == 0 and np.array_equal(c, color_host(T, 12)[0])


def test_noalloc_rejects_an_index_out_of_range():
    T = hex2tets(voxel2hex(np.ones((2, 1, 1), bool), 0.01, 0.01, 0.01)[1]).copy()
    T[3, 2] = 12
    rc, _, _, _ = color_noalloc(T, 12)
    assert rc == -4
