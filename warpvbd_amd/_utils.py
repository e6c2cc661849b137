"""Mesh generation and wall-clock helper with the API of WarpVBD's ``src/_utils.py``.

``voxel2hex`` and ``hex2tets`` produce bit-identical outputs to the reference
(``src/_utils.py:35-108``) but are vectorised over the voxel grid instead of
triple Python loops, so the 256x32x32 beam (262,144 hexes) builds in well under
a second.  ``benchmark_functions`` keeps the reference's signature
(``src/_utils.py:6-32``).
"""
from __future__ import annotations

import time

import numpy as np


def benchmark_functions(funcs: list, *args, num_samples: int = 10, skipOne=False, **kwargs) -> dict:
    """Time each callable ``num_samples`` times (plus one untimed warm-up call when
    ``skipOne``); returns ``{name: [seconds, ...]}`` (src/_utils.py:6-32)."""
    times = {f.__name__: [] for f in funcs}
    for i in range(num_samples + int(skipOne)):
        for f in funcs:
            t0 = time.perf_counter()
            f(*args, **kwargs)
            t1 = time.perf_counter()
            if not skipOne or i > 0:
                times[f.__name__].append(t1 - t0)
    return times


def voxel2hex(voxels, dx, dy, dz):
    """Voxel grid -> (vertices[N,3] float64, hexes[H,8] int64).

    Vertex ids run i fastest, then j, then k over the corners of occupied voxels;
    hexes are listed k-major, then j, then i, with corners ordered
    (i,j,k), (i+1,j,k), (i+1,j+1,k), (i,j+1,k) and the same four at k+1
    (src/_utils.py:35-89)::

             5 -------- 6
            /|         /|
           4 -------- 7 |
           | 1 -------|-2
           |/         |/
           0 -------- 3
    """
    vox = np.asarray(voxels).astype(bool)
    if vox.ndim != 3:
        raise ValueError("voxels must be a 3-D array")
    nx, ny, nz = vox.shape
    corner = np.zeros((nx + 1, ny + 1, nz + 1), dtype=bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                corner[di:di + nx, dj:dj + ny, dk:dk + nz] |= vox
    # numbering order: k outer, j, i inner  ->  C order over a (k, j, i) view
    ckji = corner.transpose(2, 1, 0)
    ids = np.full(ckji.shape, -1, dtype=np.int64)
    kk, jj, ii = np.nonzero(ckji)
    if len(ii) == 0:
        raise ValueError("need at least one occupied voxel")
    ids[kk, jj, ii] = np.arange(len(ii), dtype=np.int64)
    vertices = np.stack([dx * ii.astype(np.float64), dy * jj.astype(np.float64), dz * kk.astype(np.float64)], axis=1)

    hk, hj, hi = np.nonzero(vox.transpose(2, 1, 0))
    hexes = np.stack([
        ids[hk, hj, hi], ids[hk, hj, hi + 1], ids[hk, hj + 1, hi + 1], ids[hk, hj + 1, hi],
        ids[hk + 1, hj, hi], ids[hk + 1, hj, hi + 1], ids[hk + 1, hj + 1, hi + 1], ids[hk + 1, hj + 1, hi],
    ], axis=1)
    return vertices, hexes


# corner selection of the fixed 5-tet split (src/_utils.py:101-106)
_HEX_TO_TETS = np.array([[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]], dtype=np.int64)


def hex2tets(hex_elements):
    """Split every hex into the same 5 positive-volume tets (src/_utils.py:92-108)."""
    h = np.asarray(hex_elements, dtype=np.int64).reshape(-1, 8)
    return h[:, _HEX_TO_TETS].reshape(-1, 4)
