"""Graph colouring with the interface of WarpVBD's ``src/coloring.py``.

``graph_coloring`` and ``graph_coloring_wp`` return the reference's sequential
first-fit colouring (``src/coloring.py:40-72``) bit-exact, computed on the GPU
by libwarpvbd.so: the recurrence colour[v] = mex{colour[u] : u ~ v, u < v} is
iterated in parallel to its unique fixed point.  ``compute_adjacency_dict``,
``assign_color`` and ``find_neighbors`` are the reference's small host helpers
for building and inspecting the graph.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, i32, ptr, to_numpy

MAX_COLORS = 256  # src/coloring.py:10,50


def assign_color(coloring: np.ndarray, neighbors) -> int:
    """Lowest colour not used by an already-coloured neighbour (src/coloring.py:6-22)."""
    used = {int(coloring[n]) for n in neighbors if coloring[n] != -1}
    for c in range(MAX_COLORS):
        if c not in used:
            return c
    raise AssertionError("Exceeded maximum number of colors")


def compute_adjacency_dict(elements) -> tuple:
    """Vertex graph of the elements as ``{v: set(neighbours)}`` plus the largest
    neighbour-set size (src/coloring.py:24-38).  The vertex count is
    ``elements.max() + 1``."""
    e = np.asarray(to_numpy(elements), dtype=np.int64)
    nv = int(e.max()) + 1
    k = e.shape[1]
    pairs = np.concatenate([e[:, [i, j]] for i in range(k) for j in range(k) if i != j], axis=0)
    pairs = np.unique(pairs, axis=0)
    adjacency = {i: set() for i in range(nv)}
    starts = np.searchsorted(pairs[:, 0], np.arange(nv + 1))
    for v in range(nv):
        adjacency[v] = set(pairs[starts[v]:starts[v + 1], 1].tolist())
    maximal_degree = max((len(s) for s in adjacency.values()), default=0)
    return adjacency, maximal_degree


def _groups(coloring: np.ndarray) -> dict:
    groups = {}
    for c in range(int(coloring.max()) + 1 if len(coloring) else 0):
        members = np.nonzero(coloring == c)[0]
        if len(members):
            groups[c] = members.tolist()
    return groups


def graph_coloring(adjacency: dict, verbose: bool = True):
    """First-fit colouring of ``{v: neighbours}`` in vertex order
    (src/coloring.py:40-72).  Returns (coloring[N] int64, {colour: [vertices]})."""
    n = len(adjacency)
    off = np.zeros(n + 1, dtype=np.int64)
    lists = []
    for v in range(n):
        nb = np.fromiter((int(u) for u in adjacency[v]), dtype=np.int64)
        lists.append(nb)
        off[v + 1] = off[v] + len(nb)
    adj = i32(np.concatenate(lists) if lists else np.zeros(0))
    off32 = i32(off)
    colors = np.empty(n, np.int32)
    nc = ctypes.c_int32()
    check(_lib.load().vbd_color_graph(ptr(off32, ctypes.c_int32), ptr(adj, ctypes.c_int32) if len(adj) else None,
                                      n, ptr(colors, ctypes.c_int32), ctypes.byref(nc)), "vbd_color_graph")
    coloring = colors.astype(int)
    if verbose:
        print(f"Assigned {coloring.max() + 1} colors. Color distribution: {np.bincount(coloring)}")
    return coloring, _groups(coloring)


def find_neighbors(adjacency: np.ndarray, vertex: int) -> list:
    """Indices with a positive entry in row ``vertex`` of a dense adjacency
    matrix (src/coloring.py:75-83)."""
    return np.nonzero(np.asarray(adjacency)[vertex] > 0)[0].tolist()


def graph_coloring_wp(elements) -> np.ndarray:
    """Parallel colouring of a mesh's vertex graph on the GPU; returns the same
    colouring as ``graph_coloring(compute_adjacency_dict(elements)[0])``
    (the reference's unfinished parallel variant, src/coloring.py:85-133)."""
    e = i32(elements).reshape(len(to_numpy(elements)), -1)
    if e.shape[1] != 4:
        raise ValueError("graph_coloring_wp expects tetrahedra [M, 4]")
    n = int(e.max()) + 1
    colors = np.empty(n, np.int32)
    nc, md = ctypes.c_int32(), ctypes.c_int32()
    check(_lib.load().vbd_color(ptr(e, ctypes.c_int32), len(e), n, ptr(colors, ctypes.c_int32), ctypes.byref(nc),
                                ctypes.byref(md)), "vbd_color")
    return colors.astype(int)
