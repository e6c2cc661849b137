"""MI355X-native Vertex Block Descent.

Drop-in for WarpVBD's solver (``src/vbd_solver.py``), colouring
(``src/coloring.py``) and mesh helpers (``src/_utils.py``); the time step and
the colouring run as hand-written HIP kernels for gfx950 in libwarpvbd.so.
"""
from ._utils import benchmark_functions, hex2tets, voxel2hex
from .coloring import (assign_color, compute_adjacency_dict, find_neighbors, graph_coloring,
                       graph_coloring_wp)
from .vbd_solver import EPS, MAX_ELEMENTS_PER_VERTEX, VBDSolver, VertexArray
from .batched import BatchedVBDSolver

__all__ = ["VBDSolver", "BatchedVBDSolver", "VertexArray", "voxel2hex", "hex2tets", "benchmark_functions", "assign_color",
           "compute_adjacency_dict", "graph_coloring", "graph_coloring_wp", "find_neighbors", "EPS",
           "MAX_ELEMENTS_PER_VERTEX"]
