"""The SNH slot arithmetic of the sweeps (warpvbd_amd/csrc/vbd_device.h,
slot_eval) forms cof(F) b and J - 1 from the rest-edge constants of the slot's
dictionary entry instead of from F's rows.  These CPU checks pin the two
identities it rests on against numpy's determinant and inverse of F, on every
slot of the five voxel tets (hex2tets) under random deformations:

* with M = [b0 b1 bs] (columns), the rows of M^-1 are the rest edges
  E0 = X_o0 - X_o2, E1 = X_o1 - X_o2, R = X_self - X_o2, and k = det M;
* cof(F) bs = bs + c,  c = P x e1 + e0 x (Q + k e1),  P = k E0,  Q = k E1;
* J - 1 = c . R - cof(F) bs . d2,
  with e0 = u_o0 - u_o2, e1 = u_o1 - u_o2, d2 = u_o2 - u_self.

A swapped edge, a wrong sign of a cross product or the wrong base vertex fails
them.  The last check evaluates the form in FP32 at a strain of 1e-6, where
J - 1 has to keep its relative accuracy (the reason for the form).
"""
import numpy as np

import oracle
from warpvbd_amd._utils import hex2tets, voxel2hex


def _voxel_tets(h=0.01):
    V, H = voxel2hex(np.ones((1, 1, 1), bool), h, h, h)
    return V, hex2tets(H)


def _slots(oracle_lib):
    """(X, b of the tet's 4 vertices, local index, the others in ascending order) per slot."""
    V, T = _voxel_tets()
    invDm = oracle.rest_shape(V, T, np.ones(len(T)))[0]
    for e in range(len(T)):
        # b of local index i < 3 is row i of invDm; b of index 3 minus their sum
        b = [invDm[e][0], invDm[e][1], invDm[e][2], -invDm[e].sum(axis=0)]
        X = V[T[e]]
        assert np.allclose(sum(np.outer(X[j], b[j]) for j in range(4)), np.eye(3), atol=1e-12)
        for loc in range(4):
            yield X, b, loc, [k for k in range(4) if k != loc]


def _edge_consts(b, loc, oth):
    M = np.column_stack([b[oth[0]], b[oth[1]], b[loc]])
    k = np.linalg.det(M)
    E0, E1, R = np.linalg.inv(M)
    return k * E0, k * E1, k, R, (E0, E1, R)


def _edge_form(bs, P, Q, k, R, e0, e1, d2):
    c = np.cross(P, e1) + np.cross(e0, Q + k * e1)
    cb = bs + c
    return cb, c @ R - cb @ d2


def test_rest_edges_are_the_rows_of_the_inverse(oracle_lib):
    for X, b, loc, oth in _slots(oracle_lib):
        P, Q, k, R, (E0, E1, R) = _edge_consts(b, loc, oth)
        assert np.allclose(E0, X[oth[0]] - X[oth[2]], atol=1e-14)
        assert np.allclose(E1, X[oth[1]] - X[oth[2]], atol=1e-14)
        assert np.allclose(R, X[loc] - X[oth[2]], atol=1e-14)
        assert np.allclose(b[loc], k * np.cross(E0, E1), rtol=1e-12)
        assert np.isclose(b[loc] @ R, 1.0, rtol=1e-13)


def test_cofactor_and_determinant_from_rest_edges(oracle_lib):
    rng = np.random.default_rng(21)
    n = 0
    for X, b, loc, oth in _slots(oracle_lib):
        P, Q, k, R, _ = _edge_consts(b, loc, oth)
        for scale in (1e-4, 2e-3, 4e-3):  # up to 40% strain on a 1 cm voxel
            x = X + scale * rng.standard_normal((4, 3))
            F = sum(np.outer(x[j], b[j]) for j in range(4))
            u = x - X
            e0, e1, d2 = u[oth[0]] - u[oth[2]], u[oth[1]] - u[oth[2]], u[oth[2]] - u[loc]
            cb, jm1 = _edge_form(b[loc], P, Q, k, R, e0, e1, d2)
            J = np.linalg.det(F)
            cof = J * np.linalg.inv(F).T
            assert np.allclose(cb, cof @ b[loc], rtol=1e-11, atol=1e-11 * np.abs(b[loc]).max())
            assert np.isclose(jm1, J - 1.0, rtol=1e-9, atol=1e-14)
            n += 1
    assert n == 5 * 4 * 3


def test_edge_form_keeps_small_strains_in_fp32(oracle_lib):
    rng = np.random.default_rng(22)
    f32 = np.float32
    worst = 0.0
    for X, b, loc, oth in _slots(oracle_lib):
        P, Q, k, R, _ = _edge_consts(b, loc, oth)
        u = 1e-8 * rng.standard_normal((4, 3))  # strain about 1e-6 on a 1 cm voxel
        F = np.eye(3) + sum(np.outer(u[j], b[j]) for j in range(4))
        exact = np.linalg.det(F) - 1.0
        e0, e1, d2 = u[oth[0]] - u[oth[2]], u[oth[1]] - u[oth[2]], u[oth[2]] - u[loc]
        args = [np.asarray(a, f32) for a in (b[loc], P, Q, k, R, e0, e1, d2)]
        _, jm1 = _edge_form(*args)
        assert jm1.dtype == f32
        worst = max(worst, abs(float(jm1) - exact) / abs(exact))
    # absolute FP32 coordinates would leave J - 1 with an error near 1e-7, a
    # tenth of its size; the form keeps about FP32's relative precision
    assert worst < 1e-4, worst
