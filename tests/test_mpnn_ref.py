"""Pins the FP64 reference of the warm-start network (oracle/mpnn_ref.py):
gradients against central finite differences of the loss (BF16 rounding off,
where the loss is smooth), the neighbour mean against a dense matrix product,
and BF16 rounding against known values."""
import numpy as np

from oracle import mpnn_ref as M


def small_graph(n=40, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.random((n, n)) < 0.15
    A = np.triu(A, 1)
    A = A | A.T
    nb_off = np.concatenate([[0], np.cumsum(A.sum(1))]).astype(np.int64)
    nb = np.concatenate([np.nonzero(A[i])[0] for i in range(n)]).astype(np.int64)
    return A, nb_off, nb


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3, 0.0])
    r = M.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie 1 + 2^-8 rounds to even
    assert r[2] == 1.0078125
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8 + 1e-300)


def test_neighbour_mean_is_dense_product_and_adjoint():
    A, nb_off, nb = small_graph()
    rng = np.random.default_rng(1)
    h = rng.standard_normal((len(A), 5))
    deg = A.sum(1)
    P = np.where(deg[:, None] > 0, A / np.maximum(deg, 1)[:, None], 0.0)
    assert np.allclose(M.neighbour_mean(h, nb_off, nb), P @ h)
    da = rng.standard_normal((len(A), 5))
    assert np.allclose(M.neighbour_mean_transpose(da, nb_off, nb), P.T @ da)


def test_gradients_match_finite_differences():
    A, nb_off, nb = small_graph(30, 2)
    n = len(A)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((n, M.F_IN))
    t = rng.standard_normal((n, 3))
    mask = rng.random(n) < 0.8
    p = M.init_params(4)
    for k in p:  # non-zero biases so every ReLU branch is exercised
        if k.startswith("b_"):
            p[k] = rng.normal(0, 0.1, p[k].shape)
    loss, g, _ = M.loss_and_grads(p, f, nb_off, nb, t, mask, rounding=False)
    eps = 1e-6
    for name in p:
        flat = p[name].reshape(-1)
        idx = rng.choice(len(flat), size=min(6, len(flat)), replace=False)
        for i in idx:
            old = flat[i]
            flat[i] = old + eps
            lp = M.loss_and_grads(p, f, nb_off, nb, t, mask, rounding=False)[0]
            flat[i] = old - eps
            lm = M.loss_and_grads(p, f, nb_off, nb, t, mask, rounding=False)[0]
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            an = g[name].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, an)
