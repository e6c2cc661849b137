"""FP64 CPU reference of the warm-start message-passing network -- TEST
INFRASTRUCTURE ONLY (the checker of warpvbd_amd/csrc/vbd_mpnn.hip).

The network has no counterpart in WarpVBD; it is specified here and the GPU
kernels must reproduce it.  The GPU stores features and activations in BF16
and multiplies BF16 operands on MFMA with FP32 accumulation; this reference
rounds the same values to BF16 (round to nearest even) and accumulates in
FP64, so forward values and gradients agree to FP32 accumulation error (and
the odd BF16 rounding that lands on the other side of a tie).

Network (F = 32 input features, W = 64 width, 4 message-passing layers):
    h0      = bf(relu(bf(f) bf(W_in)^T + b_in))
    a_l     = bf(mean over graph neighbours j of h_l[j])      (0 without neighbours)
    h_{l+1} = bf(h_l + relu([h_l, a_l] bf(W_l)^T + b_l))      l = 0..3
    out     = h_4 bf(W_out)^T + b_out                         (3 outputs)
    loss    = sum_i m_i |out_i - t_i|^2 / (3 sum_i m_i)       (m: active-vertex mask)
Backward as the GPU computes it: the data gradients pass through MFMA with
BF16-rounded upstream gradients (bf(dz) bf(W)); weight gradients use the
unrounded upstream gradient against the BF16 operands the forward used.

How it is pinned (tests/test_mpnn_ref.py): central finite differences of the
loss with respect to every parameter tensor (with BF16 rounding switched off,
where the loss is smooth), and the aggregation against a dense adjacency
matrix product.
"""
from __future__ import annotations

import numpy as np

F_IN, WIDTH, LAYERS, F_OUT = 32, 64, 4, 3


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def init_params(seed: int = 0) -> dict:
    """He-style initialisation (the GPU trainer starts from the same arrays)."""
    rng = np.random.default_rng(seed)
    p = {"W_in": rng.normal(0, np.sqrt(2.0 / F_IN), (WIDTH, F_IN)), "b_in": np.zeros(WIDTH)}
    for l in range(LAYERS):
        p[f"W_{l}"] = rng.normal(0, np.sqrt(2.0 / (2 * WIDTH)) * 0.5, (WIDTH, 2 * WIDTH))
        p[f"b_{l}"] = np.zeros(WIDTH)
    p["W_out"] = rng.normal(0, np.sqrt(1.0 / WIDTH) * 0.1, (F_OUT, WIDTH))
    p["b_out"] = np.zeros(F_OUT)
    return p


def neighbour_mean(h, nb_off, nb):
    deg = np.diff(nb_off)
    a = np.zeros_like(h)
    for i in range(len(deg)):
        if deg[i]:
            a[i] = h[nb[nb_off[i]:nb_off[i + 1]]].mean(axis=0)
    return a


def neighbour_mean_transpose(da, nb_off, nb):
    """Adjoint of neighbour_mean: dh[j] += sum over i with j in N(i) of da[i]/deg(i)."""
    deg = np.diff(nb_off)
    dh = np.zeros_like(da)
    for i in range(len(deg)):
        if deg[i]:
            dh[nb[nb_off[i]:nb_off[i + 1]]] += da[i] / deg[i]
    return dh


def forward(p, f, nb_off, nb, rounding=True):
    bf = bf16_round if rounding else (lambda x: np.asarray(x, np.float64))
    cache = {"x": [], "z": []}
    z0 = bf(f) @ bf(p["W_in"]).T + p["b_in"]
    h = bf(np.maximum(z0, 0.0))
    cache["f"], cache["z0"] = f, z0
    hs = [h]
    for l in range(LAYERS):
        a = bf(neighbour_mean(h, nb_off, nb))
        x = np.concatenate([h, a], axis=1)
        z = bf(x) @ bf(p[f"W_{l}"]).T + p[f"b_{l}"]
        cache["x"].append(x)
        cache["z"].append(z)
        h = bf(h + np.maximum(z, 0.0))
        hs.append(h)
    out = bf(h) @ bf(p["W_out"]).T + p["b_out"]
    cache["h"] = hs
    return out, cache


def loss_and_grads(p, f, nb_off, nb, target, mask, rounding=True):
    bf = bf16_round if rounding else (lambda x: np.asarray(x, np.float64))
    out, c = forward(p, f, nb_off, nb, rounding)
    m = mask.astype(np.float64)[:, None]
    denom = 3.0 * max(m.sum(), 1.0)
    diff = (out - target) * m
    loss = float((diff ** 2).sum() / denom)
    dout = 2.0 * diff / denom
    g = {}
    h4 = c["h"][-1]
    g["W_out"] = dout.T @ bf(h4)
    g["b_out"] = dout.sum(axis=0)
    dh = bf(dout) @ bf(p["W_out"])
    for l in reversed(range(LAYERS)):
        z, x = c["z"][l], c["x"][l]
        dz = dh * (z > 0)
        g[f"W_{l}"] = dz.T @ bf(x)
        g[f"b_{l}"] = dz.sum(axis=0)
        dx = bf(dz) @ bf(p[f"W_{l}"])
        dh = dh + dx[:, :WIDTH] + neighbour_mean_transpose(dx[:, WIDTH:], nb_off, nb)
    dz0 = dh * (c["z0"] > 0)
    g["W_in"] = dz0.T @ bf(c["f"])
    g["b_in"] = dz0.sum(axis=0)
    return loss, g, out
