"""ORACLE / TEST INFRASTRUCTURE: float64 restatement of the learned warm-start network.

The network (24 -> 128 -> 128 -> 3, SiLU) has no counterpart in the reference; this module is
the checker for the device kernels (reactphysics3d_amd/csrc/mlp.hip).  With ``bf16=True`` it
rounds to bfloat16 at exactly the points where the device does (normalised inputs, weights,
hidden activations, and the back-propagated errors dy, dz2, dz1), and otherwise works in
float64, so the device result differs from it only by FP32 accumulation order and expf.
With ``bf16=False`` it is the exact float64 network, whose gradient is pinned against
central finite differences in tests/test_mlp_oracle.py.

Only tests use this module.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 24, 128, 3
P_W1 = 0
P_B1 = P_W1 + N_IN * N_HID
P_W2 = P_B1 + N_HID
P_B2 = P_W2 + N_HID * N_HID
P_W3 = P_B2 + N_HID
P_B3 = P_W3 + N_HID * N_OUT
N_PARAMS = P_B3 + N_OUT


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[P_W1:P_B1].reshape(N_IN, N_HID), p[P_B1:P_W2], p[P_W2:P_B2].reshape(N_HID, N_HID), p[P_B2:P_W3],
            p[P_W3:P_B3].reshape(N_HID, N_OUT), p[P_B3:N_PARAMS])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, X, mu, inv_sigma, out_scale, bf=True):
    """Returns (impulses, cache) for features X [n, 24]."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xn = (np.asarray(X, np.float32) - np.asarray(mu, np.float32)) * np.asarray(inv_sigma, np.float32)
    x = q(xn)
    z1 = x @ q(W1) + b1
    h1 = z1 * _sig(z1)
    z2 = q(h1) @ q(W2) + b2
    h2 = z2 * _sig(z2)
    y = q(h2) @ q(W3) + b3
    cache = dict(x=x, z1=z1, h1=q(h1), z2=z2, h2=q(h2), y=y, W1=q(W1), W2=q(W2), W3=q(W3))
    return y * np.asarray(out_scale, np.float64), cache


def loss_and_grad(p, X, T, mu, inv_sigma, out_scale, bf=True):
    """loss = 1/(2n) sum |y - T/out_scale|^2 and its gradient with respect to the flat parameters."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    _, c = forward(p, X, mu, inv_sigma, out_scale, bf)
    n = len(X)
    t = np.asarray(T, np.float64) / np.asarray(out_scale, np.float64)
    e = c["y"] - t
    loss = 0.5 * np.sum(e * e) / n
    dy = q(e / n)
    s2 = _sig(c["z2"])
    dz2 = q((dy @ c["W3"].T) * (s2 * (1.0 + c["z2"] * (1.0 - s2))))
    s1 = _sig(c["z1"])
    dz1 = q((dz2 @ c["W2"].T) * (s1 * (1.0 + c["z1"] * (1.0 - s1))))
    g = pack(c["x"].T @ dz1, dz1.sum(0), c["h1"].T @ dz2, dz2.sum(0), c["h2"].T @ dy, dy.sum(0))
    return loss, g


def init_params(seed=1):
    """Small random parameters for tests (not the device's initialisation)."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (N_IN, N_HID)) * np.sqrt(6.0 / (N_IN + N_HID))
    W2 = rng.uniform(-1, 1, (N_HID, N_HID)) * np.sqrt(6.0 / (2 * N_HID))
    W3 = rng.uniform(-1, 1, (N_HID, N_OUT)) * np.sqrt(6.0 / (N_HID + N_OUT))
    return pack(W1, rng.normal(0, 0.1, N_HID), W2, rng.normal(0, 0.1, N_HID), W3, rng.normal(0, 0.1, N_OUT))
