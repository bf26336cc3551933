"""CPU oracle of the rock-physics surrogate MLP (test infrastructure only).

A 5 -> 128 -> 128 -> 128 -> 3 tanh network in the augmented-parameter layout
of include/crava_hip.h (crv_mlp_*): W1a [128][16] with the layer-1 bias in
column 5, W2a / W3a [128][144] and W4a [32][144] with the bias in column 128.
It has no counterpart in the reference (the reference evaluates rplib per
sample, rplib/distributionsrockdem.cpp:125-156); the network's forward and
backward are restated here in float64 in two forms:

* ``exact``: plain float64 arithmetic.  Its gradient is pinned by central
  finite differences (tests/test_oracle.py).
* ``bf16``: the same network with the GPU kernels' roundings -- weights,
  inputs, activations, the output gradient and the hidden-layer gradients
  rounded to bfloat16 (round to nearest even), products summed in float64.
  The GPU gradient is compared against this form.
"""
from __future__ import annotations

import numpy as np

HID, KA, K1, NO = 128, 144, 16, 32
OFF_W1 = 0
OFF_W2 = OFF_W1 + HID * K1
OFF_W3 = OFF_W2 + HID * KA
OFF_W4 = OFF_W3 + HID * KA
NPARAMS = OFF_W4 + NO * KA


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)) << np.uint64(16)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_W2].reshape(HID, K1), p[OFF_W2:OFF_W3].reshape(HID, KA),
            p[OFF_W3:OFF_W4].reshape(HID, KA), p[OFF_W4:].reshape(NO, KA))


def pack(W1, W2, W3, W4):
    return np.concatenate([W1.ravel(), W2.ravel(), W3.ravel(), W4.ravel()])


def features(x, norm) -> np.ndarray:
    """Normalised inputs as the kernel forms them (float32 arithmetic)."""
    x = np.asarray(x, np.float32).copy()
    for f in range(5):
        if norm["in_log_mask"] & (1 << f):
            x[:, f] = np.log10(x[:, f])
    lo = np.asarray(norm["in_lo"], np.float32)
    sc = np.asarray(norm["in_scale"], np.float32)
    return ((x - lo) * sc - np.float32(1.0)).astype(np.float32)


def _aug(h, width):
    out = np.zeros((h.shape[0], width))
    out[:, :h.shape[1]] = h
    return out


def forward(params, x, norm, mode="bf16"):
    """Returns (y_normalised (n,3), cache)."""
    rnd = bf16 if mode == "bf16" else (lambda a: np.asarray(a, np.float64))
    W1, W2, W3, W4 = (rnd(w) for w in unpack(params))
    xt = features(x, norm)
    Xa = _aug(rnd(xt), K1)
    Xa[:, 5] = 1.0
    Z1 = Xa @ W1.T
    H1 = rnd(np.tanh(Z1))
    H1a = _aug(H1, KA)
    H1a[:, HID] = 1.0
    H2 = rnd(np.tanh(H1a @ W2.T))
    H2a = _aug(H2, KA)
    H2a[:, HID] = 1.0
    H3 = rnd(np.tanh(H2a @ W3.T))
    H3a = _aug(H3, KA)
    H3a[:, HID] = 1.0
    Y = (H3a @ W4.T)[:, :3]
    return Y, dict(Xa=Xa, H1=H1, H1a=H1a, H2=H2, H2a=H2a, H3=H3, H3a=H3a, W=(W1, W2, W3, W4))


def loss_grad(params, x, t, norm, mode="bf16"):
    """Mean squared error over n x 3 normalised outputs and its gradient with
    respect to the (augmented) parameters."""
    rnd = bf16 if mode == "bf16" else (lambda a: np.asarray(a, np.float64))
    Y, c = forward(params, x, norm, mode)
    mean = np.asarray(norm["out_mean"], np.float64)
    std = np.asarray(norm["out_std"], np.float64)
    tn = (np.asarray(t, np.float64) - mean) / std
    if mode == "bf16":
        tn = ((np.asarray(t, np.float32) - np.asarray(norm["out_mean"], np.float32))
              / np.asarray(norm["out_std"], np.float32)).astype(np.float64)
    n = Y.shape[0]
    e = Y - tn
    loss = float(np.mean(e * e))
    dY = np.zeros((n, NO))
    dY[:, :3] = rnd(np.float32(2.0 / (3.0 * n)) * e) if mode == "bf16" else 2.0 * e / (3.0 * n)
    W1, W2, W3, W4 = c["W"]
    dW4 = dY.T @ c["H3a"]
    dZ3 = rnd((dY @ W4[:, :HID]) * (1.0 - c["H3"] ** 2))
    dW3 = dZ3.T @ c["H2a"]
    dZ2 = rnd((dZ3 @ W3[:, :HID]) * (1.0 - c["H2"] ** 2))
    dW2 = dZ2.T @ c["H1a"]
    dZ1 = rnd((dZ2 @ W2[:, :HID]) * (1.0 - c["H1"] ** 2))
    dW1 = dZ1.T @ c["Xa"]
    return loss, pack(dW1, dW2, dW3, dW4)


def trainable_mask() -> np.ndarray:
    """1 where a parameter is a weight or bias of the network, 0 on padding."""
    W1, W2, W3, W4 = (np.zeros((HID, K1)), np.zeros((HID, KA)), np.zeros((HID, KA)), np.zeros((NO, KA)))
    W1[:, :6] = 1
    W2[:, :HID + 1] = 1
    W3[:, :HID + 1] = 1
    W4[:3, :HID + 1] = 1
    return pack(W1, W2, W3, W4)
