"""CPU restatement of the warm-start MLP (test infrastructure only; the product
runs photara_drender_amd/csrc/mlp.hip).

The warm start is project-specified (no reference code): per texel, features of
the 8 highest-confidence views, [R, G, B, confidence, cos(normal, view)] each
(40 floats, zero for missing views), go through 40 -> 64 -> 64 -> 3 with GELU
(tanh form) on both hidden layers; the output is added to the confidence-
weighted mean colour of the 8 samples (0 when no sample).

Parameters, one flat float32 vector of PARAMS values:
    W1 [40, 64] | b1 [64] | W2 [64, 64] | b2 [64] | W3 [64, 3] | b3 [3]
Loss over a batch: sum over valid texels and channels of (y - target)^2 divided
by 3 * max(number of valid texels, 1).  Texels without samples (confidence sum
0) predict 0, the bake's colour for them.

``bf16=True`` rounds to bfloat16 (round-to-nearest-even) exactly where the GPU
kernels feed matrix cores: features, weights, hidden activations, and in the
backward the unscaled output gradient 2 (y - target) and the pre-activation
gradients derived from it; the 1 / (3 n) scale is applied to the summed
gradients.  Everything else is float64 here (float32 accumulation on the GPU).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, VIEWS = 40, 64, 3, 8
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)), ("b3", (OUT,))]
PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 6979
GELU_C = np.sqrt(2.0 / np.pi)


def unpack(theta):
    theta = np.asarray(theta, np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(a):
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(f)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(nan, np.nan, out)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * z * z)


def base_color(x):
    """Confidence-weighted mean colour of the 8 samples."""
    f = np.asarray(x, np.float64).reshape(-1, VIEWS, 5)
    w = f[:, :, 3]
    s = w.sum(1)
    num = (f[:, :, :3] * w[:, :, None]).sum(1)
    return np.where(s[:, None] > 0, num / np.where(s > 0, s, 1.0)[:, None], 0.0)


def forward(theta, x, bf=True):
    p = unpack(theta)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    x = np.asarray(x, np.float64)
    z1 = r(x) @ r(p["W1"]) + p["b1"]
    h1 = gelu(z1)
    z2 = r(h1) @ r(p["W2"]) + p["b2"]
    h2 = gelu(z2)
    y = r(h2) @ r(p["W3"]) + p["b3"] + base_color(x)
    return y, (x, z1, h1, z2, h2)


def predict(theta, x, bf=True):
    """Warm-start colour: network output plus base colour, 0 for texels without samples."""
    y, _ = forward(theta, x, bf)
    s = np.asarray(x, np.float64).reshape(-1, VIEWS, 5)[:, :, 3].sum(1)
    return np.where(s[:, None] > 0, y, 0.0)


def loss_and_grad(theta, x, target, valid=None, bf=True):
    """Loss and its gradient w.r.t. the flat parameters (and the predictions)."""
    p = unpack(theta)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, (x, z1, h1, z2, h2) = forward(theta, x, bf)
    v = np.ones((y.shape[0], 1)) if valid is None else np.asarray(valid, np.float64).reshape(-1, 1)
    inv = 1.0 / (3.0 * max(v.sum(), 1.0))
    d = (y - np.asarray(target, np.float64)) * v
    loss = float((d * d).sum() * inv)
    dy = 2.0 * d  # unscaled output gradient
    g = {}
    g["W3"] = r(h2).T @ r(dy)
    g["b3"] = dy.sum(0)
    dz2 = (r(dy) @ r(p["W3"]).T) * gelu_grad(z2)
    g["W2"] = r(h1).T @ r(dz2)
    g["b2"] = dz2.sum(0)
    dz1 = (r(dz2) @ r(p["W2"]).T) * gelu_grad(z1)
    g["W1"] = r(x).T @ r(dz1)
    g["b1"] = dz1.sum(0)
    return loss, pack(g) * inv, y


def init_params(seed=0):
    """He-style initialisation with a zero output layer (the network starts as the base colour)."""
    rng = np.random.default_rng(seed)
    parts = {
        "W1": rng.normal(0, np.sqrt(2.0 / IN), (IN, HID)), "b1": np.zeros(HID),
        "W2": rng.normal(0, np.sqrt(2.0 / HID), (HID, HID)), "b2": np.zeros(HID),
        "W3": np.zeros((HID, OUT)), "b3": np.zeros(OUT),
    }
    return pack(parts).astype(np.float32)


def adam(theta, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1), float64; returns new (theta, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
