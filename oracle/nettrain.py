"""ORACLE for a law net's trainer (include/fe_nettrain.h): a NumPy float64 restatement, test infrastructure only
(oracle/__init__.py).

Follows the reference's generated trainer (src/field/cuda/CudaNet.cpp:279-369) as its world model drives it
(src/field/mind/WorldModel.cpp:395-417):
- rows (:291-297): a0 = (x - in_shift) * in_scale; each layer pre = W a + b (:300-305), the hidden layers' activation
  applied (:306-308), the last layer linear;
- loss (:311-315): y = pre_last * out_scale, e = y - t, 0.5 m e^2 per output; the last delta m e out_scale;
- back (:316-325): delta_below = (W^T delta) * act'(pre_below);
- gradients (:331-353): the weight's gradient sum over rows of delta_o * input_i, the bias's sum of delta_o, both
  divided by the rows (:354);
- Adam (:361-366): g clipped to +-clip, m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
  w -= lr (m / c1) / (sqrt(v / c2) + 1e-8) with c1 = 1 - b1^t, c2 = 1 - b2^t (WorldModel.cpp:406-409).
The activations are the law net's (tanh, relu, silu, linear and gelu in its tanh form), their derivatives written from
the definitions. The loss reported is the rows' mean (the card's convention; the reference sums LOSS on the host).
"""
from __future__ import annotations

import numpy as np

ACTS = ("tanh", "relu", "silu", "linear", "gelu")
GELU_C = np.sqrt(2.0 / np.pi)


def act(name: str, x):
    if name == "tanh":
        return np.tanh(x)
    if name == "relu":
        return np.maximum(x, 0.0)
    if name == "silu":
        return x / (1.0 + np.exp(-x))
    if name == "gelu":
        return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + 0.044715 * x**3)))
    return x


def act_grad(name: str, x):
    if name == "tanh":
        return 1.0 - np.tanh(x) ** 2
    if name == "relu":
        return (x > 0).astype(np.float64)
    if name == "silu":
        s = 1.0 / (1.0 + np.exp(-x))
        return s + x * s * (1.0 - s)
    if name == "gelu":
        t = np.tanh(GELU_C * (x + 0.044715 * x**3))
        return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * x * x)
    return np.ones_like(x)


def first_weight(sizes) -> int:
    """The image's first weight: 2 in + 1 floats of normalisation padded to a multiple of 4 (CudaNet.cpp:103)."""
    return (2 * sizes[0] + 1 + 3) // 4 * 4


def layout(sizes):
    """[(W offset, b offset)] per layer in the image (CudaNet.cpp:97-110), and the image's length."""
    at = first_weight(sizes)
    out = []
    for k in range(len(sizes) - 1):
        ni, no = sizes[k], sizes[k + 1]
        out.append((at, at + ni * no))
        at += ni * no + no
    return out, at


def unpack(sizes, image):
    image = np.asarray(image, np.float64)
    n = sizes[0]
    lay, total = layout(sizes)
    assert image.size == total
    Ws = [image[w:b].reshape(sizes[k + 1], sizes[k]) for k, (w, b) in enumerate(lay)]
    bs = [image[b:b + sizes[k + 1]] for k, (w, b) in enumerate(lay)]
    return image[:n], image[n:2 * n], image[2 * n], Ws, bs


def forward(sizes, name, image, x):
    """The net's output y (rows x out)."""
    shift, scale, out_scale, Ws, bs = unpack(sizes, image)
    a = (np.asarray(x, np.float64) - shift) * scale
    for k, (W, b) in enumerate(zip(Ws, bs)):
        pre = a @ W.T + b
        a = act(name, pre) if k + 1 < len(Ws) else pre
    return a * out_scale


def gradient(sizes, name, image, x, t, m=None):
    """(mean loss, gradient of the trained floats in the image's order) for rows x, targets t, row weights m."""
    shift, scale, out_scale, Ws, bs = unpack(sizes, image)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    rows = x.shape[0]
    m = np.ones(rows) if m is None else np.asarray(m, np.float64)
    inputs, pres = [(x - shift) * scale], []
    for k, (W, b) in enumerate(zip(Ws, bs)):
        pre = inputs[-1] @ W.T + b
        pres.append(pre)
        if k + 1 < len(Ws):
            inputs.append(act(name, pre))
    e = pres[-1] * out_scale - t
    loss = 0.5 * (m[:, None] * e * e).sum() / rows
    delta = m[:, None] * e * out_scale
    grads = [None] * len(Ws)
    for k in range(len(Ws) - 1, -1, -1):
        grads[k] = (delta.T @ inputs[k] / rows, delta.sum(axis=0) / rows)
        if k > 0:
            delta = (delta @ Ws[k]) * act_grad(name, pres[k - 1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat


def adam(image, grad, m1, m2, step, sizes, lr, clip, b1=0.9, b2=0.999):
    """One Adam step in place on the trained part of the image (image, m1, m2 float64); step is the new count."""
    g = np.clip(grad, -clip, clip)
    m1[:] = b1 * m1 + (1 - b1) * g
    m2[:] = b2 * m2 + (1 - b2) * g * g
    c1, c2 = 1 - b1**step, 1 - b2**step
    base = first_weight(sizes)
    image[base:] -= lr * (m1 / c1) / (np.sqrt(m2 / c2) + 1e-8)


def make_image(sizes, seed, weight_scale=None, shift=0.1, scale=1.5, out_scale=0.8):
    """A net with nonzero weights of about unit gain, a nontrivial normalisation and out_scale."""
    rng = np.random.default_rng(seed)
    lay, total = layout(sizes)
    image = np.zeros(total)
    n = sizes[0]
    image[:n] = rng.uniform(-shift, shift, n)
    image[n:2 * n] = rng.uniform(0.5 * scale, scale, n)
    image[2 * n] = out_scale
    for k, (w, b) in enumerate(lay):
        ni, no = sizes[k], sizes[k + 1]
        s = weight_scale if weight_scale is not None else 1.0 / np.sqrt(ni)
        image[w:b] = rng.uniform(-s, s, ni * no)
        image[b:b + no] = rng.uniform(-0.1, 0.1, no)
    return image
