"""ORACLE (test infrastructure only): the learned flux correction and the plain diffusion-reaction law, in NumPy float64.

Restates what online-physics-cuda_amd/csrc/kernels/learned_flux.hip computes, with bf16 rounding (round to nearest
even) at the same points as the card (inputs, weights, hidden activations and their gradients as they enter a matrix
product), so the card and this differ only by fp32 accumulation. Also: the plain law (5-point Laplacian, zero-flux
edges) stepped in FP64 (what the learned coarse lattice's error is measured against), and block averages.

Only tests/, the smoke check and bench.py call this.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 18, 64, 2
SQ2PI = 0.7978845608028654


def split(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
                        ("W3", (N_HID, N_OUT)), ("b3", (N_OUT,))):
        size = int(np.prod(shape))
        out[name] = p[o:o + size].reshape(shape)
        o += size
    assert o == p.size == 5506
    return out


ROUND = True  # False: no bf16 rounding anywhere (for checking the backward pass against finite differences)


def bf16(x):
    """Round to bf16 (nearest even), back to float64 (identity when ROUND is off)."""
    if not ROUND:
        return np.asarray(x, np.float64)
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(SQ2PI * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(SQ2PI * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * SQ2PI * (1.0 + 3 * 0.044715 * z * z)


def init_params(seed: int, last_scale: float = 0.0):
    """Glorot-uniform weights, zero biases; the last layer scaled by last_scale (0: the net starts at no correction)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out, scale in ((N_IN, N_HID, 1.0), (N_HID, N_HID, 1.0), (N_HID, N_OUT, last_scale)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        parts.append((rng.uniform(-lim, lim, (fan_in, fan_out)) * scale).reshape(-1))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def patches(u, v, nx, ny, count):
    """Inputs of every evaluation row, shape (count * nx * ny * 2 directions, 2 sides, 18): in = field * 9 + (a + 1) * 3 +
    (b + 1), a along the face's direction, b across; side 1 is the cell across the face, read mirrored (a -> -a)."""
    U = np.asarray(u, np.float64).reshape(count, ny, nx)
    V = np.asarray(v, np.float64).reshape(count, ny, nx)
    ys, xs = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((count, ny, nx, 2, 2, N_IN))
    for d in range(2):
        for side in range(2):
            ax = xs + (1 if (side and d == 0) else 0)
            ay = ys + (1 if (side and d == 1) else 0)
            for fi, F in enumerate((U, V)):
                for a in (-1, 0, 1):
                    for b in (-1, 0, 1):
                        da = -a if side else a
                        xx = ax + (da if d == 0 else b)
                        yy = ay + (b if d == 0 else da)
                        xx = np.clip(xx, 0, nx - 1)
                        yy = np.clip(yy, 0, ny - 1)
                        out[:, :, :, d, side, fi * 9 + (a + 1) * 3 + (b + 1)] = F[:, yy, xx]
    return out.reshape(-1, 2, N_IN)  # row index = (lattice, y, x, dir)


def forward(P, X):
    """Rows of 18 inputs -> (outputs, what backward needs), bf16 where the card rounds."""
    W1, W2, W3 = bf16(P["W1"]), bf16(P["W2"]), bf16(P["W3"])
    x = bf16(X)
    z1 = x @ W1 + P["b1"]
    h1 = bf16(gelu(z1))
    z2 = h1 @ W2 + P["b2"]
    h2 = bf16(gelu(z2))
    y = h2 @ W3 + P["b3"]
    return y, (x, z1, h1, z2, h2, W2, W3)


def open_faces(nx, ny, count):
    ys, xs = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    o = np.stack([xs + 1 < nx, ys + 1 < ny], -1)  # (ny, nx, 2)
    return np.broadcast_to(o, (count, ny, nx, 2)).reshape(-1)


def faces(p, u, v, nx, ny, count, scale):
    """F[(cell, dir), field] as the card's fe_lnet_faces."""
    P = split(p)
    X = patches(u, v, nx, ny, count)
    yd, _ = forward(P, X[:, 0])
    ym, _ = forward(P, X[:, 1])
    F = 0.5 * scale * (yd - ym)
    F[~open_faces(nx, ny, count)] = 0.0
    return F


def correction(F, nx, ny, count, h):
    F = F.reshape(count, ny, nx, 2, 2)  # lattice, y, x, dir, field
    c = -F[:, :, :, 0] - F[:, :, :, 1]
    c[:, :, 1:] += F[:, :, :-1, 0]
    c[:, 1:, :] += F[:, :-1, :, 1]
    return c / h  # (count, ny, nx, field)


def plain(u, v, nx, ny, count, h, du, dv, k):
    U = np.asarray(u, np.float64).reshape(count, ny, nx)
    V = np.asarray(v, np.float64).reshape(count, ny, nx)

    def lap(F):
        Fp = np.pad(F, ((0, 0), (1, 1), (1, 1)), mode="edge")
        return (Fp[:, 1:-1, :-2] + Fp[:, 1:-1, 2:] + Fp[:, :-2, 1:-1] + Fp[:, 2:, 1:-1] - 4 * F) / (h * h)

    return du * lap(U) + U - U ** 3 - k - V, dv * lap(V) + U - V


def step(p, u, v, nx, ny, count, h, du, dv, k, dt, scale, learned=True):
    ru, rv = plain(u, v, nx, ny, count, h, du, dv, k)
    if learned:
        c = correction(faces(p, u, v, nx, ny, count, scale), nx, ny, count, h)
        ru = ru + c[..., 0]
        rv = rv + c[..., 1]
    U = np.asarray(u, np.float64).reshape(count, ny, nx)
    V = np.asarray(v, np.float64).reshape(count, ny, nx)
    return (U + dt * ru).reshape(-1), (V + dt * rv).reshape(-1)


def loss_and_grad(p, u0, v0, u1, v1, nx, ny, count, h, du, dv, k, dt, scale):
    """The training loss and its gradient w.r.t. the 5506 parameters, as the card's fe_lnet_gradient."""
    P = split(p)
    X = patches(u0, v0, nx, ny, count)
    n = count * nx * ny
    yd, cd = forward(P, X[:, 0])
    ym, cm = forward(P, X[:, 1])
    op = open_faces(nx, ny, count)
    F = 0.5 * scale * (yd - ym)
    F[~op] = 0.0
    c = correction(F, nx, ny, count, h)
    ru, rv = plain(u0, v0, nx, ny, count, h, du, dv, k)
    U0, V0 = np.asarray(u0, np.float64).reshape(count, ny, nx), np.asarray(v0, np.float64).reshape(count, ny, nx)
    U1, V1 = np.asarray(u1, np.float64).reshape(count, ny, nx), np.asarray(v1, np.float64).reshape(count, ny, nx)
    tu, tv = (U1 - U0) / dt - ru, (V1 - V0) / dt - rv
    gap = np.stack([c[..., 0] - tu, c[..., 1] - tv], -1)  # (count, ny, nx, 2)
    loss = float(np.sum(gap ** 2) / (2 * n))
    E = 2 * gap / (2 * n)
    # dL/dF of the +x face of (y, x): (E[y, x+1] - E[y, x]) / h; +y face: (E[y+1, x] - E[y, x]) / h
    dF = np.zeros((count, ny, nx, 2, 2))
    dF[:, :, :-1, 0] = (E[:, :, 1:] - E[:, :, :-1]) / h
    dF[:, :-1, :, 1] = (E[:, 1:, :] - E[:, :-1, :]) / h
    dF = dF.reshape(-1, 2)
    dF[~op] = 0.0
    grads = np.zeros(5506)
    g = split(grads)  # views into grads
    for side, cache, sign in ((0, cd, 1.0), (1, cm, -1.0)):
        x, z1, h1, z2, h2, W2, W3 = cache
        dy = bf16(sign * 0.5 * scale * dF)
        dz2 = bf16((dy @ W3.T) * gelu_grad(z2))
        dz1 = bf16((dz2 @ W2.T) * gelu_grad(z1))
        g["W3"][...] += h2.T @ dy
        g["b3"][...] += dy.sum(0)
        g["W2"][...] += h1.T @ dz2
        g["b2"][...] += dz2.sum(0)
        g["W1"][...] += x.T @ dz1
        g["b1"][...] += dz1.sum(0)
    return loss, grads


def run(u, v, nx, ny, count, h, du, dv, k, dt, steps):
    """The plain law in FP64, `steps` steps of dt."""
    U = np.asarray(u, np.float64).copy()
    V = np.asarray(v, np.float64).copy()
    for _ in range(steps):
        U, V = step(None, U, V, nx, ny, count, h, du, dv, k, dt, 1.0, learned=False)
    return U, V


def coarsen(f, nx, ny, count, factor):
    F = np.asarray(f, np.float64).reshape(count, ny // factor, factor, nx // factor, factor)
    return F.mean(axis=(2, 4)).reshape(-1)
