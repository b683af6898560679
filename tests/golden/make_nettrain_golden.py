"""Writes tests/golden/nettrain/ref_cases.npz: the net trainer's loss and gradient by the REFERENCE's own CPU net code.

oracle/_ref/ref_nettrain (this project's driver, linked with the reference's objects by make -f oracle/Makefile.ref;
git-ignored) runs net_forward_taped / net_backward row by row as the reference's world model learns on the host, here,
in the build container. Each case's inputs (sizes, activation, normalisation, weights, rows, targets, row weights) and
the reference's outputs (the rows' mean loss, the mean gradient, fp32) are committed as arrays named `<case>_<what>`.
Nothing here runs on the GPU box.

    python tests/golden/make_nettrain_golden.py    (from the repository root)
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_nettrain")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nettrain", "ref_cases.npz")
ACTS = ("tanh", "relu", "silu", "linear")

# (name, sizes, activation, rows, seed): the world model's net shape at a small feature count, each activation the
# reference's net code has, a deep narrow net, a single-layer (linear) map.
CASES = [
    ("world_tanh", [8, 64, 64, 6], "tanh", 96, 1),
    ("relu", [5, 24, 17, 3], "relu", 41, 2),
    ("silu", [7, 33, 2], "silu", 37, 3),
    ("deep_linear", [4, 9, 9, 9, 9, 2], "linear", 29, 4),
    ("single", [6, 3], "tanh", 17, 5),
]


def make_inputs(sizes, rows, seed):
    rng = np.random.default_rng(seed)
    n = sizes[0]
    shift = rng.uniform(-0.2, 0.2, n).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, n).astype(np.float32)
    out_scale = np.float32(rng.uniform(0.5, 1.5))
    weights = []
    for k in range(len(sizes) - 1):
        ni, no = sizes[k], sizes[k + 1]
        s = 1.0 / np.sqrt(ni)
        weights.append(rng.uniform(-s, s, ni * no))
        weights.append(rng.uniform(-0.1, 0.1, no))
    weights = np.concatenate(weights).astype(np.float32)
    x = rng.uniform(-1, 1, (rows, n)).astype(np.float32)
    t = rng.uniform(-1, 1, (rows, sizes[-1])).astype(np.float32)
    m = rng.uniform(0.0, 2.0, rows).astype(np.float32)
    return shift, scale, out_scale, weights, x, t, m


def run(sizes, act, shift, scale, out_scale, weights, x, t, m):
    with tempfile.TemporaryDirectory() as d:
        src, dst = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(src, "wb") as f:
            f.write(np.array([len(sizes), *sizes, ACTS.index(act), x.shape[0]], np.int32).tobytes())
            for a in (shift, scale, np.array([out_scale], np.float32), weights, x, t, m):
                f.write(np.ascontiguousarray(a, np.float32).tobytes())
        subprocess.run([DRIVER, src, dst], check=True)
        raw = open(dst, "rb").read()
    loss = np.frombuffer(raw[:8], np.float64)[0]
    grad = np.frombuffer(raw[8:], np.float32).copy()
    assert grad.size == weights.size
    return loss, grad


def main() -> None:
    arrays = {}
    for name, sizes, act, rows, seed in CASES:
        shift, scale, out_scale, weights, x, t, m = make_inputs(sizes, rows, seed)
        loss, grad = run(sizes, act, shift, scale, out_scale, weights, x, t, m)
        arrays.update({f"{name}_sizes": np.array(sizes, np.int32), f"{name}_act": np.array(ACTS.index(act), np.int32),
                       f"{name}_shift": shift, f"{name}_scale": scale, f"{name}_out_scale": np.array(out_scale, np.float32),
                       f"{name}_weights": weights, f"{name}_x": x, f"{name}_t": t, f"{name}_m": m,
                       f"{name}_loss": np.array(loss), f"{name}_grad": grad})
        print(name, sizes, act, rows, "loss", loss)
    np.savez_compressed(OUT, **arrays)


if __name__ == "__main__":
    main()
