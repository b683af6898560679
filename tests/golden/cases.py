"""The golden cases: law texts written for this project, grids, steps and seeded initial states.

make_golden.py runs each through the reference's CPU engine and stores inputs and outputs as tests/golden/NAME.npz.
The tests import this module for nothing but the list of names; everything they check comes from the .npz files.
"""
from __future__ import annotations

import numpy as np

HEAT_LAW = """# 3-D conduction with one conductivity: the 7-point flux-form step
field T init 0 range -1000000 1000000

param k = 0.1

term conduction
  d T = diffuse(k, T)
  conserves T
  max_dt = 0.15 * h * h / k
"""


def _uniform(shape, seed, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    return (lo + (hi - lo) * rng.random(int(np.prod(shape)))).astype(np.float32)


def _start_value(n):
    # A deterministic spread of values 290..1290, as the reference's stencil bench starts its lattice.
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = i * np.uint64(0x9E3779B97F4A7C15)
        h ^= h >> np.uint64(31)
        h *= np.uint64(0x94D049BB133111EB)
        h ^= h >> np.uint64(29)
    return (290.0 + 1000.0 * (h >> np.uint64(40)).astype(np.float32) / np.float32(16777216.0)).astype(np.float32)


CASES = [
    # Uneven sizes on every axis, three substeps a step (dt = 0.9 over max_dt = 0.375).
    dict(name="heat_20x14x9", law=HEAT_LAW, grid=(20, 14, 9, 0.5), dt=0.9, steps=4,
         params={"k": 0.1}, init={"T": _uniform((20, 14, 9), 1)}),
    # k = h = 1, one substep: the configuration the reference's own bit-identity check uses.
    dict(name="heat_16_k1", law=HEAT_LAW, grid=(16, 16, 16, 1.0), dt=0.1, steps=5,
         params={"k": 1.0}, init={"T": _start_value(16 * 16 * 16)}),
    # The conduction configuration at a small size: h = 1/32, k = 0.1, dt = 0.4 h^2 / k (three substeps).
    dict(name="heat_32_cfg", law=HEAT_LAW, grid=(32, 32, 32, 1.0 / 32), dt=0.4 * (1.0 / 32) ** 2 / 0.1, steps=6,
         params={"k": 0.1}, init={"T": _uniform((32, 32, 32), 7)}),
    # A flat lattice, one cell thick: the z faces never exist.
    dict(name="heat_64x48x1", law=HEAT_LAW, grid=(64, 48, 1, 1.0), dt=0.1, steps=7,
         params={"k": 0.7}, init={"T": _uniform((64, 48, 1), 3, -2.0, 5.0)}),
]
