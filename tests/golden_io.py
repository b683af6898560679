"""Reads the golden vectors (tests/golden/*.npz) the reference's CPU engine wrote."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    init = {k[5:]: z[k] for k in z.files if k.startswith("init_")}
    final = {k[6:]: z[k] for k in z.files if k.startswith("final_")}
    return meta, init, final


def names(prefix: str = ""):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz") and f.startswith(prefix))
