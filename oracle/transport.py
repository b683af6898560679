"""ctypes binding of the ORACLE's FP64 three-field transport law (transport_ref.c). Test infrastructure only."""
from __future__ import annotations

import ctypes

import numpy as np

from .heat import lib as _lib

_ready = False

LAW = """# a, b, c: diffusion (one diffusivity each), upwind advection at (0.3, -0.2, 0.1), a + b <-> c
field a init 0
field b init 0
field c init 0
param Da = 0.05
param Db = 0.02
param Dc = 0.01
term transport
  d a = diffuse(Da, a) - 0.3 * (a - shift(a, -1, 0, 0)) / h + 0.2 * (shift(a, 0, 1, 0) - a) / h - 0.1 * (a - shift(a, 0, 0, -1)) / h - a * b + 0.1 * c
  d b = diffuse(Db, b) - 0.3 * (b - shift(b, -1, 0, 0)) / h + 0.2 * (shift(b, 0, 1, 0) - b) / h - 0.1 * (b - shift(b, 0, 0, -1)) / h - a * b + 0.1 * c
  d c = diffuse(Dc, c) - 0.3 * (c - shift(c, -1, 0, 0)) / h + 0.2 * (shift(c, 0, 1, 0) - c) / h - 0.1 * (c - shift(c, 0, 0, -1)) / h + a * b - 0.1 * c
"""


def loss(shape, h: float, dt: float, steps: int, D, a, b, c) -> float:
    """sum c^2 after `steps` FP64 steps of the law from (a, b, c)."""
    global _ready
    L = _lib()
    if not _ready:
        P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.oracle_transport_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           P, P, P, P]
        L.oracle_transport_run.restype = ctypes.c_double
        _ready = True
    nx, ny, nz = shape
    A, B, C = (np.array(v, dtype=np.float64, copy=True).reshape(-1) for v in (a, b, c))
    return float(L.oracle_transport_run(nx, ny, nz, h, dt, steps, np.ascontiguousarray(D, dtype=np.float64), A, B, C))
