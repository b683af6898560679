"""ctypes binding of the ORACLE's FP64 diffusion-reaction run (diffreact_ref.c). Test infrastructure only."""
from __future__ import annotations

import ctypes

import numpy as np

from .heat import lib as _lib

_ready = False


def run64(u, v, nx: int, ny: int, count: int, h: float, du: float, dv: float, k: float, dt: float, steps: int):
    """`steps` explicit steps of the plain law in FP64 from (u, v); returns the final (u, v)."""
    global _ready
    L = _lib()
    if not _ready:
        D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        c_int, c_double = ctypes.c_int, ctypes.c_double
        L.oracle_dr_run_f64.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_int, D, D]
        _ready = True
    U = np.array(u, dtype=np.float64, copy=True).reshape(-1)
    V = np.array(v, dtype=np.float64, copy=True).reshape(-1)
    L.oracle_dr_run_f64(nx, ny, count, h, du, dv, k, dt, steps, U, V)
    return U, V
