"""ctypes binding of the ORACLE's C restatement of the reference CPU solver for `d T = diffuse(k, T)` (heat_ref.c).

Test infrastructure only (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        _lib = ctypes.CDLL(path)
        f = ctypes.c_float
        P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        _lib.oracle_heat_substep.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double]
        _lib.oracle_heat_substep.restype = ctypes.c_int
        _lib.oracle_heat_run.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _lib.oracle_heat_run.restype = ctypes.c_int
        _lib.oracle_heat_substeps.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _lib.oracle_heat_substeps.restype = ctypes.c_int
        _lib.oracle_total.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        del f
    return _lib


def substep(t: np.ndarray, shape, k: float, h: float, sub_dt: float) -> np.ndarray:
    nx, ny, nz = shape
    t = np.ascontiguousarray(t, dtype=np.float32)
    out = np.empty_like(t)
    lib().oracle_heat_substep(t, out, nx, ny, nz, k, h, sub_dt)
    return out


def substeps(dt: float, coeff: float, k: float, h: float) -> int:
    return lib().oracle_heat_substeps(dt, coeff, k, h)


def run(t: np.ndarray, shape, k: float, h: float, dt: float, steps: int, coeff: float = 0.15) -> np.ndarray:
    nx, ny, nz = shape
    t = np.array(t, dtype=np.float32, copy=True)
    scratch = np.empty_like(t)
    lib().oracle_heat_run(t, scratch, nx, ny, nz, k, h, dt, coeff, steps)
    return t


def total(v: np.ndarray):
    s, m = ctypes.c_double(), ctypes.c_double()
    v = np.ascontiguousarray(v, dtype=np.float32)
    lib().oracle_total(v, v.size, ctypes.byref(s), ctypes.byref(m))
    return s.value, m.value
