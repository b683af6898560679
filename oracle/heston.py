"""ctypes binding of the ORACLE's Heston ADI restatement (heston_ref.c): one option a lattice, FP64 or FP32.

Test infrastructure only (see oracle/__init__.py). An option is 8 numbers: kappa, theta, sigma, rho, r, strike,
maturity, v0. Its lattice is ny rows of nx cells (x = ln S fastest), cell h = 6 / nx.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .heat import lib as _heat_lib

_ready = False


def lib() -> ctypes.CDLL:
    global _ready
    L = _heat_lib()
    if not _ready:
        D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        c_int, c_double = ctypes.c_int, ctypes.c_double
        L.oracle_heston_payoff.argtypes = [c_int, c_int, D, D]
        L.oracle_heston_price.argtypes = [c_int, c_int, D, D]
        L.oracle_heston_price.restype = c_double
        L.oracle_heston_stencil.argtypes = [c_int, c_int, D, ctypes.POINTER(c_int), ctypes.POINTER(c_int), D, D]
        L.oracle_heston_run_f64.argtypes = [c_int, c_int, D, D, c_int, c_int, c_double]
        L.oracle_heston_run_f32.argtypes = [c_int, c_int, D, F, c_int, c_int, c_double]
        L.oracle_heston_batch_f64.argtypes = [c_int, I, I, D, L64, D, c_int, c_int, c_double, ctypes.c_void_p]
        L.oracle_heston_batch_f32.argtypes = [c_int, I, I, D, L64, F, c_int, c_int, c_double]
        L.oracle_heston_thomas.argtypes = [c_int, c_double, c_double, c_double, D, D]
        L.oracle_heston_apply_f.argtypes = [c_int, c_int, D, D, D]
        _ready = True
    return L


CS, HV = 0, 1
CS_THETA = 0.5
HV_THETA = 0.5 + np.sqrt(3.0) / 6.0


def _p(option) -> np.ndarray:
    p = np.ascontiguousarray(option, dtype=np.float64)
    assert p.shape == (8,)
    return p


def payoff(nx: int, ny: int, option) -> np.ndarray:
    out = np.empty(nx * ny, np.float64)
    lib().oracle_heston_payoff(nx, ny, _p(option), out)
    return out


def run64(nx: int, ny: int, option, V: np.ndarray, steps: int, scheme: int = CS, theta: float = CS_THETA) -> np.ndarray:
    V = np.array(V, dtype=np.float64, copy=True)
    lib().oracle_heston_run_f64(nx, ny, _p(option), V, steps, scheme, theta)
    return V


def run32(nx: int, ny: int, option, V: np.ndarray, steps: int, scheme: int = CS, theta: float = CS_THETA) -> np.ndarray:
    V = np.array(V, dtype=np.float32, copy=True)
    lib().oracle_heston_run_f32(nx, ny, _p(option), V, steps, scheme, theta)
    return V


def price(nx: int, ny: int, option, V: np.ndarray) -> float:
    return float(lib().oracle_heston_price(nx, ny, _p(option), np.ascontiguousarray(V, dtype=np.float64)))


def stencil(nx: int, ny: int, option):
    i0, j0 = ctypes.c_int(), ctypes.c_int()
    wx, wy = np.empty(4), np.empty(4)
    lib().oracle_heston_stencil(nx, ny, _p(option), ctypes.byref(i0), ctypes.byref(j0), wx, wy)
    return i0.value, j0.value, wx, wy


def batch64(nx, ny, options, steps: int, scheme: int = CS, theta: float = CS_THETA):
    """Every option from its payoff: (prices, lattices back to back, offsets)."""
    nx = np.ascontiguousarray(nx, dtype=np.int32)
    ny = np.ascontiguousarray(ny, dtype=np.int32)
    opts = np.ascontiguousarray(options, dtype=np.float64).reshape(-1, 8)
    cells = nx.astype(np.int64) * ny
    offsets = np.concatenate([[0], np.cumsum(cells)[:-1]]).astype(np.int64)
    V = np.concatenate([payoff(int(a), int(b), o) for a, b, o in zip(nx, ny, opts)])
    prices = np.empty(len(nx), np.float64)
    lib().oracle_heston_batch_f64(len(nx), nx, ny, opts.reshape(-1), offsets, V, steps, scheme, theta,
                                  prices.ctypes.data_as(ctypes.c_void_p))
    return prices, V, offsets


def batch32(nx, ny, options, steps: int, scheme: int = CS, theta: float = CS_THETA):
    nx = np.ascontiguousarray(nx, dtype=np.int32)
    ny = np.ascontiguousarray(ny, dtype=np.int32)
    opts = np.ascontiguousarray(options, dtype=np.float64).reshape(-1, 8)
    cells = nx.astype(np.int64) * ny
    offsets = np.concatenate([[0], np.cumsum(cells)[:-1]]).astype(np.int64)
    V = np.concatenate([payoff(int(a), int(b), o) for a, b, o in zip(nx, ny, opts)]).astype(np.float32)
    lib().oracle_heston_batch_f32(len(nx), nx, ny, opts.reshape(-1), offsets, V, steps, scheme, theta)
    return V, offsets


def thomas(a: float, b: float, c: float, d: np.ndarray) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.float64)
    x = np.empty_like(d)
    lib().oracle_heston_thomas(len(d), a, b, c, d, x)
    return x


def apply_f(nx: int, ny: int, option, W: np.ndarray) -> np.ndarray:
    out = np.empty(nx * ny, np.float64)
    lib().oracle_heston_apply_f(nx, ny, _p(option), np.ascontiguousarray(W, dtype=np.float64), out)
    return out
