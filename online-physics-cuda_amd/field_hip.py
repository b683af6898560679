"""Python host side of the Field Engine's MI355X backend: ctypes over libfieldhip.so.

The library is the product (C-ABI runtime over HIP/hipRTC, the law front end, the --hip generator, the HIP backend and
the handwritten gfx950 kernels); this module only binds it. Importing it never falls back to anything: without the
built library it raises, and every compute entry point raises when the library reports an error.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfieldhip.so")

_lib: Optional[ctypes.CDLL] = None


class FieldHipError(RuntimeError):
    pass


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("ok", ctypes.c_int),
        ("why", ctypes.c_char * 512),
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int),
        ("hbm_bytes", ctypes.c_size_t),
        ("lds_bytes", ctypes.c_int),
        ("runtime_version", ctypes.c_int),
    ]


def _declare(lib: ctypes.CDLL) -> None:
    c_int, c_void_p, c_size_t, c_double, c_float = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_float
    sig = {
        "fe_hip_device_open": (c_int, [ctypes.POINTER(DeviceInfo)]),
        "fe_hip_unavailable": (ctypes.c_char_p, []),
        "fe_hip_error_string": (ctypes.c_char_p, [c_int]),
        "fe_hip_stream": (c_void_p, []),
        "fe_hip_mem_alloc": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
        "fe_hip_mem_free": (c_int, [c_void_p]),
        "fe_hip_memcpy_htod_async": (c_int, [c_void_p, c_void_p, c_size_t]),
        "fe_hip_memcpy_dtoh_async": (c_int, [c_void_p, c_void_p, c_size_t]),
        "fe_hip_memcpy_dtod_async": (c_int, [c_void_p, c_void_p, c_size_t]),
        "fe_hip_memset_d32_async": (c_int, [c_void_p, ctypes.c_uint32, c_size_t]),
        "fe_hip_stream_sync": (c_int, []),
        "fe_hip_event_create": (c_int, [ctypes.POINTER(c_void_p)]),
        "fe_hip_event_destroy": (c_int, [c_void_p]),
        "fe_hip_event_record": (c_int, [c_void_p]),
        "fe_hip_event_sync": (c_int, [c_void_p]),
        "fe_hip_event_elapsed": (c_int, [c_void_p, c_void_p, ctypes.POINTER(c_float)]),
        "fe_hip_to_bf16": (ctypes.c_uint16, [c_float]),
        "fe_heat_handwritten": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, c_double]),
        "fe_heat_handwritten_cfg": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, c_double,
                                            c_int, c_int, c_int, c_int]),
        "fe_copy_f32": (c_int, [c_void_p, c_void_p, c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib() -> ctypes.CDLL:
    """The loaded library. Raises when it is not built: there is no other path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FieldHipError(f"{LIB_PATH} is not built: run `make -C online-physics-cuda_amd` (or __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().fe_hip_error_string(status).decode() if status > 0 else "invalid arguments"
        raise FieldHipError(f"{what}: {msg} ({status})")


def device_info() -> DeviceInfo:
    info = DeviceInfo()
    lib().fe_hip_device_open(ctypes.byref(info))
    return info


def require_device() -> DeviceInfo:
    info = device_info()
    if not info.ok:
        raise FieldHipError("no HIP device: " + info.why.decode())
    return info


def sync() -> None:
    check(lib().fe_hip_stream_sync(), "stream sync")


class DeviceArray:
    """A block of device memory holding `count` items of `dtype`, freed with the object."""

    def __init__(self, count: int, dtype=np.float32):
        require_device()
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        ptr = ctypes.c_void_p()
        check(lib().fe_hip_mem_alloc(ctypes.byref(ptr), self.nbytes), "device alloc")
        self.ptr = ptr.value

    @classmethod
    def from_host(cls, values: np.ndarray) -> "DeviceArray":
        values = np.ascontiguousarray(values)
        arr = cls(values.size, values.dtype)
        arr.upload(values)
        return arr

    def upload(self, values: np.ndarray) -> None:
        values = np.ascontiguousarray(values, dtype=self.dtype)
        assert values.size == self.count
        check(lib().fe_hip_memcpy_htod_async(self.ptr, values.ctypes.data, self.nbytes), "upload")
        sync()

    def download(self) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        check(lib().fe_hip_memcpy_dtoh_async(out.ctypes.data, self.ptr, self.nbytes), "download")
        sync()
        return out

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().fe_hip_mem_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Events:
    """A pair of events on the library's stream: the card's time between begin() and end()."""

    def __init__(self):
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().fe_hip_event_create(ctypes.byref(a)), "event")
        check(lib().fe_hip_event_create(ctypes.byref(b)), "event")
        self.a, self.b = a.value, b.value

    def begin(self) -> None:
        check(lib().fe_hip_event_record(self.a), "event record")

    def end(self) -> float:
        check(lib().fe_hip_event_record(self.b), "event record")
        check(lib().fe_hip_event_sync(self.b), "event sync")
        ms = ctypes.c_float()
        check(lib().fe_hip_event_elapsed(self.a, self.b, ctypes.byref(ms)), "event elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            lib().fe_hip_event_destroy(self.a)
            lib().fe_hip_event_destroy(self.b)
        except Exception:
            pass


def heat_handwritten(src: DeviceArray, dst: DeviceArray, nx: int, ny: int, nz: int, k: float, h: float, sub_dt: float,
                     ty: Optional[int] = None, zchunk: Optional[int] = None, remap: int = 1, variant: int = 0) -> None:
    """One substep of the handwritten 7-point conduction kernel, src -> dst (queued on the library's stream).

    variant 0 picks (four cells a lane when nx % 4 == 0), 1 a lane a column, 2 four cells a lane."""
    n = nx * ny * nz
    if src.count != n or dst.count != n or src.dtype != np.float32 or dst.dtype != np.float32:
        raise FieldHipError("heat_handwritten: arrays must hold nx*ny*nz floats")
    if ty is None and zchunk is None and variant == 0:
        check(lib().fe_heat_handwritten(src.ptr, dst.ptr, nx, ny, nz, k, h, sub_dt), "heat_handwritten")
        return
    x4 = variant == 2 or (variant == 0 and nx % 4 == 0)
    ty = ty if ty is not None else (6 if x4 else 4)
    zchunk = zchunk if zchunk is not None else (8 if x4 else 32)
    check(lib().fe_heat_handwritten_cfg(src.ptr, dst.ptr, nx, ny, nz, k, h, sub_dt, ty, zchunk, remap,
                                        2 if x4 else 1), "heat_handwritten")


def copy_f32(src: DeviceArray, dst: DeviceArray) -> None:
    if src.count != dst.count:
        raise FieldHipError("copy_f32: sizes differ")
    check(lib().fe_copy_f32(src.ptr, dst.ptr, src.count), "copy_f32")
