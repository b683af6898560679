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
        "fe_roof_f32": (c_int, [c_void_p, c_void_p, c_size_t, c_int, c_int, c_int, c_int]),
        "fe_hip_compile": (c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), c_int,
                                   ctypes.POINTER(c_void_p), ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p),
                                   ctypes.POINTER(c_int)]),
        "fe_hip_free": (None, [c_void_p]),
        "fe_law_compile": (c_void_p, [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p), c_int,
                                      ctypes.c_char_p, c_size_t]),
        "fe_law_free": (None, [c_void_p]),
        "fe_law_field_count": (c_int, [c_void_p]),
        "fe_law_field_id": (c_int, [c_void_p, ctypes.c_char_p]),
        "fe_law_field_name": (ctypes.c_char_p, [c_void_p, c_int]),
        "fe_law_param_count": (c_int, [c_void_p]),
        "fe_law_term_count": (c_int, [c_void_p]),
        "fe_law_ir_json": (c_void_p, [c_void_p]),
        "fe_law_describe": (c_void_p, [c_void_p]),
        "fe_law_text": (c_void_p, [c_void_p]),
        "fe_law_hip_source": (c_void_p, [c_void_p, ctypes.c_char_p, c_size_t]),
        "fe_string_free": (None, [c_void_p]),
        "fe_solver_create": (c_void_p, [c_void_p, c_int, c_int, c_int, c_double, ctypes.c_char_p, c_size_t]),
        "fe_solver_free": (None, [c_void_p]),
        "fe_solver_set_field": (c_int, [c_void_p, c_int, c_void_p]),
        "fe_solver_get_field": (c_int, [c_void_p, c_int, c_void_p]),
        "fe_solver_set_param": (c_int, [c_void_p, ctypes.c_char_p, c_double]),
        "fe_solver_step": (c_int, [c_void_p, c_double]),
        "fe_solver_run": (c_int, [c_void_p, c_double, c_int]),
        "fe_solver_shift_window": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p]),
        "fe_solver_fill_field": (c_int, [c_void_p, c_int, ctypes.c_float]),
        "fe_solver_set_net_weights": (c_int, [c_void_p, ctypes.c_char_p, c_void_p, ctypes.c_size_t]),
        "fe_solver_total_accumulate": (c_int, [c_void_p, c_int, c_void_p]),
        "fe_net_half_forward": (c_int, [c_void_p, c_int, c_int, ctypes.c_float, c_void_p, c_void_p, ctypes.c_longlong, c_void_p]),
        "fe_solver_refresh_derived": (c_int, [c_void_p]),
        "fe_solver_set_lazy_readback": (c_int, [c_void_p, c_int]),
        "fe_solver_check_contracts_every": (c_int, [c_void_p, c_int]),
        "fe_solver_set_learned": (c_int, [c_void_p, c_int]),
        "fe_solver_set_respect_max_dt": (c_int, [c_void_p, c_int]),
        "fe_solver_violations_json": (c_void_p, [c_void_p]),
        "fe_solver_last_substeps": (c_int, [c_void_p, ctypes.POINTER(c_int), c_int]),
        "fe_solver_last_changed": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_ubyte), c_int]),
        "fe_solver_profile_json": (c_void_p, [c_void_p]),
        "fe_solver_digest": (ctypes.c_uint64, [c_void_p]),
        "fe_solver_error": (ctypes.c_char_p, [c_void_p]),
        "fe_solver_field_device": (c_void_p, [c_void_p, c_int]),
        "fe_solver_set_field_device": (c_int, [c_void_p, c_int, c_void_p]),
        "fe_solver_bind_field_device": (c_int, [c_void_p, c_int, c_void_p]),
        "fe_solver_total": (c_int, [c_void_p, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_double)]),
        "fe_law_adjoint": (c_void_p, [c_void_p, ctypes.c_char_p, ctypes.c_char_p, c_size_t]),
        "fe_heston_create": (c_void_p, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_double, ctypes.c_char_p,
                                        c_size_t]),
        "fe_heston_free": (None, [c_void_p]),
        "fe_heston_error": (ctypes.c_char_p, [c_void_p]),
        "fe_heston_cells": (ctypes.c_longlong, [c_void_p]),
        "fe_heston_reset": (c_int, [c_void_p]),
        "fe_heston_step": (c_int, [c_void_p, c_int]),
        "fe_heston_capture": (c_int, [c_void_p, c_int]),
        "fe_heston_launch": (c_int, [c_void_p]),
        "fe_heston_sync": (c_int, [c_void_p]),
        "fe_heston_prices": (c_int, [c_void_p, c_void_p]),
        "fe_heston_lattice": (c_int, [c_void_p, c_int, c_void_p]),
        "fe_heston_classes": (c_int, [c_void_p]),
        "fe_heston_get_shape": (c_int, [c_void_p, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                        ctypes.POINTER(c_int)]),
        "fe_heston_set_shape": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
        "fe_heston_tune": (c_int, [c_void_p, c_int, ctypes.c_char_p, c_size_t]),
        "fe_lnet_create": (c_void_p, [c_void_p, c_float, ctypes.c_char_p, c_size_t]),
        "fe_lnet_free": (None, [c_void_p]),
        "fe_lnet_error": (ctypes.c_char_p, [c_void_p]),
        "fe_lnet_param_count": (c_int, []),
        "fe_lnet_get_params": (c_int, [c_void_p, c_void_p]),
        "fe_lnet_set_params": (c_int, [c_void_p, c_void_p]),
        "fe_lnet_faces": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_void_p]),
        "fe_lnet_step": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_double,
                                 c_int, c_double, c_void_p, c_void_p]),
        "fe_lnet_step_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, c_double, c_double,
                                         c_double, c_double, c_void_p, c_void_p]),
        "fe_lnet_fallbacks": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_uint)]),
        "fe_lnet_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double,
                                     c_double, c_double, c_double, c_double, c_void_p, c_void_p]),
        "fe_lnet_adam": (c_int, [c_void_p, c_double, c_double]),
        "fe_lnet_coarsen": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
        "fe_mppi_create": (c_void_p, [c_int, c_int, c_int, c_int, ctypes.c_char_p, c_size_t]),
        "fe_mppi_free": (None, [c_void_p]),
        "fe_mppi_error": (ctypes.c_char_p, [c_void_p]),
        "fe_mppi_set_dynamics": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
        "fe_mppi_set_cost": (c_int, [c_void_p, c_void_p, c_void_p, c_float]),
        "fe_mppi_plan": (c_int, [c_void_p, c_int, c_void_p, c_void_p, ctypes.c_uint64, c_float, c_float, c_void_p, c_void_p]),
        "fe_nettrain_create": (c_void_p, [c_void_p, c_int, c_int, ctypes.c_char_p, c_size_t]),
        "fe_nettrain_free": (None, [c_void_p]),
        "fe_nettrain_error": (ctypes.c_char_p, [c_void_p]),
        "fe_nettrain_image_size": (c_int, [c_void_p]),
        "fe_nettrain_set": (c_int, [c_void_p, c_void_p]),
        "fe_nettrain_get": (c_int, [c_void_p, c_void_p]),
        "fe_nettrain_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
        "fe_nettrain_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_float, c_float, c_void_p]),
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


def roof_f32(src: DeviceArray, dst: Optional[DeviceArray], inflight: int, nontemporal: bool, read_only: bool,
             blocks: int) -> None:
    """The HBM roof sweep's kernel (include/fe_field.h fe_roof_f32): a copy, or a read of src alone."""
    if dst is not None and dst.count != src.count:
        raise FieldHipError("roof_f32: sizes differ")
    check(lib().fe_roof_f32(src.ptr, dst.ptr if dst is not None else None, src.count, inflight, int(nontemporal),
                            int(read_only), blocks), "roof_f32")


def _take_string(ptr) -> Optional[str]:
    if not ptr:
        return None
    try:
        return ctypes.string_at(ptr).decode()
    finally:
        lib().fe_string_free(ptr)


def compile_hip(source: str, name: str = "probe.hip", options=()) -> bytes:
    """hipRTC-compiles HIP source for the device's arch (gfx950 without a device) through the runtime's cache."""
    opts = (ctypes.c_char_p * max(1, len(options)))(*[o.encode() for o in options])
    code, size, log, cached = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_void_p(), ctypes.c_int()
    rc = lib().fe_hip_compile(source.encode(), name.encode(), opts, len(options), ctypes.byref(code), ctypes.byref(size),
                              ctypes.byref(log), ctypes.byref(cached))
    text = ctypes.string_at(log.value).decode(errors="replace") if log.value else ""
    if log.value:
        lib().fe_hip_free(log)
    if rc != 0:
        raise FieldHipError(f"hipRTC: {text}")
    blob = ctypes.string_at(code.value, size.value)
    lib().fe_hip_free(code)
    return blob


class Law:
    """A compiled law: the front end and its IR (no device needed)."""

    def __init__(self, *texts: str, names=None):
        names = names or [f"law{i}" for i in range(len(texts))]
        n = len(texts)
        c_names = (ctypes.c_char_p * n)(*[x.encode() for x in names])
        c_texts = (ctypes.c_char_p * n)(*[t.encode() for t in texts])
        err = ctypes.create_string_buffer(1 << 16)
        self.ptr = lib().fe_law_compile(c_names, c_texts, n, err, len(err))
        if not self.ptr:
            raise FieldHipError("the law does not compile:\n" + err.value.decode())
        self.warnings = err.value.decode()

    @classmethod
    def from_files(cls, *paths: str) -> "Law":
        return cls(*[open(p).read() for p in paths], names=list(paths))

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().fe_law_free(self.ptr)
        except Exception:
            pass

    @property
    def field_count(self) -> int:
        return lib().fe_law_field_count(self.ptr)

    def field_id(self, name: str) -> int:
        return lib().fe_law_field_id(self.ptr, name.encode())

    def field_names(self):
        return [lib().fe_law_field_name(self.ptr, i).decode() for i in range(self.field_count)]

    def ir_json(self) -> str:
        return _take_string(lib().fe_law_ir_json(self.ptr))

    def describe(self) -> str:
        return _take_string(lib().fe_law_describe(self.ptr))

    def text(self) -> str:
        """The law written back as law text (compiles to the same IR)."""
        out = _take_string(lib().fe_law_text(self.ptr))
        if out is None:
            raise FieldHipError("this law cannot be written as law text")
        return out

    def hip_source(self) -> str:
        why = ctypes.create_string_buffer(4096)
        src = _take_string(lib().fe_law_hip_source(self.ptr, why, len(why)))
        if src is None:
            raise FieldHipError("no HIP source: " + why.value.decode())
        return src

    def adjoint(self, term: str = "") -> str:
        """The adjoint law of a term (its text): a step of it is one vector-Jacobian product of one step of the term."""
        why = ctypes.create_string_buffer(4096)
        text = _take_string(lib().fe_law_adjoint(self.ptr, term.encode(), why, len(why)))
        if text is None:
            raise FieldHipError("no adjoint: " + why.value.decode())
        return text


class Solver:
    """A law integrating on the card: the reference's Solver calls over the HIP backend (no CPU path)."""

    def __init__(self, law: Law, shape, h: float):
        require_device()
        self.law = law  # keep the law alive
        self.shape = tuple(int(v) for v in shape)
        self.h = float(h)
        why = ctypes.create_string_buffer(1 << 16)
        nx, ny, nz = self.shape
        self.ptr = lib().fe_solver_create(law.ptr, nx, ny, nz, self.h, why, len(why))
        if not self.ptr:
            raise FieldHipError("the law does not run on the card: " + why.value.decode())
        self.cells = nx * ny * nz

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().fe_solver_free(self.ptr)
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise FieldHipError(f"{what}: " + (lib().fe_solver_error(self.ptr) or b"invalid arguments").decode())

    def _field(self, field) -> int:
        f = self.law.field_id(field) if isinstance(field, str) else int(field)
        if f < 0:
            raise FieldHipError(f"no field {field!r}")
        return f

    def set_field(self, field, values: np.ndarray) -> None:
        values = np.ascontiguousarray(values, dtype=np.float32).ravel()
        if values.size != self.cells:
            raise FieldHipError("set_field: wrong number of cells")
        self._check(lib().fe_solver_set_field(self.ptr, self._field(field), values.ctypes.data), "set_field")

    def get_field(self, field) -> np.ndarray:
        out = np.empty(self.cells, dtype=np.float32)
        self._check(lib().fe_solver_get_field(self.ptr, self._field(field), out.ctypes.data), "get_field")
        return out

    def set_param(self, name: str, value: float) -> None:
        if lib().fe_solver_set_param(self.ptr, name.encode(), float(value)) != 0:
            raise FieldHipError(f"no param {name!r}")

    def step(self, dt: float) -> None:
        self._check(lib().fe_solver_step(self.ptr, float(dt)), "step")

    def run(self, dt: float, steps: int) -> None:
        """`steps` steps of dt in one call (with lazy readback, pairs of plain steps replay a captured graph)."""
        self._check(lib().fe_solver_run(self.ptr, float(dt), int(steps)), "run")

    def refresh_derived(self) -> None:
        self._check(lib().fe_solver_refresh_derived(self.ptr), "refresh_derived")

    def set_lazy_readback(self, on: bool) -> None:
        self._check(lib().fe_solver_set_lazy_readback(self.ptr, int(on)), "set_lazy_readback")

    def check_contracts_every(self, steps: int) -> None:
        self._check(lib().fe_solver_check_contracts_every(self.ptr, int(steps)), "check_contracts_every")

    def set_learned(self, on: bool) -> None:
        self._check(lib().fe_solver_set_learned(self.ptr, int(on)), "set_learned")

    def set_respect_max_dt(self, on: bool) -> None:
        self._check(lib().fe_solver_set_respect_max_dt(self.ptr, int(on)), "set_respect_max_dt")

    def violations(self):
        import json
        return json.loads(_take_string(lib().fe_solver_violations_json(self.ptr)))

    def last_substeps(self):
        n = lib().fe_solver_last_substeps(self.ptr, (ctypes.c_int * 1)(), 0)
        buf = (ctypes.c_int * max(1, n))()
        lib().fe_solver_last_substeps(self.ptr, buf, n)
        return list(buf[:n])

    def last_changed(self):
        n = lib().fe_solver_last_changed(self.ptr, (ctypes.c_ubyte * 1)(), 0)
        buf = (ctypes.c_ubyte * max(1, n))()
        lib().fe_solver_last_changed(self.ptr, buf, n)
        return [bool(v) for v in buf[:n]]

    def profile(self):
        import json
        return json.loads(_take_string(lib().fe_solver_profile_json(self.ptr)))

    def digest(self) -> int:
        return int(lib().fe_solver_digest(self.ptr))

    def field_device(self, field) -> int:
        return lib().fe_solver_field_device(self.ptr, self._field(field))

    def total(self, field):
        """(total, total magnitude) of a field, fixed-order compensated on the card."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().fe_solver_total(self.ptr, self._field(field), ctypes.byref(a), ctypes.byref(b)), "total")
        return a.value, b.value

    def total_accumulate(self, field, acc) -> None:
        """Adds the field's (sum, magnitude) into acc (a DeviceArray of at least 4 floats, read as two doubles)."""
        self._check(lib().fe_solver_total_accumulate(self.ptr, self._field(field), acc.ptr), "total_accumulate")

    def set_net_weights(self, net: str, weights) -> None:
        import numpy as np

        w = np.ascontiguousarray(weights, np.float32)
        self._check(lib().fe_solver_set_net_weights(self.ptr, net.encode(), w.ctypes.data, w.size), "set_net_weights")

    def shift_window(self, dx: int, dy: int, dz: int, boxes, arriving) -> None:
        """The window moved by (dx, dy, dz): every field's cells slide (what is now at c was at c + d) and the arriving
        boxes ((x0, y0, z0, x1, y1, z1), half-open) take `arriving`, one array of each field's box cells (x fastest) in
        field order then box order. The card slides its copy and receives only the boxes."""
        import numpy as np

        b = np.ascontiguousarray(np.asarray(boxes, np.int32).reshape(-1, 6))
        a = np.ascontiguousarray(np.asarray(arriving, np.float32).ravel())
        cells = int(sum((q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]) for q in b))
        nf = len(self.law.field_names())
        if a.size != cells * nf:
            raise FieldHipError(f"shift_window: {a.size} arriving values for {cells} box cells of {nf} fields")
        self._check(lib().fe_solver_shift_window(self.ptr, dx, dy, dz, b.ctypes.data if len(b) else None, len(b),
                                                 a.ctypes.data if a.size else None), "shift_window")

    def fill_field(self, field, value: float) -> None:
        self._check(lib().fe_solver_fill_field(self.ptr, self._field(field), float(value)), "fill_field")

    def set_field_device(self, field, src) -> None:
        """Copies a whole field from device memory (a DeviceArray or a device pointer) into the solver on the card."""
        ptr = src.ptr if isinstance(src, DeviceArray) else src
        self._check(lib().fe_solver_set_field_device(self.ptr, self._field(field), ptr), "set_field_device")

    def bind_field_device(self, field, buf) -> None:
        """A static field read from device memory (a DeviceArray, a pointer, or None for the solver's own) without a
        copy; the caller keeps the memory alive while it is bound."""
        ptr = buf.ptr if isinstance(buf, DeviceArray) else buf
        self._check(lib().fe_solver_bind_field_device(self.ptr, self._field(field), ptr), "bind_field_device")


class HestonBatch:
    """Options of uneven lattices stepped together on the card by ADI (include/fe_adi.h). options: n x 8 (kappa, theta,
    sigma, rho, r, strike, maturity, v0); scheme 0 Craig-Sneyd, 1 Hundsdorfer-Verwer."""

    CS, HV = 0, 1

    def __init__(self, nx, ny, options, steps: int, scheme: int = 0, theta: float = 0.5):
        import numpy as np

        self.nx = np.ascontiguousarray(nx, dtype=np.int32)
        self.ny = np.ascontiguousarray(ny, dtype=np.int32)
        self.options = np.ascontiguousarray(options, dtype=np.float64).reshape(-1, 8)
        self.n = len(self.nx)
        self.steps = steps
        why = ctypes.create_string_buffer(2048)
        self.ptr = lib().fe_heston_create(self.n, self.nx.ctypes.data, self.ny.ctypes.data, self.options.ctypes.data, steps,
                                          scheme, theta, why, len(why))
        if not self.ptr:
            raise FieldHipError("batch refused: " + why.value.decode())

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fe_heston_free(self.ptr)
            self.ptr = None

    def _check(self, status: int, what: str) -> None:
        if status != 0:
            raise FieldHipError(f"{what}: {lib().fe_heston_error(self.ptr).decode()}")

    @property
    def cells(self) -> int:
        return int(lib().fe_heston_cells(self.ptr))

    def reset(self) -> None:
        self._check(lib().fe_heston_reset(self.ptr), "reset")

    def step(self, count: int = 1) -> None:
        self._check(lib().fe_heston_step(self.ptr, count), "step")

    def capture(self, count: int) -> None:
        self._check(lib().fe_heston_capture(self.ptr, count), "capture")

    def launch(self) -> None:
        self._check(lib().fe_heston_launch(self.ptr), "launch")

    def sync(self) -> None:
        self._check(lib().fe_heston_sync(self.ptr), "sync")

    def prices(self):
        import numpy as np

        out = np.empty(self.n, np.float64)
        self._check(lib().fe_heston_prices(self.ptr, out.ctypes.data), "prices")
        return out

    def lattice(self, option: int):
        import numpy as np

        out = np.empty(int(self.nx[option]) * int(self.ny[option]), np.float32)
        self._check(lib().fe_heston_lattice(self.ptr, option, out.ctypes.data), "lattice")
        return out

    def shapes(self):
        out = []
        for c in range(lib().fe_heston_classes(self.ptr)):
            v = [ctypes.c_int() for _ in range(4)]
            lib().fe_heston_get_shape(self.ptr, c, *[ctypes.byref(x) for x in v])
            out.append({"nx": v[0].value, "lanes": v[1].value, "row_waves": v[2].value, "col_threads": v[3].value})
        return out

    def set_shape(self, cls: int, lanes: int, row_waves: int, col_threads: int) -> None:
        self._check(lib().fe_heston_set_shape(self.ptr, cls, lanes, row_waves, col_threads), "set_shape")

    def tune(self, reps: int = 2):
        """Picks each class's launch shapes by timing them (the lattices return to their payoffs)."""
        import json

        report = ctypes.create_string_buffer(8192)
        self._check(lib().fe_heston_tune(self.ptr, reps, report, len(report)), "tune")
        return json.loads(report.value.decode())


class LearnedNet:
    """The learned flux correction of a coarse 2-D two-field lattice (include/fe_learned.h). Device arrays in and out
    (DeviceArray or raw device pointers); `scale` multiplies the net's outputs into face fluxes."""

    def __init__(self, params, scale: float):
        import numpy as np

        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != lib().fe_lnet_param_count():
            raise FieldHipError(f"{p.size} parameters given, the net has {lib().fe_lnet_param_count()}")
        why = ctypes.create_string_buffer(1024)
        self.ptr = lib().fe_lnet_create(p.ctypes.data, scale, why, len(why))
        if not self.ptr:
            raise FieldHipError("net refused: " + why.value.decode())
        self.scale = scale

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fe_lnet_free(self.ptr)
            self.ptr = None

    def _check(self, status: int, what: str) -> None:
        if status != 0:
            raise FieldHipError(f"{what}: {lib().fe_lnet_error(self.ptr).decode()}")

    @staticmethod
    def _p(a):
        return a.ptr if isinstance(a, DeviceArray) else a

    def params(self):
        import numpy as np

        out = np.empty(lib().fe_lnet_param_count(), np.float32)
        self._check(lib().fe_lnet_get_params(self.ptr, out.ctypes.data), "get_params")
        return out

    def set_params(self, p) -> None:
        import numpy as np

        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(lib().fe_lnet_set_params(self.ptr, p.ctypes.data), "set_params")

    def faces(self, u, v, nx, ny, count, h, out) -> None:
        self._check(lib().fe_lnet_faces(self.ptr, self._p(u), self._p(v), nx, ny, count, h, self._p(out)), "faces")

    def step(self, u, v, nx, ny, count, h, du, dv, k, dt, u2, v2, learned=True, bound=1e3) -> None:
        self._check(lib().fe_lnet_step(self.ptr, self._p(u), self._p(v), nx, ny, count, h, du, dv, k, dt, int(learned), bound,
                                       self._p(u2), self._p(v2)), "step")

    def step_noalloc(self, u, v, nx, ny, count, h, du, dv, k, dt, u2, v2, bound=1e3) -> None:
        """The learned step allocating nothing (its scratch is the library's static device array; at most 2^20 cells)."""
        self._check(lib().fe_lnet_step_noalloc(self.ptr, self._p(u), self._p(v), nx, ny, count, h, du, dv, k, dt, bound,
                                               self._p(u2), self._p(v2)), "step_noalloc")

    def fallbacks(self) -> int:
        c = ctypes.c_uint()
        self._check(lib().fe_lnet_fallbacks(self.ptr, ctypes.byref(c)), "fallbacks")
        return int(c.value)

    def gradient(self, u0, v0, u1, v1, nx, ny, count, h, du, dv, k, dt, want=True, weight=1.0):
        import numpy as np

        g = np.empty(lib().fe_lnet_param_count(), np.float32) if want else None
        loss = np.empty(1, np.float32) if want else None
        self._check(lib().fe_lnet_gradient(self.ptr, self._p(u0), self._p(v0), self._p(u1), self._p(v1), nx, ny, count, h, du, dv,
                                           k, dt, float(weight), g.ctypes.data if want else None,
                                           loss.ctypes.data if want else None), "gradient")
        return (float(loss[0]), g) if want else None

    def adam(self, lr: float, clip: float = 1e-2) -> None:
        self._check(lib().fe_lnet_adam(self.ptr, lr, clip), "adam")


def net_half_forward(sizes, act: int, out_scale: float, weights, rows):
    """A `half` law net's forward pass on the matrix cores (fp16 operands, fp32 accumulation): rows (n x inputs,
    already normalised) -> n x outputs."""
    import numpy as np

    sizes = np.ascontiguousarray(sizes, np.int32)
    w = np.ascontiguousarray(weights, np.float32)
    x = np.ascontiguousarray(rows, np.float32).reshape(-1, int(sizes[0]))
    out = np.empty((x.shape[0], int(sizes[-1])), np.float32)
    rc = lib().fe_net_half_forward(sizes.ctypes.data, len(sizes), act, out_scale, w.ctypes.data, x.ctypes.data, x.shape[0],
                                   out.ctypes.data)
    if rc != 0:
        raise FieldHipError(f"net_half_forward failed ({rc})")
    return out


def coarsen(fine, fnx, fny, count, factor, coarse) -> None:
    p = lambda a: a.ptr if isinstance(a, DeviceArray) else a  # noqa: E731
    if lib().fe_lnet_coarsen(p(fine), fnx, fny, count, factor, p(coarse)) != 0:
        raise FieldHipError("coarsen: invalid arguments")


class MppiPlanner:
    """A world model's MPPI planner on the card (include/fe_mppi.h): dynamics F + 2 -> hidden -> hidden -> F (tanh),
    `samples` futures of `horizon` steps for every agent, a weighted quadratic cost about a target."""

    def __init__(self, features: int, hidden: int, horizon: int, samples: int):
        why = ctypes.create_string_buffer(1024)
        self.ptr = lib().fe_mppi_create(features, hidden, horizon, samples, why, len(why))
        if not self.ptr:
            raise FieldHipError("planner refused: " + why.value.decode())
        self.features, self.hidden, self.horizon, self.samples = features, hidden, horizon, samples

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fe_mppi_free(self.ptr)
            self.ptr = None

    def _check(self, status: int, what: str) -> None:
        if status != 0:
            raise FieldHipError(f"{what}: {lib().fe_mppi_error(self.ptr).decode()}")

    def set_dynamics(self, weights, in_shift, in_scale, scale) -> None:
        import numpy as np

        arrays = [np.ascontiguousarray(a, np.float32) for a in (weights, in_shift, in_scale, scale)]
        self._keep = arrays
        self._check(lib().fe_mppi_set_dynamics(self.ptr, *[a.ctypes.data for a in arrays]), "set_dynamics")

    def set_cost(self, target, weight, action_weight: float) -> None:
        import numpy as np

        t, w = np.ascontiguousarray(target, np.float32), np.ascontiguousarray(weight, np.float32)
        self._check(lib().fe_mppi_set_cost(self.ptr, t.ctypes.data, w.ctypes.data, float(action_weight)), "set_cost")

    def plan(self, obs, ids, mean, tick: int, sigma: float, dt: float, costs: bool = False):
        """(new mean agents x horizon x 2, costs agents x samples or None)."""
        import numpy as np

        obs = np.ascontiguousarray(obs, np.float32)
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.ascontiguousarray(mean, np.float32).copy()
        agents = ids.size
        c = np.empty((agents, self.samples), np.float32) if costs else None
        self._check(lib().fe_mppi_plan(self.ptr, agents, obs.ctypes.data, ids.ctypes.data, int(tick), float(sigma), float(dt),
                                       out.ctypes.data, c.ctypes.data if costs else None), "plan")
        return out, c


ACTIVATIONS = ("tanh", "relu", "silu", "linear", "gelu")


class NetTrainer:
    """A law net's trainer on the card (include/fe_nettrain.h): sizes (inputs, hidden widths, outputs), a hidden
    activation by name, the weights in the reference's GPU net image (in_shift, in_scale, out_scale, zeros to a multiple
    of 4 floats, then W[out][in] and b[out] per layer); the loss is the mean over rows of 0.5 m |out_scale f(x) - t|^2,
    Adam clips each gradient first."""

    def __init__(self, sizes, act: str = "tanh"):
        import numpy as np

        if act not in ACTIVATIONS:
            raise FieldHipError(f"activation {act!r}: one of {', '.join(ACTIVATIONS)}")
        self.sizes = [int(v) for v in sizes]
        self.act = act
        arr = np.ascontiguousarray(self.sizes, np.int32)
        why = ctypes.create_string_buffer(1024)
        self.ptr = lib().fe_nettrain_create(arr.ctypes.data, arr.size, ACTIVATIONS.index(act), why, len(why))
        if not self.ptr:
            raise FieldHipError("trainer refused: " + why.value.decode())
        self.image_size = lib().fe_nettrain_image_size(self.ptr)
        self.base = (2 * self.sizes[0] + 1 + 3) // 4 * 4  # the first weight
        self.trained = self.image_size - self.base

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fe_nettrain_free(self.ptr)
            self.ptr = None

    def _check(self, status: int, what: str) -> None:
        if status != 0:
            raise FieldHipError(f"{what}: {lib().fe_nettrain_error(self.ptr).decode()}")

    def set_image(self, image) -> None:
        import numpy as np

        a = np.ascontiguousarray(image, np.float32)
        if a.size != self.image_size:
            raise FieldHipError(f"image of {a.size} floats, the net's is {self.image_size}")
        self._check(lib().fe_nettrain_set(self.ptr, a.ctypes.data), "set")

    def image(self):
        import numpy as np

        out = np.empty(self.image_size, np.float32)
        self._check(lib().fe_nettrain_get(self.ptr, out.ctypes.data), "get")
        return out

    def _rows(self, x, target, m):
        import numpy as np

        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        rows = x.shape[0] if x.ndim == 2 else 0
        if x.ndim != 2 or x.shape[1] != self.sizes[0] or t.shape != (rows, self.sizes[-1]):
            raise FieldHipError(f"rows: x {x.shape} and targets {t.shape} for sizes {self.sizes}")
        mm = None if m is None else np.ascontiguousarray(m, np.float32)
        if mm is not None and mm.shape != (rows,):
            raise FieldHipError(f"row weights {mm.shape} for {rows} rows")
        return x, t, mm, rows

    def gradient(self, x, target, m=None):
        """(loss, gradient of the trained floats) for the rows; no update."""
        import numpy as np

        x, t, mm, rows = self._rows(x, target, m)
        g = np.empty(self.trained, np.float32)
        loss = ctypes.c_float()
        self._check(lib().fe_nettrain_gradient(self.ptr, x.ctypes.data, t.ctypes.data, None if mm is None else mm.ctypes.data, rows,
                                               ctypes.byref(loss), g.ctypes.data), "gradient")
        return loss.value, g

    def step(self, x, target, m=None, lr: float = 2e-3, clip: float = 5.0) -> float:
        """One Adam step on the rows; the loss before it."""
        x, t, mm, rows = self._rows(x, target, m)
        loss = ctypes.c_float()
        self._check(lib().fe_nettrain_step(self.ptr, x.ctypes.data, t.ctypes.data, None if mm is None else mm.ctypes.data, rows,
                                           float(lr), float(clip), ctypes.byref(loss)), "step")
        return loss.value
