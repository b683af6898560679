"""The C ABI: the library loads and exports every entry point include/*.h declares (no device needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for header in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not header.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", header)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(fe_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "fe_hip_device_open" in names and "fe_heat_handwritten" in names
    assert len(names) > 20


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports(name):
    import field_hip

    lib = field_hip.lib()
    assert hasattr(lib, name), f"{name} is declared in include/ but libfieldhip.so does not export it"


def test_bf16_rounding():
    import field_hip

    lib = field_hip.lib()
    # round to nearest even at the 16-bit boundary, and a NaN stays a NaN
    assert lib.fe_hip_to_bf16(1.0) == 0x3F80
    assert lib.fe_hip_to_bf16(ctypes.c_float(1.0 + 2 ** -8).value) == 0x3F80  # exactly half way: to even
    assert lib.fe_hip_to_bf16(ctypes.c_float(1.0 + 3 * 2 ** -8).value) == 0x3F82
    assert lib.fe_hip_to_bf16(float("nan")) & 0x7FC0 == 0x7FC0


def test_no_device_here_is_reported_in_words():
    import field_hip

    info = field_hip.device_info()
    if info.ok:
        pytest.skip("a device is present")
    assert info.why and b"device" in info.why


def test_field_backend_cpu_keeps_the_law_off_the_card(monkeypatch):
    """FIELD_BACKEND=cpu is refused by name before any device is touched (src/field/Solver.cpp:173-200 reads it the
    same way); this build has no CPU integrator to fall back on."""
    import field_hip

    monkeypatch.setenv("FIELD_BACKEND", "cpu")
    law = field_hip.Law("field T init 0\nterm t\n  d T = diffuse(0.1, T)\n")
    why = ctypes.create_string_buffer(4096)
    ptr = field_hip.lib().fe_solver_create(law.ptr, 4, 4, 4, 1.0, why, len(why))
    assert not ptr
    assert b"FIELD_BACKEND=cpu" in why.value
