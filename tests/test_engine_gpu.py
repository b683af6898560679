"""The generated HIP backend against the reference CPU engine's own outputs, bit for bit.

Each golden case is a law, a lattice, a seeded state and N steps of dt, run once through the reference's CPU engine
(tests/golden/make_golden.py). Here the same law runs on the card through the C ABI (fe_solver_*), and every field
must come back with the reference's bits.
"""
import numpy as np
import pytest

import golden_io

pytestmark = pytest.mark.gpu


def run_engine(field_hip, meta, init, lazy=False, contracts=1):
    nx, ny, nz, h = meta["grid"]
    law = field_hip.Law(meta["law"])
    solver = field_hip.Solver(law, (nx, ny, nz), h)
    for name, value in meta["params"].items():
        solver.set_param(name, value)
    for name, values in init.items():
        solver.set_field(name, values)
    solver.refresh_derived()
    solver.set_lazy_readback(lazy)
    solver.check_contracts_every(contracts)
    for _ in range(meta["steps"]):
        solver.step(meta["dt"])
    return law, solver


def exact(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("name", golden_io.names())
def test_engine_matches_reference_bits(device, name):
    import field_hip

    meta, init, final = golden_io.load(name)
    print(f"case {name} grid {meta['grid']} steps {meta['steps']}", flush=True)
    law, solver = run_engine(field_hip, meta, init)
    for field, expected in final.items():
        got = solver.get_field(field)
        assert exact(got, expected), f"{name}/{field}: max |diff| {np.max(np.abs(got - expected))}"
    assert solver.violations() == [] or all(v["kind"] != "conservation" for v in solver.violations())


@pytest.mark.parametrize("name", ["heat_20x14x9", "multi", "diffreact_rk4"])
def test_two_runs_and_lazy_readback_agree_to_the_bit(device, name):
    import field_hip

    meta, init, final = golden_io.load(name)
    _, a = run_engine(field_hip, meta, init)
    _, b = run_engine(field_hip, meta, init, lazy=True, contracts=1000)
    assert a.digest() == b.digest()


def test_contracts_report_a_leak_and_a_range_breach(device):
    import field_hip

    text = ("field T init 0 range 0 1\nparam k = 0.2\nterm leaky\n  d T = diffuse(k, T) + 0.5\n  conserves T\n")
    law = field_hip.Law(text)
    solver = field_hip.Solver(law, (16, 8, 4), 1.0)
    solver.set_field("T", np.linspace(0.0, 0.9, 16 * 8 * 4, dtype=np.float32))
    solver.step(1.0)
    kinds = sorted(v["kind"] for v in solver.violations())
    assert kinds == ["conservation", "range"], solver.violations()
    # without the leak the same term holds both contracts
    law2 = field_hip.Law(text.replace(" + 0.5", ""))
    ok = field_hip.Solver(law2, (16, 8, 4), 1.0)
    ok.set_field("T", np.linspace(0.0, 0.9, 16 * 8 * 4, dtype=np.float32))
    ok.step(1.0)
    assert ok.violations() == []


def test_substeps_follow_max_dt_and_can_be_ignored(device):
    import field_hip

    meta, init, _ = golden_io.load("heat_20x14x9")
    law, solver = run_engine(field_hip, dict(meta, steps=1), init)
    assert solver.last_substeps() == [3]
    solver.set_respect_max_dt(False)
    solver.step(meta["dt"])
    assert solver.last_substeps() == [1]


def test_profile_reports_card_time(device):
    import field_hip

    meta, init, _ = golden_io.load("multi")
    _, solver = run_engine(field_hip, dict(meta, steps=1), init)
    p = solver.profile()
    assert p["backend"] == "HIP" and p["launches"] > 0 and p["gpu_ms"] > 0.0
    assert [t["name"] for t in p["terms"]] == ["exchange", "decay"]
