"""Solver.run (fe_solver_run): pairs of plain steps replayed from a captured graph give the bits of as many eager steps.

Each case runs twice from the same start: eager (one fe_solver_step a step) and through fe_solver_run in chunks of
several sizes, the host reading every field back between the chunks (so the graph is replayed after eager work and
waits of the program's own). Every field must agree to the bit after every chunk, and the run must end on the
reference engine's golden bits.
"""
import numpy as np
import pytest

import golden_io

pytestmark = pytest.mark.gpu

CASES = ["heat_20x14x9", "multi", "op_poisson", "op_distance_moving", "kitchen", "net_flux_relu", "net_ldiffuse_half", "adi_yz"]


def make(field_hip, meta, init):
    nx, ny, nz, h = meta["grid"]
    solver = field_hip.Solver(golden_io.make_law(field_hip, meta), (nx, ny, nz), h)
    for name, value in meta["params"].items():
        solver.set_param(name, value)
    for name, values in init.items():
        solver.set_field(name, values)
    solver.refresh_derived()
    solver.set_lazy_readback(True)
    solver.check_contracts_every(1 << 30)
    return solver


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", CASES)
def test_replayed_pairs_match_eager_steps(device, name):
    import field_hip

    meta, init, _ = golden_io.load(name)
    print(f"case {name}", flush=True)
    eager, replay = make(field_hip, meta, init), make(field_hip, meta, init)
    fields = list(init)
    for chunk in (3, 2, 4, 1, 6, 2, 35, 33):
        for _ in range(chunk):
            eager.step(meta["dt"])
        replay.run(meta["dt"], chunk)
        for f in fields:
            assert np.array_equal(bits(eager.get_field(f)), bits(replay.get_field(f))), f"{name}/{f} after a run of {chunk}"


@pytest.mark.parametrize("name", CASES)
def test_run_ends_on_the_reference_bits(device, name):
    import field_hip

    meta, init, final = golden_io.load(name)
    solver = make(field_hip, meta, init)
    solver.run(meta["dt"], meta["steps"])
    for f, expected in final.items():
        assert golden_io.same(meta, solver.get_field(f), expected), f"{name}/{f}"
