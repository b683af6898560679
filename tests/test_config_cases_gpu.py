"""The bench's Dirichlet heat (64^3, 200 steps of 3 substeps) and Black-Scholes (4096 cells x 512 planes, 10 steps)
runs on the card, held to the reference CPU engine's final bits (SHA-256 digests in tests/golden/bench_reference.json,
written by tests/golden/make_bench_reference.py from the reference's own engine)."""
import hashlib
import json
import os

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

REF = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bench_reference.json")))


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


@pytest.mark.parametrize("how", ["step", "run"])
def test_dirichlet_heat_64_is_the_reference_bits(device, how):
    import field_hip

    n = 64
    h = 1.0 / n
    t0, nb = bench.sine_start(n, 1e-3)
    s = field_hip.Solver(field_hip.Law(bench.DIRICHLET_LAW), (n, n, n), h)
    s.set_field("T", t0)
    s.set_field("nb", nb)
    s.set_lazy_readback(True)
    s.check_contracts_every(1 << 30)
    dt = 0.4 * h * h / 0.1
    if how == "run":
        s.run(dt, 200)
    else:
        for _ in range(200):
            s.step(dt)
    assert s.last_substeps() == [3]
    assert digest(s.get_field("T")) == REF["dirichlet_heat_64"]["sha256"]


def test_black_scholes_4096x512_is_the_reference_bits(device):
    import field_hip

    v0, h, dt = bench.black_scholes_start(4096, 512)
    s = field_hip.Solver(field_hip.Law(bench.BLACK_SCHOLES_LAW), (4096, 1, 512), h)
    s.set_field("V", v0)
    for _ in range(10):
        s.step(dt)
    assert s.violations() == []
    assert digest(s.get_field("V")) == REF["black_scholes_4096x512"]["sha256"]
