"""The handwritten gfx950 conduction kernel against the reference CPU engine's bits (golden) and the ORACLE."""
import numpy as np
import pytest

import golden_io
from oracle import heat

pytestmark = pytest.mark.gpu


def run_handwritten(field_hip, t0, shape, k, h, dt, steps, coeff=0.15, **cfg):
    nx, ny, nz = shape
    subs = heat.substeps(dt, coeff, k, h)
    a = field_hip.DeviceArray.from_host(t0)
    b = field_hip.DeviceArray(a.count)
    for _ in range(steps * subs):
        field_hip.heat_handwritten(a, b, nx, ny, nz, k, h, dt / subs, **cfg)
        a, b = b, a
    return a.download()


@pytest.mark.parametrize("name", golden_io.names("heat_"))
def test_handwritten_matches_reference_bits(device, name):
    import field_hip

    meta, init, final = golden_io.load(name)
    nx, ny, nz, h = meta["grid"]
    print(f"case {name} grid {nx}x{ny}x{nz}", flush=True)
    out = run_handwritten(field_hip, init["T"], (nx, ny, nz), meta["params"]["k"], h, meta["dt"], meta["steps"])
    assert np.array_equal(out.view(np.uint32), final["T"].view(np.uint32)), \
        f"{name}: max |diff| {np.max(np.abs(out - final['T']))}"


SHAPES = [
    # variant 1: a lane a column (any nx)
    (1, (130, 37, 19), 1, 1, 0), (1, (130, 37, 19), 2, 7, 1), (1, (130, 37, 19), 4, 32, 1), (1, (130, 37, 19), 8, 64, 0),
    (1, (130, 37, 19), 4, 1000, 1),
    # variant 2: four cells a lane (nx % 4 == 0), tiles of 256 columns: nx below, at and above one tile
    (2, (132, 37, 19), 2, 1, 0), (2, (132, 37, 19), 6, 3, 1), (2, (132, 37, 19), 14, 8, 1), (2, (256, 9, 11), 6, 4, 1),
    (2, (516, 13, 10), 6, 5, 0), (2, (8, 3, 2), 2, 1, 1), (2, (260, 20, 33), 14, 64, 1),
]


@pytest.mark.parametrize("variant,shape,ty,zchunk,remap", SHAPES)
def test_every_launch_shape_gives_the_oracle_bits(device, variant, shape, ty, zchunk, remap):
    import field_hip

    print(f"variant {variant} shape {shape} ty {ty} zchunk {zchunk} remap {remap}", flush=True)
    rng = np.random.default_rng(11)
    t = rng.random(int(np.prod(shape))).astype(np.float32)
    for k, h, sub in ((0.2, 0.37, 0.05), (0.1, 0.25, 0.1)):  # h*h not a power of two, then one
        ref = heat.substep(t, shape, k, h, sub)
        a = field_hip.DeviceArray.from_host(t)
        b = field_hip.DeviceArray(a.count)
        field_hip.heat_handwritten(a, b, *shape, k, h, sub, ty=ty, zchunk=zchunk, remap=remap, variant=variant)
        out = b.download()
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), f"max |diff| {np.max(np.abs(out - ref))}"


def test_denormals_flush_as_on_the_cpu(device):
    import field_hip

    # values near the smallest normal float: the CPU's FTZ/DAZ threads flush results below it to signed zero
    shape = (64, 8, 4)
    rng = np.random.default_rng(3)
    t = (rng.random(int(np.prod(shape))) * 4e-38 - 2e-38).astype(np.float32)
    for variant in (1, 2):
        ref = heat.substep(t, shape, 0.3, 1.0, 0.5)
        a = field_hip.DeviceArray.from_host(t)
        b = field_hip.DeviceArray(a.count)
        field_hip.heat_handwritten(a, b, *shape, 0.3, 1.0, 0.5, variant=variant)
        out = b.download()
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), variant


def test_full_size_substep_is_the_oracle_and_repeats(device):
    import field_hip

    n = 256
    shape = (n, n, n)
    rng = np.random.default_rng(2026)
    t = rng.random(n ** 3).astype(np.float32)
    k, h = 0.1, 1.0 / n
    sub = 0.4 * h * h / k / 3
    ref = heat.substep(t, shape, k, h, sub)
    a = field_hip.DeviceArray.from_host(t)
    b = field_hip.DeviceArray(a.count)
    field_hip.heat_handwritten(a, b, *shape, k, h, sub)
    first = b.download()
    field_hip.heat_handwritten(a, b, *shape, k, h, sub)
    second = b.download()
    assert np.array_equal(first.view(np.uint32), second.view(np.uint32)), "two runs differ"
    assert np.array_equal(first.view(np.uint32), ref.view(np.uint32)), f"max |diff| {np.max(np.abs(first - ref))}"


def test_copy_kernel(device):
    import field_hip

    x = np.arange(1000003, dtype=np.float32)
    a = field_hip.DeviceArray.from_host(x)
    b = field_hip.DeviceArray(a.count)
    field_hip.copy_f32(a, b)
    assert np.array_equal(b.download(), x)


FULL_SIZE_LAW = """field T init 0 range -1000000 1000000
param k = 0.1
term conduction
  d T = diffuse(k, T)
  conserves T
  max_dt = 0.15 * h * h / k
"""


@pytest.mark.parametrize("n,steps", [(256, 40), (512, 10)])
@pytest.mark.parametrize("contracts", [0, 8])
def test_engine_run_at_full_size_is_the_oracle(device, n, steps, contracts):
    """The generated engine at production size, through Solver.run (captured graph replays between the steps that
    check contracts), against the oracle's bits: 256^3 crosses a 32-step graph and its remainder, 512^3 runs the
    streaming-store window. contracts=0: never checked after the first step; 8: checked every 8th step."""
    import field_hip

    shape = (n, n, n)
    k, h = 0.1, 1.0 / n
    sub = 0.4 * h * h / k / 3
    print(f"engine full size: {n}^3, {steps} steps, contracts every {contracts or 'never'}", flush=True)
    t0 = np.random.default_rng(2026 + n).random(n ** 3, dtype=np.float32)
    s = field_hip.Solver(field_hip.Law(FULL_SIZE_LAW), shape, h)
    s.set_field("T", t0)
    s.set_lazy_readback(True)
    s.check_contracts_every(contracts or (1 << 30))
    s.run(sub, steps)
    got = s.get_field("T")
    assert s.violations() == [], s.violations()
    ref = t0
    for _ in range(steps):
        ref = heat.substep(ref, shape, k, h, sub)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"max |diff| {np.max(np.abs(got - ref))}"
    # The conserved total the card reports is the oracle's fixed-order compensated total of the same bits.
    card_total, _ = s.total("T")
    oracle_total, _ = heat.total(ref)
    assert abs(card_total - oracle_total) <= 1e-9 * abs(oracle_total)
