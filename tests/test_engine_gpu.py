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
    law = golden_io.make_law(field_hip, meta)
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
        assert golden_io.same(meta, got, expected), f"{name}/{field}: max |diff| {np.max(np.abs(got - expected))}"
    if "final_violations" in meta:
        # The last step's broken contracts are the reference's: the same terms, fields, kinds and cell counts.
        got_v = {(v["term"], v["field"], v["kind"], int(v["detail"].split()[0]) if v["kind"] == "range" else -1)
                 for v in solver.violations()}
        want_v = {(v["term"], v["field"], v["kind"], v["cells"] if v["kind"] == "range" else -1) for v in meta["final_violations"]}
        assert got_v == want_v
    else:
        assert solver.violations() == [] or all(v["kind"] != "conservation" for v in solver.violations())


@pytest.mark.parametrize("name", ["heat_20x14x9", "multi", "diffreact_rk4", "net_flux_tanh", "net_ldiffuse_relu"])
def test_two_runs_and_lazy_readback_agree_to_the_bit(device, name):
    import field_hip

    meta, init, final = golden_io.load(name)
    _, a = run_engine(field_hip, meta, init)
    _, b = run_engine(field_hip, meta, init, lazy=True, contracts=1000)
    assert a.digest() == b.digest()


@pytest.mark.parametrize("how", ["step", "lazy", "run"])
def test_clamped_fields_are_clamped_on_steps_that_check_no_contract(device, how):
    """The reference CPU clamps a clamped field after every term that writes it and after the derives, every step;
    here contracts are checked once in 1000 steps, so only the clamp kernel can keep the fields in range."""
    import field_hip

    meta, init, final = golden_io.load("clamp_every_step")
    nx, ny, nz, h = meta["grid"]
    solver = field_hip.Solver(golden_io.make_law(field_hip, meta), (nx, ny, nz), h)
    for name, values in init.items():
        solver.set_field(name, values)
    solver.refresh_derived()
    solver.set_lazy_readback(how != "step")
    solver.check_contracts_every(1000)
    solver.step(meta["dt"])  # step 0 checks contracts; the others do not
    if how == "run":
        solver.run(meta["dt"], meta["steps"] - 1)
    else:
        for _ in range(meta["steps"] - 1):
            solver.step(meta["dt"])
    for field, expected in final.items():
        assert exact(solver.get_field(field), expected), field


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


def test_judged_steps_reuse_a_total_only_while_it_stands(device):
    # Each judged step's "before" total is the last judged "after" one when nothing wrote the field since; an upload,
    # a device write, another term's write or a clamp must make it be taken again (else a false leak is reported).
    import field_hip

    text = ("field T init 0 range -1000000 1000000\nfield S init 0 range 0 1 clamp\nparam k = 0.2\nparam leak = 0\n"
            "term conduct\n  d T = diffuse(k, T) + leak\n  d S = diffuse(k, S) + 0.3\n  conserves T\n"
            "term source\n  d T = 0.01 * S\n")
    law = field_hip.Law(text)
    solver = field_hip.Solver(law, (16, 8, 4), 1.0)
    rng = np.random.default_rng(5)
    solver.set_field("T", rng.random(16 * 8 * 4).astype(np.float32))
    solver.set_field("S", rng.random(16 * 8 * 4).astype(np.float32))
    solver.check_contracts_every(1)

    def kinds():
        return sorted(v["kind"] for v in solver.violations())

    for _ in range(3):
        solver.step(1.0)
        assert kinds() == [], solver.violations()
    solver.set_param("leak", 0.5)
    solver.step(1.0)
    assert kinds() == ["conservation"], solver.violations()
    solver.set_param("leak", 0.0)
    solver.step(1.0)
    assert kinds() == []
    solver.set_field("T", rng.random(16 * 8 * 4).astype(np.float32) * 7)
    solver.step(1.0)
    assert kinds() == [], solver.violations()
    src = field_hip.DeviceArray.from_host(rng.random(16 * 8 * 4).astype(np.float32) * 3)
    solver.set_field_device("T", src)
    solver.step(1.0)
    assert kinds() == [], solver.violations()
    solver.run(1.0, 5)
    solver.step(1.0)
    assert kinds() == [], solver.violations()


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


# Every gather the generator emits for a law, picked through the launch-shape overrides: the direct kernel (a thread a
# cell), the window kernel with a cell a lane (64 x 6 tiles) and with four cells a lane (256 x 2 and 256 x 6 tiles).
GATHERS = {"direct": {"FIELD_HIP_DIRECT": "1"}, "w1x6": {"FIELD_HIP_TILE": "1x6"}, "w4x2": {"FIELD_HIP_TILE": "4x2"},
           "w4x4": {"FIELD_HIP_TILE": "4x4"}, "w4x6": {"FIELD_HIP_TILE": "4x6"},
           "w4x2z5": {"FIELD_HIP_TILE": "4x2", "FIELD_HIP_ZCHUNK": "5"},
           # launch order and streaming stores (what lattices past the last-level cache get)
           "w4x6s": {"FIELD_HIP_TILE": "4x6", "FIELD_HIP_REMAP": "0", "FIELD_HIP_STREAM": "1"}}


def use_gather(monkeypatch, gather):
    for key in ("FIELD_HIP_DIRECT", "FIELD_HIP_TILE", "FIELD_HIP_ZCHUNK", "FIELD_HIP_REMAP", "FIELD_HIP_STREAM"):
        monkeypatch.delenv(key, raising=False)
    for key, value in GATHERS[gather].items():
        monkeypatch.setenv(key, value)


@pytest.mark.parametrize("gather", sorted(GATHERS))
@pytest.mark.parametrize("name", golden_io.names())
def test_every_gather_matches_reference_bits(device, monkeypatch, gather, name):
    import field_hip

    use_gather(monkeypatch, gather)
    meta, init, final = golden_io.load(name)
    print(f"case {name} gather {gather} grid {meta['grid']}", flush=True)
    _, solver = run_engine(field_hip, meta, init)
    for field, expected in final.items():
        assert golden_io.same(meta, solver.get_field(field), expected), f"{name}/{field} through {gather}"


CONDUCTION = "field T init 0\nparam k = 0.1\nterm conduction\n  d T = diffuse(k, T)\n"

# Rows of several tiles, partial tiles, rows that are not whole float4s, flat and thin lattices.
LATTICES = [(260, 9, 7), (516, 5, 6), (200, 7, 9), (13, 11, 6), (4, 3, 2), (2, 1, 1), (64, 1, 1), (8, 1, 70), (1, 5, 9)]


@pytest.mark.parametrize("gather", sorted(GATHERS))
@pytest.mark.parametrize("shape", LATTICES, ids=lambda s: "x".join(map(str, s)))
def test_conduction_matches_the_oracle_on_uneven_lattices(device, monkeypatch, gather, shape):
    import field_hip
    from oracle import heat

    use_gather(monkeypatch, gather)
    nx, ny, nz = shape
    h, k = 1.0 / 16, 0.1
    dt = 0.4 * h * h / k / 3
    t0 = np.random.default_rng(nx * 1000 + ny * 10 + nz).random(nx * ny * nz, dtype=np.float32)
    print(f"lattice {shape} gather {gather}", flush=True)
    solver = field_hip.Solver(field_hip.Law(CONDUCTION), shape, h)
    solver.set_field("T", t0)
    expect = t0
    for _ in range(3):
        solver.step(dt)
        expect = heat.substep(expect, shape, k, h, dt)
    assert exact(solver.get_field("T"), expect)


@pytest.mark.parametrize("gather", sorted(GATHERS))
def test_conduction_with_infinities_and_nan_follows_the_oracle(device, monkeypatch, gather):
    import field_hip
    from oracle import heat

    use_gather(monkeypatch, gather)
    shape = (24, 7, 6)
    h, k = 1.0 / 16, 0.1
    dt = 0.4 * h * h / k / 3
    t0 = np.random.default_rng(5).random(24 * 7 * 6, dtype=np.float32)
    t0[3 * 24 * 7 + 3 * 24 + 5] = np.inf    # interior
    t0[0] = -np.inf                         # a corner: three faces missing
    t0[2 * 24 * 7 + 6 * 24 + 23] = np.nan   # an edge cell
    t0[5 * 24 * 7 + 24 + 0] = np.inf        # a face cell
    solver = field_hip.Solver(field_hip.Law(CONDUCTION), shape, h)
    solver.set_field("T", t0)
    solver.step(dt)
    got = solver.get_field("T")
    expect = heat.substep(t0, shape, k, h, dt)
    nan = np.isnan(expect)
    assert np.array_equal(np.isnan(got), nan)  # NaN payloads differ between the CPU and the card; which cells do not
    assert exact(got[~nan], expect[~nan])


IN_PLACE = """field a init 0 range -1e30 1e30
field b init 0 range -1e30 1e30
term t
  d b = diffuse(0.1, a) - 0.5 * b + shift(a, 1, 0, -1)
"""


@pytest.mark.parametrize("gather", sorted(g for g in GATHERS if g != "direct"))
def test_flagged_planes_of_an_in_place_term_are_redone_exactly(device, monkeypatch, gather):
    """A term writing b in place (b read only at its own cell) and a read around it through the window: infinities
    and NaN in a flag lanes, which the window takes again after the tile in the exact forms while its in-place
    stores wait. The direct gather (exact forms everywhere) gives the bits."""
    import field_hip

    shape = (40, 9, 11)
    n = int(np.prod(shape))
    rng = np.random.default_rng(17)
    a, b = rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32)
    for i, v in ((4 * 360 + 4 * 40 + 7, np.inf), (0, -np.inf), (7 * 360 + 8 * 40 + 39, np.nan), (10 * 360 + 5, np.inf)):
        a[i] = v
    runs = {}
    for g in ("direct", gather):
        use_gather(monkeypatch, g)
        solver = field_hip.Solver(field_hip.Law(IN_PLACE), shape, 0.25)
        solver.set_field("a", a)
        solver.set_field("b", b)
        for _ in range(3):
            solver.step(0.01)
        runs[g] = solver.get_field("b")
    want, got = runs["direct"], runs[gather]
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert exact(got[~nan], want[~nan])


@pytest.mark.parametrize("name,net", [("net_flux_relu", "g"), ("net_ldiffuse_half", "w")])
def test_new_net_weights_are_taken_up(device, name, net):
    # Weights written after the law was built (what a trainer does) reach the card at the next step: a solver whose
    # weights were replaced runs to the bits of one built from a .net file holding the new weights.
    import field_hip
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from cases import net_text

    meta, init, _ = golden_io.load(name)
    old = meta["nets"][net]
    sizes = [int(v) for v in old.split("sizes ")[1].split("\n")[0].split()]
    act = old.split("act ")[1].split()[0]
    new_text = net_text(net, sizes, act, 777, 0.9)
    weights = np.array([float(v) for v in new_text.split("weights\n")[1].split()], np.float32)
    shift_scale = {k: old.split(k + " ")[1].split("\n")[0] for k in ("in_shift", "in_scale", "out_scale")}
    # the new file keeps the old normalisation (set_net_weights changes weights only)
    lines = new_text.splitlines()
    lines = [("%s %s" % (l.split()[0], shift_scale[l.split()[0]]) if l.split() and l.split()[0] in shift_scale else l) for l in lines]
    new_meta = dict(meta, nets={net: "\n".join(lines) + "\n"})

    def run(m, swap):
        nx, ny, nz, h = m["grid"]
        s = field_hip.Solver(golden_io.make_law(field_hip, m), (nx, ny, nz), h)
        for f, v in init.items():
            s.set_field(f, v)
        s.refresh_derived()
        s.set_lazy_readback(True)
        s.check_contracts_every(1 << 30)
        s.step(m["dt"])
        if swap:
            s.set_net_weights(net, weights)
        s.run(m["dt"], 3)
        return {f: s.get_field(f) for f in init}

    changed, fresh, kept = run(meta, True), None, run(meta, False)
    # the reference run: new weights from the start would differ in the first step, so compare from the same point
    nx, ny, nz, h = meta["grid"]
    s = field_hip.Solver(golden_io.make_law(field_hip, new_meta), (nx, ny, nz), h)
    s0 = field_hip.Solver(golden_io.make_law(field_hip, meta), (nx, ny, nz), h)
    for f, v in init.items():
        s0.set_field(f, v)
    s0.refresh_derived()
    s0.step(meta["dt"])
    for f in init:
        s.set_field(f, s0.get_field(f))
    s.refresh_derived()
    s.set_lazy_readback(True)
    s.check_contracts_every(1 << 30)
    s.run(meta["dt"], 3)
    fresh = {f: s.get_field(f) for f in init}
    for f in init:
        assert np.array_equal(changed[f].view(np.uint32), fresh[f].view(np.uint32)), f
    assert any(not np.array_equal(changed[f], kept[f]) for f in init)


@pytest.mark.parametrize("lazy", [False, True])
def test_idle_terms_are_not_dispatched(device, lazy):
    """A term whose every rate is provably zero is skipped, as the CPU engine skips it: with no fire anywhere, `spread`
    and `settle` stay idle (and `ash`, all -0, keeps its sign: the golden case holds the CPU's bits)."""
    import field_hip

    meta, init, final = golden_io.load("idle_cold")
    _, solver = run_engine(field_hip, meta, init, lazy=lazy, contracts=1000 if lazy else 1)
    idle = {t["name"]: t["idle"] for t in solver.profile()["terms"]}
    assert idle == {"spread": True, "heat": False, "settle": True}
    for field, expected in final.items():
        assert exact(solver.get_field(field), expected), field


@pytest.mark.parametrize("name", ["idle_cold", "idle_spark"])
def test_run_equals_steps_when_terms_wake_or_sleep(device, name):
    """Solver.run replays graphs only while no term is idle; the state after a run is the state after the same steps
    one at a time, and the reference's."""
    import field_hip

    meta, init, final = golden_io.load(name)
    nx, ny, nz, h = meta["grid"]
    s = field_hip.Solver(golden_io.make_law(field_hip, meta), (nx, ny, nz), h)
    for f, values in init.items():
        s.set_field(f, values)
    s.set_lazy_readback(True)
    s.check_contracts_every(1000)
    s.run(meta["dt"], meta["steps"])
    for field, expected in final.items():
        assert exact(s.get_field(field), expected), field


def test_many_term_world_law_runs_as_the_reference(device):
    """The valley law (12 fields, 14 terms, the world law's op mix) on the 48x40x24 scene, 12 steps: the reference
    holds its CUDA backend to its CPU there to 1e-4 and two runs to the same digest (src/tests/cuda_parity_smoke.cpp:
    83-121); here a lazy Solver.run (graph replays where no contract falls due) must give the reference CPU's bits,
    and so must a second run."""
    import field_hip

    meta, init, final = golden_io.load("valley")
    nx, ny, nz, h = meta["grid"]
    digests = []
    for _ in range(2):
        s = field_hip.Solver(golden_io.make_law(field_hip, meta), (nx, ny, nz), h)
        for f, values in init.items():
            s.set_field(f, values)
        s.refresh_derived()
        s.set_lazy_readback(True)
        s.check_contracts_every(5)
        s.run(meta["dt"], meta["steps"])
        for field, expected in final.items():
            got = s.get_field(field)
            assert golden_io.same(meta, got, expected), f"{field}: max |diff| {np.max(np.abs(got - expected))}"
        digests.append(s.digest())
    assert digests[0] == digests[1]


@pytest.mark.parametrize("nx,window", [(144, False), (192, True), (256, True), (132, False)])
def test_a_window_only_where_rows_fill_its_tiles(device, monkeypatch, nx, window):
    """The gather a term gets: a window only when a row fills at least 90% of its tiles' width (64 cells with a cell a
    lane, 256 with four), the direct gather otherwise; both give the reference's bits, so this pins the choice itself."""
    import field_hip

    for key in ("FIELD_HIP_DIRECT", "FIELD_HIP_TILE", "FIELD_HIP_ZCHUNK", "FIELD_HIP_REMAP", "FIELD_HIP_STREAM"):
        monkeypatch.delenv(key, raising=False)
    law = field_hip.Law("field T init 0\nfield k init 1 static\nterm t\n  d T = diffuse(k, T) + 0.1 * lap(T)\n")
    s = field_hip.Solver(law, (nx, 12, 10), 0.5)
    s.set_field("T", np.random.default_rng(nx).random(nx * 120, dtype=np.float32))
    s.step(0.01)
    kernel = s.profile()["terms"][0]["kernel"]
    assert ("_w" in kernel) == window, (nx, kernel)
