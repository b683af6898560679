"""The engine's guard on a learned scheme, on the card (the reference's gate: src/tests/guarded_smoke.cpp:40-88).

A conduction law carries a learned scheme (`ldiffuse` through an untrained net, whose last layer is zero, so it is the
explicit scheme) with a learned max_dt twice the classical one. At a step of twice the classical stable step the
untrained scheme is unstable: every step that breaks T's declared range or its total must be undone on the card and
taken by the classical scheme, so no violation reaches the caller, T stays in its range, the result stays near a
purely classical run, and the profile counts the fallbacks. With the learned switch off nothing learned runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GUARDED_LAW = """\
field T init 290 range 0 6000 unit K
param k_heat = 1.0
net cond_ld in 5 hidden 32 32 out 4 act tanh file none
term conduction
  d T = diffuse(k_heat, T)
  conserves T
  max_dt = 0.15 * h * h / k_heat
  learned d T = ldiffuse(cond_ld, k_heat, T)
  learned max_dt = 0.3 * h * h / k_heat
"""


def hot_spots(n: int) -> np.ndarray:
    """Two hot discs and a warm bar on a 290 K plate (n x n x 1, h = 1)."""
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    t = np.full((n, n), 290.0)
    t[(x - 30) ** 2 + (y - 40) ** 2 <= 81] = 1500.0
    t[(x - 64) ** 2 + (y - 60) ** 2 <= 144] = 900.0
    t[10:34, 70:90] = 1200.0
    return t.astype(np.float32).ravel()


@pytest.mark.parametrize("lazy", [False, True])
def test_unstable_learned_scheme_is_caught_and_undone(device, lazy):
    import field_hip

    n, dt, steps = 96, 0.3, 40
    print(f"guard: {n}x{n}x1, dt {dt}, {steps} steps, lazy {lazy}", flush=True)
    law = field_hip.Law(GUARDED_LAW)
    learned = field_hip.Solver(law, (n, n, 1), 1.0)
    classical = field_hip.Solver(law, (n, n, 1), 1.0)
    t0 = hot_spots(n)
    for s in (learned, classical):
        s.set_field("T", t0)
        s.set_lazy_readback(lazy)
    learned.set_learned(True)
    fallbacks = held = 0
    for _ in range(steps):
        learned.step(dt)
        classical.step(dt)
        prof = learned.profile()
        fallbacks += prof["fallbacks"]
        held += 1 if prof["terms"][0]["learned"] else 0
        assert learned.violations() == [], learned.violations()
        assert prof["terms"][0]["learned"] != prof["terms"][0]["fell_back"]
    got, ref = learned.get_field("T"), classical.get_field("T")
    print(f"guard: {fallbacks} of {steps} steps undone, {held} held; T in [{got.min()}, {got.max()}], "
          f"worst off the classical run {np.max(np.abs(got - ref))}", flush=True)
    assert fallbacks > 0, "the guard catches the unstable step"
    assert fallbacks + held == steps
    assert got.min() >= 0.0 and got.max() <= 6000.0
    assert np.max(np.abs(got.astype(np.float64) - ref)) < 400.0
    # Conservation holds whichever scheme took each step (both are in flux form).
    assert abs(float(np.sum(got, dtype=np.float64)) - float(np.sum(t0, dtype=np.float64))) < 1e-3 * float(np.sum(t0))


def test_an_undone_step_equals_the_classical_step_bit_for_bit(device):
    """A step whose learned scheme is undone leaves exactly the classical scheme's state: the snapshot restores what
    the term wrote, and the classical rerun starts from it. Each step starts a classical solver from the guarded
    solver's state; on every fallback the two must agree to the bit."""
    import field_hip

    n, dt, steps = 96, 0.3, 30
    law = field_hip.Law(GUARDED_LAW)
    learned = field_hip.Solver(law, (n, n, 1), 1.0)
    classical = field_hip.Solver(law, (n, n, 1), 1.0)
    learned.set_field("T", hot_spots(n))
    learned.set_learned(True)
    undone = 0
    for _ in range(steps):
        classical.set_field("T", learned.get_field("T"))
        learned.step(dt)
        classical.step(dt)
        if learned.profile()["fallbacks"]:
            undone += 1
            a, b = learned.get_field("T"), classical.get_field("T")
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert undone > 0


def test_learned_off_runs_nothing_learned(device):
    import field_hip

    n = 48
    s = field_hip.Solver(field_hip.Law(GUARDED_LAW), (n, n, 1), 1.0)
    s.set_field("T", hot_spots(96).reshape(96, 96)[:n, :n].ravel().copy())
    s.step(0.3)
    prof = s.profile()
    assert prof["fallbacks"] == 0 and not prof["terms"][0]["learned"] and not prof["terms"][0]["fell_back"]


def test_stable_learned_step_holds_and_matches_the_explicit_scheme(device):
    """At a step inside the learned max_dt the untrained net is the explicit scheme exactly (its last layer is zero):
    the learned step holds (no fallback) and takes one pass where the classical scheme takes two substeps."""
    import field_hip

    n, dt = 48, 0.25
    law = field_hip.Law(GUARDED_LAW)
    s = field_hip.Solver(law, (n, n, 1), 1.0)
    t0 = hot_spots(96).reshape(96, 96)[:n, :n].ravel().copy()
    s.set_field("T", t0)
    s.set_learned(True)
    s.step(dt)
    prof = s.profile()
    assert prof["fallbacks"] == 0 and prof["terms"][0]["learned"]
    assert s.last_substeps() == [1]
    # One explicit pass of dt on the same start, through the plain law with a max_dt that allows it.
    plain = field_hip.Solver(field_hip.Law("field T init 290\nparam k_heat = 1.0\nterm c\n  d T = diffuse(k_heat, T)\n"),
                             (n, n, 1), 1.0)
    plain.set_field("T", t0)
    plain.step(dt)
    assert np.max(np.abs(s.get_field("T") - plain.get_field("T"))) <= 1e-3
