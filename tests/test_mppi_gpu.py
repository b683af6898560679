"""The MPPI planner on the card (include/fe_mppi.h) against its oracle (oracle/mppi.py, pinned by closed forms).

The card runs the dynamics on fp16 MFMA with fp32 accumulation and its own transcendentals; the oracle rounds where
the card rounds and sums in float64, so futures' costs agree to 2e-3 relative and the plans to 2e-3 absolute. Two runs
give the same plan bit for bit."""
import numpy as np
import pytest

from oracle import mppi

pytestmark = pytest.mark.gpu

F, HID, H, K = 6, 32, 8, 128


def planner(field_hip, p):
    pl = field_hip.MppiPlanner(F, HID, H, K)
    pl.set_dynamics(p["flat"], p["shift"], p["inscale"], p["scale"])
    pl.set_cost(p["target"], p["weight"], p["action_weight"])
    return pl


def test_plan_matches_the_oracle_and_repeats(device):
    import field_hip

    p = mppi.make_params(F, HID, 7)
    rng = np.random.default_rng(8)
    agents = 5
    obs = rng.uniform(-1, 1, (agents, F)).astype(np.float32)
    ids = np.array([3, 17, 99, 1000, 4242], np.uint32)
    mean = rng.uniform(-0.3, 0.3, (agents, H, 2)).astype(np.float32)
    pl = planner(field_hip, p)
    got, costs = pl.plan(obs, ids, mean, tick=12, sigma=0.3, dt=0.1, costs=True)
    want, wcost = mppi.plan(p, obs.astype(np.float64), ids, mean.astype(np.float64), 12, 0.3, 0.1, K)
    rel = np.abs(costs - wcost) / np.maximum(np.abs(wcost), 1e-3)
    assert rel.max() < 2e-3, rel.max()
    assert np.max(np.abs(got - want)) < 2e-3, np.max(np.abs(got - want))
    again, _ = pl.plan(obs, ids, mean, tick=12, sigma=0.3, dt=0.1)
    assert np.array_equal(got.view(np.uint32), again.view(np.uint32))
    other, _ = pl.plan(obs, ids, mean, tick=13, sigma=0.3, dt=0.1)
    assert not np.array_equal(got, other)  # another tick, other noise


def test_planning_lowers_the_plans_cost(device):
    """Planning repeatedly from where it left off: the plan's own cost (future 0) falls."""
    import field_hip

    p = mppi.make_params(F, HID, 9, rate_scale=2.0)
    obs = np.zeros((1, F), np.float32)
    pl = planner(field_hip, p)
    mean = np.zeros((1, H, 2), np.float32)
    first = None
    for it in range(12):
        mean, costs = pl.plan(obs, np.array([5], np.uint32), mean, tick=it, sigma=0.5, dt=0.2, costs=True)
        if first is None:
            first = float(costs[0, 0])
    _, costs = pl.plan(obs, np.array([5], np.uint32), mean, tick=99, sigma=0.5, dt=0.2, costs=True)
    assert float(costs[0, 0]) < 0.8 * first, (first, float(costs[0, 0]))


def test_refuses_shapes_it_does_not_take(device):
    import field_hip

    with pytest.raises(field_hip.FieldHipError):
        field_hip.MppiPlanner(F, 24, H, K)  # hidden not a multiple of 16
    with pytest.raises(field_hip.FieldHipError):
        field_hip.MppiPlanner(F, HID, H, 100)  # samples not a multiple of 16
