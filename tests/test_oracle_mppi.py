"""The MPPI oracle (oracle/mppi.py) pinned by closed forms: constant-rate dynamics, the softmin update's limits, the
noise's statistics and determinism."""
import numpy as np

from oracle import mppi

F, HID, H, K = 5, 32, 6, 64


def constant_rate_params(rate):
    p = mppi.make_params(F, HID, 1)
    p["W3"] = np.zeros((F, HID))
    p["b3"] = np.asarray(rate, np.float64)
    return p


def test_constant_rate_rollout_is_the_closed_form():
    rate = np.array([0.5, -1.0, 0.25, 0.0, 2.0])
    p = constant_rate_params(rate)
    p["scale"] = np.array([1.0, 2.0, 0.5, 1.0, 1.0])
    obs = np.array([0.1, -0.2, 0.3, 0.0, -0.5])
    dt = 0.1
    acts = np.random.default_rng(2).uniform(-1, 1, (K, H, 2))
    J = mppi.rollout(p, obs, acts, dt)
    want = np.zeros(K)
    for t in range(H):
        s = obs + (t + 1) * dt * rate * p["scale"]
        want += (p["weight"] * (s - p["target"]) ** 2).sum() + p["action_weight"] * (acts[:, t, :] ** 2).sum(axis=1)
    np.testing.assert_allclose(J, want, rtol=1e-12)


def test_update_limits():
    acts = np.random.default_rng(3).uniform(-1, 1, (K, H, 2))
    # equal costs: every future weighs the same
    np.testing.assert_allclose(mppi.update(np.full(K, 2.0), acts), acts.mean(axis=0), rtol=1e-12)
    # two cost levels: one future at 0, the rest at 100; lambda = (mean - least) / 4, so each of the others weighs
    # exp(-100 / lambda) against the cheap one's 1
    J = np.full(K, 100.0)
    J[17] = 0.0
    lam = 0.25 * 100.0 * (K - 1) / K
    r = np.exp(-100.0 / lam)
    others = acts.sum(axis=0) - acts[17]
    np.testing.assert_allclose(mppi.update(J, acts), (acts[17] + r * others) / (1 + (K - 1) * r), rtol=1e-12)


def test_future_zero_is_the_plan_and_noise_is_standard_normal():
    mean = np.random.default_rng(4).uniform(-0.5, 0.5, (H, 2))
    a = mppi.actions(mean, 9, 3, K, H, 0.4)
    np.testing.assert_allclose(a[0], mean, atol=1e-7)
    assert np.all(np.abs(a) <= 1.0)
    g = np.array([mppi.gauss(mppi.noise_key(5, 11, k, t, d)) for k in range(1, 400) for t in range(8) for d in range(2)])
    assert abs(g.mean()) < 0.05 and abs(g.std() - 1.0) < 0.05
    assert mppi.gauss(mppi.noise_key(5, 11, 3, 2, 1)) == mppi.gauss(mppi.noise_key(5, 11, 3, 2, 1))
    assert mppi.gauss(mppi.noise_key(5, 11, 3, 2, 1)) != mppi.gauss(mppi.noise_key(5, 12, 3, 2, 1))
