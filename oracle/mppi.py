"""ORACLE for the MPPI planner (include/fe_mppi.h): a NumPy restatement, test infrastructure only (oracle/__init__.py).

Follows the reference's planner (src/field/mind/WorldModel.cpp:192-240, WorldModel::plan :521-595): futures of the
learned ODE s += dt * f(s, a) * scale under actions a = clamp(plan + sigma * noise, -1, 1) (future 0 without noise), a
cost per future summed over the steps, and the plan's update weighted by exp(-(J - J_min) / lambda), lambda a quarter of
(mean cost - least cost) and at least 1e-4. The cost is this build's weighted quadratic (the card's), and the card's
roundings are kept where it rounds: the layer inputs and the weights to fp16; sums in float64 (the card accumulates in
fp32). The noise is the card's hash (splitmix64's finaliser) through Box-Muller, in float64.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1


def mix64(h: int) -> int:
    h &= M64
    h ^= h >> 30
    h = (h * 0xBF58476D1CE4E5B9) & M64
    h ^= h >> 27
    h = (h * 0x94D049BB133111EB) & M64
    h ^= h >> 31
    return h


def noise_key(agent_id: int, tick: int, k: int, t: int, d: int) -> int:
    return mix64(agent_id * 0x9E3779B97F4A7C15 + 1) ^ mix64(tick * 0xD1B54A32D192ED03 + 2) ^ (k << 24) ^ (t << 8) ^ d


def gauss(key: int) -> float:
    a = mix64(key)
    b = mix64(a ^ 0x632BE59BD9B4E019)
    u1 = ((a >> 40) + 0.5) / 16777216.0
    u2 = ((b >> 40) + 0.5) / 16777216.0
    return float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))


def split(weights, F: int, hid: int):
    """(W1 [hid][F+2], b1, W2 [hid][hid], b2, W3 [F][hid], b3) from the flat .net layout."""
    w = np.asarray(weights, np.float64)
    sizes = [(hid, F + 2), (hid,), (hid, hid), (hid,), (F, hid), (F,)]
    out, at = [], 0
    for s in sizes:
        n = int(np.prod(s))
        out.append(w[at:at + n].reshape(s))
        at += n
    return out


def f16(x):
    return np.asarray(x, np.float64).astype(np.float16).astype(np.float64)


def actions(mean, agent_id: int, tick: int, K: int, H: int, sigma: float):
    """K x H x 2 actions of one agent (future 0 is the plan itself)."""
    a = np.empty((K, H, 2))
    for k in range(K):
        for t in range(H):
            for d in range(2):
                e = 0.0 if k == 0 else gauss(noise_key(agent_id, tick, k, t, d))
                a[k, t, d] = min(max(float(np.float32(mean[t, d] + np.float32(sigma) * np.float32(e))), -1.0), 1.0)
    return a


def rollout(params, obs, acts, dt: float):
    """Costs (K,) of the futures of one agent under acts (K x H x 2)."""
    W1, b1, W2, b2, W3, b3 = (f16(params["W1"]), params["b1"], f16(params["W2"]), params["b2"], f16(params["W3"]), params["b3"])
    F = obs.size
    K, H, _ = acts.shape
    s = np.tile(np.asarray(obs, np.float64), (K, 1))
    J = np.zeros(K)
    for t in range(H):
        x = np.concatenate([s, acts[:, t, :]], axis=1)
        x = f16((x - params["shift"]) * params["inscale"])
        h1 = f16(np.tanh(x @ W1.T + b1))
        h2 = f16(np.tanh(h1 @ W2.T + b2))
        rate = h2 @ W3.T + b3
        s = s + dt * rate * params["scale"]
        J = J + ((s - params["target"]) ** 2 * params["weight"]).sum(axis=1) + params["action_weight"] * (acts[:, t, :] ** 2).sum(axis=1)
    return J


def update(J, acts):
    """The new plan (H x 2) from the futures' costs and actions."""
    lo = J.min()
    lam = max(1e-4, 0.25 * (J.mean() - lo))
    w = np.exp(-(J - lo) / lam)
    return np.tensordot(w, acts, axes=(0, 0)) / w.sum()


def plan(params, obs, ids, mean, tick: int, sigma: float, dt: float, K: int):
    """(new mean agents x H x 2, costs agents x K)."""
    agents, H = mean.shape[0], mean.shape[1]
    out, costs = np.empty_like(mean, dtype=np.float64), np.empty((agents, K))
    for i in range(agents):
        a = actions(mean[i], int(ids[i]), tick, K, H, sigma)
        costs[i] = rollout(params, obs[i], a, dt)
        out[i] = update(costs[i], a)
    return out, costs


def make_params(F: int, hid: int, seed: int, rate_scale: float = 0.5):
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (hid, F + 2)) * np.sqrt(3.0 / (F + 2))
    W2 = rng.uniform(-1, 1, (hid, hid)) * np.sqrt(3.0 / hid)
    W3 = rng.uniform(-1, 1, (F, hid)) * rate_scale * np.sqrt(3.0 / hid)
    b1, b2, b3 = rng.uniform(-0.1, 0.1, hid), rng.uniform(-0.1, 0.1, hid), rng.uniform(-0.05, 0.05, F)
    flat = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)
    return {"flat": flat, "W1": W1.astype(np.float32).astype(np.float64), "b1": b1.astype(np.float32).astype(np.float64),
            "W2": W2.astype(np.float32).astype(np.float64), "b2": b2.astype(np.float32).astype(np.float64),
            "W3": W3.astype(np.float32).astype(np.float64), "b3": b3.astype(np.float32).astype(np.float64),
            "shift": np.zeros(F + 2), "inscale": np.ones(F + 2), "scale": np.ones(F),
            "target": rng.uniform(-0.5, 0.5, F).astype(np.float32).astype(np.float64), "weight": np.ones(F), "action_weight": 0.01}
