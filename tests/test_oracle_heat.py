"""The ORACLE's conduction restatement against the reference CPU engine's own outputs and against the mathematics."""
import math

import numpy as np
import pytest

import golden_io
from oracle import heat


@pytest.mark.parametrize("name", golden_io.names("heat_"))
def test_oracle_reproduces_reference_bits(name):
    meta, init, final = golden_io.load(name)
    nx, ny, nz, h = meta["grid"]
    k = meta["params"]["k"]
    out = heat.run(init["T"], (nx, ny, nz), k, h, meta["dt"], meta["steps"], coeff=0.15)
    # bit for bit: the restatement computes what the reference computes, in the same order
    assert np.array_equal(out.view(np.uint32), final["T"].view(np.uint32)), \
        f"max |diff| {np.max(np.abs(out - final['T']))}"


def test_a_plausible_mistake_is_caught():
    # The same step with the faces summed in another order (z first) differs from the reference in its last bits
    # somewhere: the golden check above is sensitive to the arithmetic's order, not only to its value.
    meta, init, final = golden_io.load("heat_20x14x9")
    nx, ny, nz, h = meta["grid"]
    k = meta["params"]["k"]
    t = init["T"].reshape(nz, ny, nx).astype(np.float64)
    kf = 2.0 * k * k / (k + k)
    acc = np.zeros_like(t)
    for axis in (0, 1, 2):  # z, y, x: the wrong order
        for d in (-1, 1):
            nb = np.roll(t, -d, axis=axis)
            idx = [slice(None)] * 3
            idx[axis] = -1 if d == 1 else 0
            nb[tuple(idx)] = t[tuple(idx)]
            acc = acc + kf * (nb - t)
    rate = (acc / (h * h)).astype(np.float32)
    sub = meta["dt"] / heat.substeps(meta["dt"], 0.15, k, h)
    wrong = (t.astype(np.float32) + (rate.astype(np.float64) * sub).astype(np.float32)).ravel()
    right = heat.substep(init["T"], (nx, ny, nz), k, h, sub)
    assert not np.array_equal(wrong.view(np.uint32), right.view(np.uint32))


def test_substeps_follow_max_dt():
    h, k = 1.0 / 64, 0.1
    assert heat.substeps(0.4 * h * h / k, 0.15, k, h) == 3
    assert heat.substeps(0.1 * h * h / k, 0.15, k, h) == 1
    assert heat.substeps(1e9, 0.15, k, h) == 4096  # capped


def test_conserves_to_rounding():
    rng = np.random.default_rng(5)
    t = rng.random(24 * 20 * 16).astype(np.float32)
    before, mag = heat.total(t)
    out = heat.run(t, (24, 20, 16), 0.3, 1.0, 0.4, 10)
    after, _ = heat.total(out)
    assert abs(after - before) <= 1e-5 * (mag + 1.0) * math.sqrt(3 * 10)


@pytest.mark.parametrize("n", [16, 32, 64])
def test_cosine_mode_decays_at_the_discrete_rate(n):
    # T = cos(pi x) on [0,1] with zero-flux ends is an eigenvector of the discrete operator (the edge cell repeating
    # is the mirror ghost): each substep multiplies it by 1 - dt k (2 - 2 cos(pi h)) / h^2, exactly up to float rounding.
    h, k = 1.0 / n, 0.1
    x = (np.arange(n) + 0.5) * h
    t = np.tile(np.cos(np.pi * x), (2, 2, 1)).ravel().astype(np.float32)  # nz = 2, ny = 2, nx = n
    dt = 0.1 * h * h / k
    out = heat.run(t, (n, 2, 2), k, h, dt, 20)
    factor = (1.0 - dt * k * (2.0 - 2.0 * math.cos(math.pi * h)) / (h * h)) ** 20
    assert np.max(np.abs(out - factor * t)) < 2e-6


def test_order_two_against_the_exact_solution():
    # zero-flux cosine mode, fixed t_end, dt = 0.1 h^2/k: the error against exp(-pi^2 k t) cos(pi x) falls as h^2
    k, t_end = 0.1, 0.5
    errs = []
    for n in (16, 32, 64):
        h = 1.0 / n
        x = (np.arange(n) + 0.5) * h
        steps = int(round(t_end / (0.1 * h * h / k)))
        dt = t_end / steps
        t = np.tile(np.cos(np.pi * x), (1, 1, 1)).ravel().astype(np.float32)
        out = heat.run(t, (n, 1, 1), k, h, dt, steps, coeff=0.0)
        exact = math.exp(-math.pi ** 2 * k * t_end) * np.cos(np.pi * x)
        errs.append(float(np.sqrt(np.mean((out - exact) ** 2))))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(1.8 < o < 2.2 for o in orders), (errs, orders)
