#!/usr/bin/env python3
"""Throughput of the Field Engine's MI355X backend on its headline workload: the fused 7-point conduction step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256]

One "step" is one step of the law `d T = diffuse(k, T)` (conserves T, max_dt = 0.15 h^2/k) on an N^3 fp32 lattice,
taken by the engine (law text -> IR -> generated gfx950 kernel -> HIP backend) at dt = 0.4 h^2/k / 3, which is one
fused substep: one read and one write of every cell. The field is seeded uniform [0,1) and resident in HBM; readback
is lazy and contracts are off in the timed steps (the same step with contracts on is reported beside it). The unit is
cell-updates/s. Also measured: the handwritten kernel the engine must equal to the bit (and whether it does), a copy
kernel (the bandwidth roof), the same step at 512^3 (larger than the 256 MB Infinity Cache), and the reference's own
OpenMP CPU engine on this host's cores on a bounded sample. Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "online-physics-cuda_amd"))
sys.path.insert(0, ROOT)

import field_hip  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
BYTES_PER_CELL = 8.0   # one float read + one float written per cell-update: the algorithmic minimum

HEAT_LAW = """field T init 0 range -1000000 1000000
param k = 0.1
term conduction
  d T = diffuse(k, T)
  conserves T
  max_dt = 0.15 * h * h / k
"""


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def time_launches(launch, steps: int, warmup: int):
    """Whole-run card time of `steps` back-to-back calls (ms per call) and the median of per-call event times."""
    for _ in range(warmup):
        launch()
    field_hip.sync()
    ev = field_hip.Events()
    ev.begin()
    for _ in range(steps):
        launch()
    whole = ev.end() / steps
    per = []
    for _ in range(min(steps, 50)):
        ev.begin()
        launch()
        per.append(ev.end())
    return whole, float(np.median(per))


def conduction(n: int, steps: int, warmup: int, seed: int = 2026, verify_steps: int = 3):
    cells = n ** 3
    k, h = 0.1, 1.0 / n
    sub_dt = 0.4 * h * h / k / 3
    t0 = np.random.default_rng(seed).random(cells, dtype=np.float32)
    out = {"n": n, "cells": cells}

    # The roof: a copy of the same bytes.
    a = field_hip.DeviceArray.from_host(t0)
    b = field_hip.DeviceArray(cells)
    _, copy_ms = time_launches(lambda: field_hip.copy_f32(a, b), steps, warmup)
    out["copy_ms"] = copy_ms
    out["copy_gbs"] = BYTES_PER_CELL * cells / (copy_ms * 1e6)

    # The handwritten kernel.
    bufs = [a, b]

    def hand():
        field_hip.heat_handwritten(bufs[0], bufs[1], n, n, n, k, h, sub_dt)
        bufs.reverse()

    hand_whole, hand_ms = time_launches(hand, steps, warmup)
    out.update(hand_ms_per_step=hand_whole, hand_kernel_ms=hand_ms, hand_gbs=BYTES_PER_CELL * cells / (hand_ms * 1e6))

    # The engine: the law compiled, generated and run by the HIP backend.
    law = field_hip.Law(HEAT_LAW)
    solver = field_hip.Solver(law, (n, n, n), h)
    solver.set_field("T", t0)
    solver.set_lazy_readback(True)
    solver.check_contracts_every(1 << 30)
    solver.step(sub_dt)  # the first step uploads the field and takes the contracts' step
    # The step loop as a caller runs it: Solver.run (fe_solver_run), pairs of steps replayed from a captured graph.
    # Warm: the same run first (it captures the graphs the timed run replays), twice when odd so the fields are back
    # where the captured runs start.
    solver.run(sub_dt, steps)
    if steps % 2:
        solver.run(sub_dt, steps)
    field_hip.sync()
    ev = field_hip.Events()
    ev.begin()
    solver.run(sub_dt, steps)
    eng_whole = ev.end() / steps
    # One fe_solver_step call a step (host launches, per-term events).
    eng_eager, _ = time_launches(lambda: solver.step(sub_dt), steps, warmup)
    out["engine_ms_per_step_eager"] = eng_eager
    per_kernel = []
    for _ in range(min(steps, 50)):
        solver.step(sub_dt)
        per_kernel.append(solver.profile()["terms"][0]["ms"])
    eng_kernel = float(np.median(per_kernel))
    out["engine_kernel"] = solver.profile()["terms"][0]["kernel"]
    out.update(engine_ms_per_step=eng_whole, engine_kernel_ms=eng_kernel,
               engine_gbs=BYTES_PER_CELL * cells / (eng_kernel * 1e6),
               engine_cell_updates_per_s=cells / (eng_whole * 1e-3))
    solver.check_contracts_every(1)
    with_contracts = []
    for _ in range(min(steps, 20)):
        t = time.perf_counter()
        solver.step(sub_dt)
        with_contracts.append((time.perf_counter() - t) * 1e3)
    out["engine_ms_per_step_contracts_on"] = float(np.median(with_contracts))
    out["engine_contracts_violations"] = solver.violations()

    # Bit identity: the same steps from the same start, engine and handwritten.
    s2 = field_hip.Solver(law, (n, n, n), h)
    s2.set_field("T", t0)
    s2.set_lazy_readback(True)
    s2.check_contracts_every(1 << 30)
    for _ in range(verify_steps):
        s2.step(sub_dt)
    engine_t = s2.get_field("T")
    a.upload(t0)
    bufs = [a, b]
    for _ in range(verify_steps):
        hand()
    hand_t = bufs[0].download()
    out["engine_equals_handwritten_bits"] = bool(np.array_equal(engine_t.view(np.uint32), hand_t.view(np.uint32)))
    out["verify_steps"] = verify_steps
    return out


ADI_SHAPES = ((128, 64), (256, 128), (512, 256))


def heston_options(n: int, seed: int = 2026):
    """n options: lattice S x v drawn from 128x64, 256x128, 512x256 (a third each, log-uniform in cells); kappa in
    [0.5, 5], theta [0.01, 0.1], sigma [0.1, 0.8], rho [-0.9, -0.1], v0 [0.01, 0.1], strike [80, 120], maturity
    [0.25, 2], r = 0.02, seeded uniform."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, 3, n)
    nx = np.array([ADI_SHAPES[i][0] for i in idx], np.int32)
    ny = np.array([ADI_SHAPES[i][1] for i in idx], np.int32)
    opts = np.stack([rng.uniform(0.5, 5, n), rng.uniform(0.01, 0.1, n), rng.uniform(0.1, 0.8, n), rng.uniform(-0.9, -0.1, n),
                     np.full(n, 0.02), rng.uniform(80, 120, n), rng.uniform(0.25, 2, n), rng.uniform(0.01, 0.1, n)], 1)
    return nx, ny, opts


def heston_adi(n: int, steps: int, sample_per_class: int, cpu: bool):
    """The batch on the card (Craig-Sneyd, theta 1/2, fp32): tuned, every step captured in one graph, one launch timed.
    With `cpu`, a sample of the same mix is run by the oracle in FP64 (the error) and in FP32 on the host's cores (the
    baseline)."""
    nx, ny, opts = heston_options(n)
    b = field_hip.HestonBatch(nx, ny, opts, steps, 0, 0.5)
    cells = b.cells
    t = time.time()
    tuned = b.tune(2)
    tune_s = time.time() - t
    b.capture(steps)
    b.launch()          # warm: one whole run
    b.sync()
    b.reset()
    field_hip.sync()
    ev = field_hip.Events()
    ev.begin()
    b.launch()
    run_ms = ev.end()
    prices = b.prices()
    out = {"options": n, "steps": steps, "cells": cells, "scheme": "Craig-Sneyd theta 1/2, fp32, delta form",
           "run_ms": run_ms, "solves_per_s": n / (run_ms * 1e-3), "cell_updates_per_s": cells * steps / (run_ms * 1e-3),
           "achieved_gbs_8B": 8.0 * cells * steps / (run_ms * 1e6), "launches_in_graph": 4 * 3 * steps,
           "tuned": tuned, "tune_s": tune_s}
    out["frac_hbm_8B"] = out["achieved_gbs_8B"] / HBM_PEAK_GBS
    if cpu:
        sample = np.concatenate([np.flatnonzero(nx == s[0])[:sample_per_class] for s in ADI_SHAPES])
        base = cpu_baseline_adi(nx[sample], ny[sample], opts[sample], steps)
        err = np.abs(prices[sample] - base["fp64_prices"])
        out["error_vs_fp64"] = {"options": int(sample.size), "max_abs": float(err.max()),
                                "max_rel": float(np.max(err / np.maximum(np.abs(base["fp64_prices"]), 1e-3))),
                                "rms_abs": float(np.sqrt(np.mean(err ** 2)))}
        out["cpu_baseline"] = base["baseline"]
        out["speedup_vs_cpu"] = out["solves_per_s"] / base["baseline"]["value"]
    return out


def cpu_baseline_adi(nx, ny, opts, steps):
    """The ORACLE's restatement of the same scheme on the host: FP64 (the prices the card's are judged against) and
    FP32 on all cores (the baseline), on a sample of the batch's mix."""
    from oracle import heston
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ref, _, _ = heston.batch64(nx, ny, opts, steps)
    t = time.time()
    heston.batch32(nx, ny, opts, steps)
    wall = time.time() - t
    return {"fp64_prices": ref,
            "baseline": {"value": len(nx) / wall, "unit": "solves/s", "cores": cores, "kind": "port",
                         "sample": f"oracle FP32 restatement of the same scheme (OpenMP over options), {len(nx)} options of the "
                                   f"batch's mix, {steps} steps, {wall:.1f} s"}}


DIRICHLET_LAW = """field T init 0 range -1000000 1000000
field nb init 0 range 0 6 static
param k = 0.1
term conduction
  d T = diffuse(k, T) - 2 * k * nb * T / (h * h)
  max_dt = 0.15 * h * h / k
"""

BLACK_SCHOLES_LAW = """field V init 0 range -1 1000000
param sig = 0.2
param r = 0.05
term black_scholes
  d V = 0.5 * sig * sig * (shift(V, 1, 0, 0) - 2 * V + shift(V, -1, 0, 0)) / (h * h) + (r - 0.5 * sig * sig) * gx(V) - r * V
  max_dt = 0.4 * h * h / (sig * sig)
"""


def sine_start(n: int, noise: float, seed: int = 5):
    """prod sin(pi x_i) at cell centres of [0, 1]^3, plus seeded uniform noise of amplitude `noise`; and the count of
    each cell's boundary faces (zero Dirichlet: each such face holds a ghost of -T)."""
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    t = np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    if noise:
        t = t + noise * np.random.default_rng(seed).random(t.shape)
    i = np.arange(n)
    edge = ((i == 0) | (i == n - 1)).astype(np.float32)
    nb = edge[None, None, :] + edge[None, :, None] + edge[:, None, None]
    return t.astype(np.float32).reshape(-1), nb.astype(np.float32).reshape(-1)


def dirichlet_heat():
    """Zero-Dirichlet 3-D conduction (k = 0.1, dt = 0.4 h^2/k, substeps from max_dt = 0.15 h^2/k): the error against the
    decaying sine mode exp(-3 pi^2 k t) prod sin(pi x) and its order under refinement 16 -> 32 -> 64 at a fixed end
    time (200 steps of 64^3), and the 64^3 run with noise: a fixed-order compensated total every step."""
    k = 0.1
    t_end = 200 * 0.4 * (1.0 / 64) ** 2 / k
    errors = {}
    for n in (16, 32, 64):
        h = 1.0 / n
        steps = int(np.ceil(t_end / (0.4 * h * h / k) - 1e-9))
        t0, nb = sine_start(n, 0.0)
        s = field_hip.Solver(field_hip.Law(DIRICHLET_LAW), (n, n, n), h)
        s.set_field("T", t0)
        s.set_field("nb", nb)
        s.set_lazy_readback(True)
        s.check_contracts_every(1 << 30)
        for _ in range(steps):
            s.step(t_end / steps)
        exact = np.exp(-3 * np.pi ** 2 * k * t_end) * t0.astype(np.float64)
        errors[n] = float(np.max(np.abs(s.get_field("T") - exact)))
    orders = [float(np.log2(errors[16] / errors[32])), float(np.log2(errors[32] / errors[64]))]
    n, h = 64, 1.0 / 64
    dt = 0.4 * h * h / k
    t0, nb = sine_start(n, 1e-3)
    s = field_hip.Solver(field_hip.Law(DIRICHLET_LAW), (n, n, n), h)
    s.set_field("T", t0)
    s.set_field("nb", nb)
    s.set_lazy_readback(True)
    s.check_contracts_every(1 << 30)
    totals = []
    ev = field_hip.Events()
    ev.begin()
    for _ in range(200):
        s.step(dt)
        totals.append(s.total("T")[0])
    ms = ev.end()
    out = {"errors_max_vs_exact": errors, "orders": orders, "substeps_per_step": s.last_substeps(),
           "steps": 200, "cells": n ** 3, "ms_per_step_with_total": ms / 200,
           "total_first_last": [totals[0], totals[-1]], "totals_decreasing": bool(np.all(np.diff(totals) < 0))}
    out.update(reference_check("dirichlet_heat_64", s.get_field("T")))
    return out


def black_scholes_start(nx: int = 4096, planes: int = 512):
    """x = ln S on ln 100 +- 4 (h = 8 / nx), the put payoff with strike 100 averaged over each cell, the same line in
    every plane; returns (V0, h, dt) with dt = max_dt = 0.4 h^2 / sig^2 at sig = 0.2, computed as the law computes it
    (sig * sig is 0.04000000000000001 in double: a dt of 0.4 h^2 / 0.04 would be a hair above max_dt, two substeps)."""
    h = 8.0 / nx
    x0 = np.log(100.0) - 4.0
    edges = x0 + h * np.arange(nx + 1)
    lo, hi = edges[:-1], edges[1:]
    a = np.minimum(hi, np.log(100.0))
    pay = np.where(a > lo, (100.0 * (a - lo) - (np.exp(a) - np.exp(lo))) / h, 0.0)
    sig = 0.2
    return np.tile(pay, planes).astype(np.float32), h, 0.4 * h * h / (sig * sig)


def black_scholes(steps: int):
    """The Black-Scholes law in x = ln S on 4096 cells x 512 planes (every plane an option), sig 0.2, r 0.05, put K =
    100 averaged over each cell, one substep a step at dt = max_dt = 0.4 h^2 / sig^2."""
    nx, planes = 4096, 512
    v0, h, dt = black_scholes_start(nx, planes)
    s = field_hip.Solver(field_hip.Law(BLACK_SCHOLES_LAW), (nx, 1, planes), h)
    s.set_field("V", v0)
    s.set_lazy_readback(True)
    s.check_contracts_every(1 << 30)
    for _ in range(10):
        s.step(dt)
    mid = s.get_field("V")
    # Timed as a caller runs it: Solver.run (captured graphs of whole steps), after the same run once to capture them.
    s.run(dt, steps)
    field_hip.sync()
    ev = field_hip.Events()
    ev.begin()
    s.run(dt, steps)
    ms = ev.end() / steps
    cells = nx * planes
    out = {"cells": cells, "ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3), "gbs_8B": 8.0 * cells / (ms * 1e6),
           "kernel": s.profile()["terms"][0]["kernel"]}
    out.update(reference_check("black_scholes_4096x512", mid))
    return out


IMPLICIT_X_LAW = """field T init 0 range -1000000 1000000
param k = 0.01
term crank_nicolson
  let a = 0.5 * k * dt / (h * h)
  d T = (tridiag_x(0 - a, 1 + 2 * a, 0 - a, T + a * (shift(T, 1, 0, 0) - 2 * T + shift(T, -1, 0, 0))) - T) / dt
"""

IMPLICIT_YZ_LAW = """field T init 0 range -1000000 1000000
field w init 1 range 0 10 static
param k = 0.05
term implicit_yz
  let a = k * w * dt / (h * h)
  let y1 = tridiag_y(0 - a, 1 + 2 * a, 0 - a, T)
  d T = (tridiag_z(0 - 0.5 * a, 1 + a, 0 - 0.5 * a, y1) - T) / dt
"""


def implicit_start(n: int):
    """The implicit-lines start: T seeded uniform [0, 1), w uniform [0.5, 2)."""
    rng = np.random.default_rng(11)
    return rng.random(n ** 3, dtype=np.float32), (0.5 + 1.5 * rng.random(n ** 3)).astype(np.float32)


def implicit_lines(n: int, steps: int, check_steps: int = 3):
    """Laws whose step solves a tridiagonal system along every lattice line (Thomas in double, the reference CPU's
    arithmetic): Crank-Nicolson along x (one number a coefficient, so c' stays on chip) and an implicit y-then-z step
    with a field-dependent operator, on n^3. Each is held to the reference CPU engine's bits after check_steps steps,
    then timed through Solver.run."""
    h = 1.0 / n
    t0, w0 = implicit_start(n)
    out = {}
    for name, law, dt, fields in (("crank_nicolson_x", IMPLICIT_X_LAW, 0.5 * h * h / 0.01, {"T": t0}),
                                  ("implicit_yz", IMPLICIT_YZ_LAW, 0.5 * h * h / 0.05, {"T": t0, "w": w0})):
        s = field_hip.Solver(field_hip.Law(law), (n, n, n), h)
        for f, v in fields.items():
            s.set_field(f, v)
        s.refresh_derived()
        s.set_lazy_readback(True)
        s.check_contracts_every(1 << 30)
        for _ in range(check_steps):
            s.step(dt)
        rec = reference_check(f"{name}_{n}", s.get_field("T"))
        per_term = []
        for _ in range(5):
            s.step(dt)
            per_term.append(s.profile()["terms"][0]["ms"])
        s.run(dt, steps)
        field_hip.sync()
        ev = field_hip.Events()
        ev.begin()
        s.run(dt, steps)
        ms = ev.end() / steps
        cells = n ** 3
        out[name] = {"cells": cells, "ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3),
                     "term_ms": float(np.median(per_term)), "gbs_8B": 8.0 * cells / (ms * 1e6), **rec}
    return out


def mppi_plan(cpu: bool, plans: int = 20):
    """The world model's MPPI planner (include/fe_mppi.h): 256 agents, 512 futures of 32 steps each, dynamics 16 + 2 ->
    64 -> 64 -> 16 (tanh) on fp16 MFMA; seeded random weights and observations. A plan is one fe_mppi_plan call (it
    uploads the observations and plans, runs the futures and the update, and brings the plans back)."""
    F, HID, H, K, agents = 16, 64, 32, 512, 256
    rng = np.random.default_rng(21)
    w1 = rng.uniform(-1, 1, (HID, F + 2)) * np.sqrt(3.0 / (F + 2))
    w2 = rng.uniform(-1, 1, (HID, HID)) * np.sqrt(3.0 / HID)
    w3 = rng.uniform(-1, 1, (F, HID)) * 0.5 * np.sqrt(3.0 / HID)
    flat = np.concatenate([w1.ravel(), rng.uniform(-0.1, 0.1, HID), w2.ravel(), rng.uniform(-0.1, 0.1, HID), w3.ravel(),
                           rng.uniform(-0.05, 0.05, F)]).astype(np.float32)
    shift, inscale, scale = np.zeros(F + 2, np.float32), np.ones(F + 2, np.float32), np.ones(F, np.float32)
    target, weight = rng.uniform(-0.5, 0.5, F).astype(np.float32), np.ones(F, np.float32)
    pl = field_hip.MppiPlanner(F, HID, H, K)
    pl.set_dynamics(flat, shift, inscale, scale)
    pl.set_cost(target, weight, 0.01)
    obs = rng.uniform(-1, 1, (agents, F)).astype(np.float32)
    ids = np.arange(agents, dtype=np.uint32)
    mean = np.zeros((agents, H, 2), np.float32)
    for tick in range(3):
        mean, _ = pl.plan(obs, ids, mean, tick=tick, sigma=0.3, dt=0.1)
    times = []
    for tick in range(3, 3 + plans):
        t = time.perf_counter()
        mean, _ = pl.plan(obs, ids, mean, tick=tick, sigma=0.3, dt=0.1)
        times.append((time.perf_counter() - t) * 1e3)
    ms = float(np.median(times))
    flop = 2.0 * agents * K * H * ((F + 2) * HID + HID * HID + HID * F)
    out = {"agents": agents, "samples": K, "horizon": H, "features": F, "hidden": HID, "ms_per_plan": ms,
           "futures_per_s": agents * K / (ms * 1e-3), "mfma_tflops": flop / (ms * 1e9),
           "of_fp16_dense_peak": flop / (ms * 1e9) / 2500.0, "plan_finite": bool(np.all(np.isfinite(mean)))}
    if cpu:
        from oracle import mppi as ref
        W1, b1, W2, b2, W3, b3 = ref.split(flat, F, HID)
        params = {"W1": W1, "b1": b1, "W2": W2, "b2": b2, "W3": W3, "b3": b3, "shift": shift.astype(np.float64),
                  "inscale": inscale.astype(np.float64), "scale": scale.astype(np.float64),
                  "target": target.astype(np.float64), "weight": weight.astype(np.float64), "action_weight": 0.01}
        sample = 2
        t = time.perf_counter()
        ref.plan(params, obs[:sample].astype(np.float64), ids[:sample], mean[:sample].astype(np.float64), 99, 0.3, 0.1, K)
        sec = time.perf_counter() - t
        out["cpu_baseline"] = {"value": sample * K / sec, "unit": "futures/s", "cores": 1, "kind": "port",
                               "sample": f"oracle NumPy restatement (float64), {sample} agents x {K} futures x {H} steps, "
                                         f"{sec:.1f} s"}
    return out


def net_train(cpu: bool, steps: int = 50):
    """A law net's trainer (include/fe_nettrain.h) on the world model's dynamics net, 16 + 2 -> 64 -> 64 -> 16 (tanh):
    one Adam step a call (it uploads the rows and targets, runs the tiles, the two fixed-order sums and Adam, and reads
    the loss back), at the world model's batch of 1024 rows and at 65 536 rows; seeded random weights, rows a smooth
    map of the inputs."""
    from oracle import nettrain as ref

    F = 16
    sizes = [F + 2, 64, 64, F]
    flop_row = 2.0 * (3 * ((F + 2) * 64 + 64 * 64 + 64 * F) - (F + 2) * 64)  # forward, weight gradients, deltas below layer 1
    rng = np.random.default_rng(31)
    A = rng.uniform(-1, 1, (F, F + 2)) / np.sqrt(F + 2)
    image = ref.make_image(sizes, 32, shift=0.0, scale=1.0, out_scale=1.0).astype(np.float32)
    out = {"sizes": sizes, "activation": "tanh"}
    for rows in (1024, 65536):
        x = rng.uniform(-1, 1, (rows, F + 2)).astype(np.float32)
        t = np.sin(x @ A.T).astype(np.float32)
        tr = field_hip.NetTrainer(sizes, "tanh")
        tr.set_image(image)
        first = tr.step(x, t)
        for _ in range(3):
            tr.step(x, t)
        times = []
        for _ in range(steps):
            s = time.perf_counter()
            last = tr.step(x, t)
            times.append((time.perf_counter() - s) * 1e3)
        ms = float(np.median(times))
        out[f"rows_{rows}"] = {"ms_per_step": ms, "rows_per_s": rows / (ms * 1e-3), "tflops": flop_row * rows / (ms * 1e9),
                               "first_loss": first, "loss_after": last}
    if cpu:
        rows = 1024
        x = rng.uniform(-1, 1, (rows, F + 2))
        t = np.sin(x @ A.T)
        img, base = image.astype(np.float64), ref.first_weight(sizes)
        m1, m2 = np.zeros(img.size - base), np.zeros(img.size - base)
        n = 0
        s = time.perf_counter()
        while time.perf_counter() - s < 5.0:
            n += 1
            _, g = ref.gradient(sizes, "tanh", img, x, t)
            ref.adam(img, g, m1, m2, n, sizes, 2e-3, 5.0)
        sec = time.perf_counter() - s
        out["cpu_baseline"] = {"value": n * rows / sec, "unit": "rows/s", "cores": 1, "kind": "port",
                               "sample": f"oracle NumPy restatement (float64), {n} Adam steps of {rows} rows, {sec:.1f} s"}
    return out


def hbm_roof(n: int = 512, reps: int = 10):
    """The attainable HBM rate at n^3 floats (512^3: 512 MiB each way, twice the last-level cache): a copy and a read
    swept over loads in flight a lane (1, 2, 4, 8), ordinary or nontemporal accesses, and 2, 4 or 8 blocks of 256
    threads a CU; the best of each, by events over `reps` back-to-back launches, beside the copy kernel the bench uses."""
    cells = n ** 3
    a = field_hip.DeviceArray.from_host(np.random.default_rng(3).random(cells, dtype=np.float32))
    b = field_hip.DeviceArray(cells)
    cus = field_hip.require_device().compute_units
    ev = field_hip.Events()

    def timed(launch):
        for _ in range(3):
            launch()
        field_hip.sync()
        ev.begin()
        for _ in range(reps):
            launch()
        return ev.end() / reps

    runs = []
    for read_only in (False, True):
        for inflight in (1, 2, 4, 8):
            for nt in (False, True):
                for per_cu in (2, 4, 8):
                    ms = timed(lambda: field_hip.roof_f32(a, None if read_only else b, inflight, nt, read_only, per_cu * cus))
                    moved = (1 if read_only else 2) * 4.0 * cells
                    runs.append({"read_only": read_only, "inflight": inflight, "nontemporal": nt, "blocks_per_cu": per_cu,
                                 "ms": ms, "gbs": moved / (ms * 1e6)})
    copy_ms = timed(lambda: field_hip.copy_f32(a, b))
    best_copy = max((r for r in runs if not r["read_only"]), key=lambda r: r["gbs"])
    best_read = max((r for r in runs if r["read_only"]), key=lambda r: r["gbs"])
    return {"n": n, "bytes_each_way": 4 * cells, "bench_copy_gbs": 8.0 * cells / (copy_ms * 1e6), "best_copy": best_copy,
            "best_read": best_read, "runs": runs}


def valley_start(nx: int, ny: int, nz: int):
    """The valley scene (rock floor, sand bank, fuel with a fire, a pond, ice) laid out in proportion on nx x ny x nz:
    the golden case's start (tests/golden/cases.py valley_init, 48 x 40 x 24) with every coordinate scaled."""
    z, y, x = np.meshgrid(np.arange(nz) * (24.0 / nz), np.arange(ny) * (40.0 / ny), np.arange(nx) * (48.0 / nx), indexing="ij")
    rock = (z < 5).astype(np.float32)
    sand = ((z >= 5) & (z < 7) & (x < 20)).astype(np.float32)
    fuel = ((z >= 5) & (z < 8) & (x >= 12) & (x < 30) & (y >= 10) & (y < 30)).astype(np.float32)
    water = ((z >= 5) & (z < 10) & (x >= 34) & (x < 44) & (y >= 4) & (y < 14)).astype(np.float32)
    ice = ((z >= 5) & (z < 7) & (x >= 40) & (y >= 30)).astype(np.float32)
    t = np.full(x.shape, 290.0)
    t[(x - 20) ** 2 + (y - 20) ** 2 + (z - 7) ** 2 <= 9] = 1400.0
    t[ice > 0] = 250.0
    t = t + np.random.default_rng(71).random(x.shape) * 0.5
    flame = np.where((x - 20) ** 2 + (y - 20) ** 2 + (z - 7) ** 2 <= 4, 0.6, 0.0)
    out = {"rock": rock, "sand": sand, "water": water, "ice": ice, "fuel": fuel, "T": t, "flame": flame}
    return {k: np.ascontiguousarray(v, np.float32).ravel() for k, v in out.items()}


VALLEY_FIELDS = ["rock", "sand", "water", "ice", "steam", "fuel", "flame", "smoke", "ash", "T", "wet", "near"]


def valley_law() -> str:
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import cases  # the law text the golden case holds the card to the reference CPU engine with
    return cases.VALLEY_LAW


def world_valley(n=(144, 144, 48), steps: int = 50, check_steps: int = 3):
    """A many-term world law (the valley: 14 terms over 12 fields, the shipped world law's op mix) on 144 x 144 x 48
    (1 M cells, the size the reference times its world law at), h = 0.5, dt = 0.25: held to the reference CPU
    engine's bits after `check_steps` steps, then timed with contracts checked every step (lazy readback, one
    fe_solver_step a step, wall clock) and through Solver.run with contracts every 8 steps."""
    nx, ny, nz = n
    law = field_hip.Law(valley_law())
    start = valley_start(nx, ny, nz)

    def fresh():
        s = field_hip.Solver(law, (nx, ny, nz), 0.5)
        for f, v in start.items():
            s.set_field(f, v)
        s.refresh_derived()
        s.set_lazy_readback(True)
        return s

    s = fresh()
    s.check_contracts_every(1)
    for _ in range(check_steps):
        s.step(0.25)
    rec = reference_check(f"valley_{nx}x{ny}x{nz}", np.concatenate([s.get_field(f) for f in VALLEY_FIELDS]))
    walls = []
    for _ in range(steps):
        t = time.perf_counter()
        s.step(0.25)
        walls.append((time.perf_counter() - t) * 1e3)
    prof = s.profile()
    idle = sum(1 for t in prof["terms"] if t.get("idle"))
    r = fresh()
    r.check_contracts_every(8)
    r.run(0.25, steps)
    field_hip.sync()
    ev = field_hip.Events()
    ev.begin()
    r.run(0.25, steps)
    run_ms = ev.end() / steps
    out = {"cells": nx * ny * nz, "terms": len(prof["terms"]), "idle_terms_last_step": idle,
           "ms_per_step_contracts_every_step": float(np.median(walls)), "p99_ms": float(np.percentile(walls, 99)),
           "ms_per_step_run_contracts_every_8": run_ms, "violations_last_step": len(s.violations())}
    out.update(rec)
    return out


def reference_check(name: str, final: np.ndarray) -> dict:
    """The card's field against the reference CPU engine's on the same law, start and steps: tests/golden/
    bench_reference.json holds the SHA-256 of the reference's final field bits and its time a step, both taken in the
    build container by tests/golden/make_bench_reference.py (the reference's engine does not run on this host)."""
    path = os.path.join(ROOT, "tests", "golden", "bench_reference.json")
    if not os.path.exists(path):
        return {}
    rec = json.load(open(path)).get(name)
    if rec is None:
        return {}
    digest = hashlib.sha256(np.ascontiguousarray(final, np.float32).tobytes()).hexdigest()
    cells = int(np.asarray(final).size)
    return {"equals_reference_cpu_bits": digest == rec["sha256"],
            "cpu_baseline": {"value": cells / (rec["median_ms"] * 1e-3), "unit": "cell-updates/s", "cores": rec["cores"],
                             "kind": "reference", "sample": rec["sample"]}}


def block_mean(a, n: int, count: int, f: int) -> np.ndarray:
    """count n x n lattices averaged over f x f blocks (float64 sums), as float32."""
    m = n // f
    return np.asarray(a, np.float64).reshape(count, m, f, m, f).mean(axis=(2, 4)).reshape(-1).astype(np.float32)


def learned_coarse(cpu: bool, train_steps: int, test_seeds: int):
    """The learned coarse lattice (diffusion-reaction, Du 1e-3, Dv 5e-3, k 5e-3): 256 training seeds of the fine 512^2
    law run by the engine, 100 snapshots each coarsened to 128^2 on the card; the net trained on the card; then on
    unseen seeds the coarse law with and without the net, and the plain law at 256^2 and 512^2, all against the fine
    law in FP64 (the oracle), each with its card time for the same simulated time."""
    import learned_coarse as lc
    s = lc.Setup()
    stats = {}
    data = lc.generate(s, range(256 + test_seeds), base_seed=7000, stats=stats)
    scale = lc.target_scale(data)
    net = field_hip.LearnedNet(lc.initial_params(1, last_scale=0.0), scale)
    t = time.time()
    # Each snapshot pair's loss relative to the plain law's on it: the first pairs (the start's sharp blocks relaxing)
    # miss by 10^7 times more than the last and would otherwise be all the net learns.
    plain = lc.plain_losses(data, slice(0, 16), scale)
    weights = plain.mean() / plain
    losses = lc.train(net, data, train_steps, batch=16, lr=3e-3, train_seeds=256, log_every=max(1, train_steps // 5),
                      weights=weights)
    train_s = time.time() - t
    test = slice(256, 256 + test_seeds)
    snaps = s.snapshots
    t_end = snaps * s.dt_coarse
    times = {}
    tl, tp = [], []
    lu, lv = lc.rollout(net, data, test, snaps, learned=True, times=tl)
    pu, pv = lc.rollout(net, data, test, snaps, learned=False, times=tp)
    times["learned_128"], times["plain_128"] = tl[-1], tp[-1]
    u0, v0 = lc.initial(s, range(256, 256 + test_seeds), base_seed=7000)
    runs = {"learned_128": (lu, lv), "plain_128": (pu, pv)}
    for n, sub in ((256, 3 * snaps), (512, s.every * snaps)):
        f = n // s.coarse
        tt = []
        fu, fv = lc.plain_run(lc.upsample(u0, f).reshape(-1), lc.upsample(v0, f).reshape(-1), n, test_seeds, s, t_end, sub, tt)
        times[f"plain_{n}"] = tt[-1]
        runs[f"plain_{n}"] = (block_mean(fu, n, test_seeds, f), block_mean(fv, n, test_seeds, f))
    out = {"fine": s.fine, "coarse": s.coarse, "train_seeds": 256, "test_seeds": test_seeds, "snapshots": snaps,
           "data": stats, "train_steps": train_steps, "train_s": train_s, "loss_log_weighted": losses, "scale": scale,
           "loss_vs_plain_by_snapshot": {int(t): float(net.gradient(data.at("u", t, 0), data.at("v", t, 0), data.at("u", t + 1, 0),
                                                                   data.at("v", t + 1, 0), s.coarse, s.coarse, 16, s.h_coarse,
                                                                   s.du, s.dv, s.k, s.dt_coarse)[0] / plain[t])
                                         for t in (0, 1, 5, 20, 50, 99)},
           "card_ms_for_t_end": times, "fallbacks": net.fallbacks()}
    if cpu:
        truth = cpu_baseline_learned(s, u0, v0, test_seeds, t_end)
        errs = {}
        for name, (ru, rv) in runs.items():
            errs[name] = float(np.sqrt(np.mean((ru - truth["u"]) ** 2) + np.mean((rv - truth["v"]) ** 2)))
        out["rms_error_vs_fp64_fine"] = errs
        out["cpu_baseline"] = truth["baseline"]
        # equal error: the cheapest plain resolution whose error is no larger than the learned lattice's
        fit = [n for n in (128, 256, 512) if errs[f"plain_{n}"] <= errs["learned_128"]]
        if fit:
            n = fit[0]
            out["equal_error"] = {"plain": n, "speedup_learned_vs_plain": times[f"plain_{n}"] / times["learned_128"]}
        out["speedup_learned_128_vs_plain_512"] = times["plain_512"] / times["learned_128"]
    return out


def cpu_baseline_learned(s, u0, v0, count: int, t_end: float):
    """The fine law in FP64 on the host (the ORACLE's restatement, OpenMP): the truth the coarse lattices are judged
    against, and its time as the CPU baseline."""
    import learned_coarse as lc
    from oracle import diffreact, learned as L
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    steps = s.every * s.snapshots
    t = time.time()
    U, V = diffreact.run64(lc.upsample(u0, s.factor), lc.upsample(v0, s.factor), s.fine, s.fine, count, s.h_fine, s.du, s.dv,
                           s.k, t_end / steps, steps)
    wall = time.time() - t
    return {"u": L.coarsen(U, s.fine, s.fine, count, s.factor), "v": L.coarsen(V, s.fine, s.fine, count, s.factor),
            "baseline": {"value": 2 * s.fine ** 2 * count * steps / wall, "unit": "cell-updates/s", "cores": cores, "kind": "port",
                         "sample": f"oracle FP64 fine law, {count} seeds of {s.fine}^2, {steps} steps ({wall:.1f} s)"}}


TRANSPORT_LAW = """field a init 0
field b init 0
field c init 0
param Da = 0.05
param Db = 0.02
param Dc = 0.01
term transport
  d a = diffuse(Da, a) - 0.3 * (a - shift(a, -1, 0, 0)) / h + 0.2 * (shift(a, 0, 1, 0) - a) / h - 0.1 * (a - shift(a, 0, 0, -1)) / h - a * b + 0.1 * c
  d b = diffuse(Db, b) - 0.3 * (b - shift(b, -1, 0, 0)) / h + 0.2 * (shift(b, 0, 1, 0) - b) / h - 0.1 * (b - shift(b, 0, 0, -1)) / h - a * b + 0.1 * c
  d c = diffuse(Dc, c) - 0.3 * (c - shift(c, -1, 0, 0)) / h + 0.2 * (shift(c, 0, 1, 0) - c) / h - 0.1 * (c - shift(c, 0, 0, -1)) / h + a * b - 0.1 * c
"""


def adjoint_transport(n: int, cpu: bool):
    """The gradient of L = sum c^2 after 64 steps of the three-field transport law on n^3, with respect to the start and
    the three diffusivities: the generated adjoint law, checkpoints every 8 steps, recomputation, on the card."""
    import adjoint
    h = 1.0 / n
    dt = 0.8 * h * h / (6 * 0.05)
    rng = np.random.default_rng(11)
    init = {f: rng.uniform(0.2, 1.0, n ** 3).astype(np.float32) for f in ("a", "b", "c")}
    res = adjoint.vjp(TRANSPORT_LAW, (n, n, n), h, {"Da": 0.05, "Db": 0.02, "Dc": 0.01}, init, dt, 64, "c", every=8)
    cells = n ** 3
    # a backward step: the adjoint law reads 3 primal and 6 adjoint fields and writes 6 (60 B a cell); the recomputed
    # forward steps (56 of them) read and write 3 fields (24 B a cell)
    bytes_back = (64 * 60 + 56 * 24) * cells
    out = {"n": n, "steps": 64, "every": 8, "loss": res.loss, "param_grads": res.param_grads, "forward_ms": res.forward_ms,
           "backward_ms": res.backward_ms, "backward_cell_updates_per_s": 64 * cells / (res.backward_ms * 1e-3),
           "backward_gbs": bytes_back / (res.backward_ms * 1e6), **res.stats}
    out["backward_frac_hbm"] = out["backward_gbs"] / HBM_PEAK_GBS
    if cpu:
        out["fd_check"] = cpu_baseline_adjoint()
    return out


def cpu_baseline_adjoint():
    """Central differences of the ORACLE's FP64 run on a small lattice against the card's checkpointed adjoint."""
    import adjoint
    from oracle import transport
    shape, h, dt, steps = (10, 9, 8), 0.125, 0.01, 12
    D = [0.05, 0.02, 0.01]
    rng = np.random.default_rng(3)
    init = {f: rng.uniform(0.2, 1.0, int(np.prod(shape))) for f in ("a", "b", "c")}
    res = adjoint.vjp(transport.LAW, shape, h, {"Da": D[0], "Db": D[1], "Dc": D[2]}, init, dt, steps, "c", every=4)
    worst = 0.0
    for i, name in enumerate(("Da", "Db", "Dc")):
        e = 1e-4 * D[i]  # central: truncation ~(e/D)^2, rounding ~ulp(L)/e
        up, dn = list(D), list(D)
        up[i] += e
        dn[i] -= e
        fd = (transport.loss(shape, h, dt, steps, up, init["a"], init["b"], init["c"])
              - transport.loss(shape, h, dt, steps, dn, init["a"], init["b"], init["c"])) / (2 * e)
        worst = max(worst, abs(res.param_grads[name] - fd) / abs(fd))
    return {"lattice": list(shape), "steps": steps, "every": 4, "max_rel_error_param_grads": worst}


def measured_traffic(kernel: str, n: int):
    """HBM bytes a launch of the engine's gather moves, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (scaled by a copy kernel's; profiles/conduction_256_512_rocprof.md), when they were taken on this kernel and size."""
    path = os.path.join(ROOT, "profiles", f"conduction_{n}_traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    return rec["bytes_per_launch"] if rec.get("kernel") == kernel else None


def cpu_baseline(n: int, budget_s: float):
    """The oracle's C restatement of the reference CPU path (OpenMP over this host's cores) on the same step, a bounded
    sample: substeps of the N^3 lattice until about budget_s seconds have passed."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    h = 1.0 / n
    dt = 0.4 * h * h / 0.1 / 3
    from oracle import heat
    t0 = np.random.default_rng(7).random(n ** 3, dtype=np.float32)
    times, cur, t_start = [], t0, time.time()
    while time.time() - t_start < budget_s or len(times) < 2:
        t = time.time()
        cur = heat.substep(cur, (n, n, n), 0.1, h, dt)
        times.append(time.time() - t)
    ms = 1e3 * float(np.median(times))
    return {"value": n ** 3 / (ms * 1e-3), "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle C restatement (OpenMP), {n}^3, {len(times)} substeps, median {ms:.1f} ms"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-large", action="store_true", help="skip the 512^3 case")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adi", action="store_true", help="skip the batched Heston ADI case")
    ap.add_argument("--adi-options", type=int, default=2048)
    ap.add_argument("--adi-steps", type=int, default=512)
    ap.add_argument("--adi-sample", type=int, default=16, help="options of each shape checked against FP64")
    ap.add_argument("--no-extra", action="store_true", help="skip the Dirichlet, Black-Scholes, learned and adjoint cases")
    ap.add_argument("--learned-steps", type=int, default=3000)
    ap.add_argument("--adjoint-size", type=int, default=256)
    ap.add_argument("--section", help="run one of the other cases alone and print its record (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        return 2
    info = field_hip.require_device()
    log(f"device {info.name.decode()} ({info.arch.decode()}, {info.compute_units} CUs)")

    sections = {"dirichlet_heat_64": dirichlet_heat,
                "black_scholes_4096x512": lambda: black_scholes(200),
                "implicit_lines_256": lambda: implicit_lines(256, 50),
                "mppi_plan": lambda: mppi_plan(not args.no_cpu),
                "net_train_world_model": lambda: net_train(not args.no_cpu),
                "hbm_roof_512": lambda: hbm_roof(512),
                "world_valley_144x144x48": lambda: world_valley(),
                "learned_coarse": lambda: learned_coarse(not args.no_cpu, args.learned_steps, 4),
                "adjoint_transport": lambda: adjoint_transport(args.adjoint_size, not args.no_cpu)}
    if args.section:
        if args.section not in sections:
            print(f"no case {args.section!r} (cases: {', '.join(sections)})", file=sys.stderr)
            return 2
        print(json.dumps({args.section: sections[args.section]()}, default=float), flush=True)
        return 0

    main_case = conduction(args.size, args.steps, args.warmup)
    log(f"{args.size}^3: {json.dumps(main_case)}")
    large = None if args.no_large else conduction(512, 64, max(3, args.warmup // 4), verify_steps=1)
    if large:
        log(f"512^3: {json.dumps(large)}")
    cpu = None if args.no_cpu else cpu_baseline(args.size, args.cpu_seconds)
    if cpu:
        log(f"cpu baseline: {json.dumps(cpu)}")
    adi = None
    if not args.no_adi:
        adi = heston_adi(args.adi_options, args.adi_steps, args.adi_sample, not args.no_cpu)
        log(f"heston adi: {json.dumps(adi)}")
    extra = {}
    if not args.no_extra:
        for name, fn in sections.items():
            t = time.time()
            extra[name] = fn()
            log(f"{name} ({time.time() - t:.1f} s): {json.dumps(extra[name], default=float)}")

    achieved = BYTES_PER_CELL * main_case["cells"] / (main_case["engine_kernel_ms"] * 1e6)
    line = {
        "metric": "cell-updates/s (fused 7-point conduction step through the engine)",
        "value": main_case["engine_cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_case["engine_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage, f64 per-cell arithmetic (the reference CPU's)",
        "data": "synthetic (seeded uniform [0,1))",
        "config": {"workload": f"law d T = diffuse(k, T) on {args.size}^3 fp32, h=1/{args.size}, k=0.1, one substep of "
                               f"dt=0.4h^2/k/3 a step, ping-pong, lazy readback, contracts off",
                   "engine_equals_handwritten_bits": main_case["engine_equals_handwritten_bits"],
                   "handwritten": {k: main_case[k] for k in ("hand_kernel_ms", "hand_gbs", "hand_ms_per_step")},
                   "engine_kernel": main_case["engine_kernel"], "engine_kernel_ms": main_case["engine_kernel_ms"],
                   "engine_ms_per_step_contracts_on": main_case["engine_ms_per_step_contracts_on"],
                   "copy_roof_gbs": main_case["copy_gbs"], "large_512": large, "heston_adi": adi, **extra},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(main_case["engine_kernel"], args.size)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line, default=float), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
