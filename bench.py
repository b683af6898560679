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
import json
import os
import shutil
import subprocess
import sys
import tempfile
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
    eng_whole, _ = time_launches(lambda: solver.step(sub_dt), steps, warmup)
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
    """The batch on the card (Craig-Sneyd, theta 1/2, fp32): tuned, every step captured in one graph, one launch timed;
    prices of a sample (the same mix) against the FP64 oracle; the FP32 oracle on the host's cores on that sample."""
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
    sample = np.concatenate([np.flatnonzero(nx == s[0])[:sample_per_class] for s in ADI_SHAPES])
    if sample.size:
        from oracle import heston
        t = time.time()
        ref, _, _ = heston.batch64(nx[sample], ny[sample], opts[sample], steps)
        fp64_s = time.time() - t
        err = np.abs(prices[sample] - ref)
        out["error_vs_fp64"] = {"options": int(sample.size), "max_abs": float(err.max()),
                                "max_rel": float(np.max(err / np.maximum(np.abs(ref), 1e-3))), "rms_abs": float(np.sqrt(np.mean(err ** 2))),
                                "oracle_s": fp64_s}
        if cpu:
            cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
            t = time.time()
            heston.batch32(nx[sample], ny[sample], opts[sample], steps)
            wall = time.time() - t
            cpu_rate = sample.size / wall
            out["cpu_baseline"] = {"value": cpu_rate, "unit": "solves/s", "cores": cores, "kind": "port",
                                   "sample": f"oracle FP32 restatement of the same scheme (OpenMP over options), {sample.size} "
                                             f"options ({sample_per_class} of each lattice shape), {steps} steps, {wall:.1f} s"}
            out["speedup_vs_cpu"] = out["solves_per_s"] / cpu_rate
    return out


def measured_traffic(kernel: str, n: int):
    """HBM bytes a launch of the engine's gather moves, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (scaled by a copy kernel's; profiles/conduction_256_512_rocprof.md), when they were taken on this kernel and size."""
    path = os.path.join(ROOT, "profiles", f"conduction_{n}_traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    return rec["bytes_per_launch"] if rec.get("kernel") == kernel else None


def cpu_baseline(n: int, budget_s: float):
    """The reference's own CPU engine (oracle/_ref/lawc, built from its sources) on this host, a bounded sample of the
    same step; the C restatement (oracle/) when the reference binary is absent."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    lawc = os.path.join(ROOT, "oracle", "_ref", "lawc")
    h = 1.0 / n
    dt = 0.4 * h * h / 0.1 / 3
    if os.path.exists(lawc):
        tmp = tempfile.mkdtemp()
        try:
            law = os.path.join(tmp, "heat.law")
            open(law, "w").write(HEAT_LAW)
            scene = os.path.join(tmp, "s.scene")
            open(scene, "w").write(f"grid {n} {n} {n} {h!r}\nnoise T 1.0 7\n")
            out = os.path.join(tmp, "t.json")
            steps = 5
            env = dict(os.environ, OMP_NUM_THREADS=str(cores))
            t = time.time()
            subprocess.run([lawc, "--scene", scene, "--steps", str(steps + 1), "--warmup", "1", "--dt", repr(dt), "--json",
                            out, law], check=True, capture_output=True, env=env, timeout=max(60, 10 * budget_s))
            wall = time.time() - t
            res = json.load(open(out))
            ms = res["median_ms"]
            return {"value": n ** 3 / (ms * 1e-3), "unit": "cell-updates/s", "cores": cores, "kind": "reference",
                    "sample": f"reference lawc (its CPU engine, OpenMP {cores} threads), d T = diffuse(k,T) with its "
                              f"contracts, {n}^3, {steps} timed steps of one substep after 1 warm-up, median {ms:.1f} "
                              f"ms/step ({wall:.1f} s wall)"}
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
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
    ap.add_argument("--adi-sample", type=int, default=4, help="options of each shape checked against FP64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        return 2
    info = field_hip.require_device()
    log(f"device {info.name.decode()} ({info.arch.decode()}, {info.compute_units} CUs)")

    main_case = conduction(args.size, args.steps, args.warmup)
    log(f"{args.size}^3: {json.dumps(main_case)}")
    large = None if args.no_large else conduction(512, max(20, args.steps // 8), max(3, args.warmup // 4), verify_steps=1)
    if large:
        log(f"512^3: {json.dumps(large)}")
    cpu = None if args.no_cpu else cpu_baseline(args.size, args.cpu_seconds)
    if cpu:
        log(f"cpu baseline: {json.dumps(cpu)}")
    adi = None
    if not args.no_adi:
        adi = heston_adi(args.adi_options, args.adi_steps, args.adi_sample, not args.no_cpu)
        log(f"heston adi: {json.dumps(adi)}")

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
                   "copy_roof_gbs": main_case["copy_gbs"], "large_512": large, "heston_adi": adi},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(main_case["engine_kernel"], args.size)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
