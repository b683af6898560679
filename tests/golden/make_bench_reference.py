"""Writes tests/golden/bench_reference.json: the reference CPU engine's results for the bench's Dirichlet,
Black-Scholes and implicit-lines cases, as SHA-256 digests of its final field bits plus its time a step.

The bench holds the card's fields to these digests (bit identity) and reports the time as the CPU baseline of those
cases. The reference's `lawc` is built from its sources (make -f oracle/Makefile.ref) and runs only here, in the build
container; only this JSON travels. Laws and starts are the bench's own (bench.py).

    python tests/golden/make_bench_reference.py     (from the repository root)
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LAWC = os.path.join(ROOT, "oracle", "_ref", "lawc")
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def run(law: str, grid, init: dict, steps: int, dt: float, field, cores: int):
    """The reference's final `field` (or, for a list of names, their concatenation) and its median ms a step."""
    nx, ny, nz, h = grid
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "law.law")
        open(path, "w").write(law)
        init_dir = os.path.join(tmp, "init")
        os.makedirs(init_dir)
        for name, values in init.items():
            np.asarray(values, np.float32).tofile(os.path.join(init_dir, name + ".bin"))
        out_dir, js = os.path.join(tmp, "out"), os.path.join(tmp, "t.json")
        subprocess.run([LAWC, "--grid", str(nx), str(ny), str(nz), repr(float(h)), "--init", init_dir, "--steps", str(steps),
                        "--dt", repr(float(dt)), "--dump", out_dir, "--json", js, "--warmup", "1", path],
                       check=True, capture_output=True, env=dict(os.environ, OMP_NUM_THREADS=str(cores)), timeout=1200)
        names = [field] if isinstance(field, str) else list(field)
        final = np.concatenate([np.fromfile(os.path.join(out_dir, f"step_{steps}", f + ".bin"), dtype=np.float32) for f in names])
        return final, json.load(open(js))["median_ms"]


def main() -> None:
    if not os.path.exists(LAWC):
        sys.exit("build the reference first: make -f oracle/Makefile.ref REF=/root/reference")
    cores = len(os.sched_getaffinity(0))
    out = {}
    n = 64
    h = 1.0 / n
    t0, nb = bench.sine_start(n, 1e-3)
    final, ms = run(bench.DIRICHLET_LAW, (n, n, n, h), {"T": t0, "nb": nb}, 200, 0.4 * h * h / 0.1, "T", cores)
    out["dirichlet_heat_64"] = {"sha256": hashlib.sha256(final.tobytes()).hexdigest(), "steps": 200, "median_ms": ms,
                                "cores": cores, "sample": f"reference lawc (its CPU engine), OpenMP {cores} threads in the "
                                f"build container, 64^3 with nb, 200 steps of 3 substeps, median {ms:.3f} ms/step"}
    v0, h, dt = bench.black_scholes_start(4096, 512)
    final, ms = run(bench.BLACK_SCHOLES_LAW, (4096, 1, 512, h), {"V": v0}, 10, dt, "V", cores)
    out["black_scholes_4096x512"] = {"sha256": hashlib.sha256(final.tobytes()).hexdigest(), "steps": 10, "median_ms": ms,
                                     "cores": cores, "sample": f"reference lawc (its CPU engine), OpenMP {cores} threads in "
                                     f"the build container, 4096x1x512, 10 steps, median {ms:.3f} ms/step"}
    n = 256
    h = 1.0 / n
    t0, w0 = bench.implicit_start(n)
    for name, law, dt, init in (("crank_nicolson_x", bench.IMPLICIT_X_LAW, 0.5 * h * h / 0.01, {"T": t0}),
                                ("implicit_yz", bench.IMPLICIT_YZ_LAW, 0.5 * h * h / 0.05, {"T": t0, "w": w0})):
        final, ms = run(law, (n, n, n, h), init, 3, dt, "T", cores)
        out[f"{name}_{n}"] = {"sha256": hashlib.sha256(final.tobytes()).hexdigest(), "steps": 3, "median_ms": ms,
                              "cores": cores, "sample": f"reference lawc (its CPU engine), OpenMP {cores} threads in the "
                              f"build container, {n}^3, 3 steps, median {ms:.1f} ms/step"}
    nx, ny, nz = 144, 144, 48
    final, ms = run(bench.valley_law(), (nx, ny, nz, 0.5), bench.valley_start(nx, ny, nz), 3, 0.25, bench.VALLEY_FIELDS, cores)
    out[f"valley_{nx}x{ny}x{nz}"] = {"sha256": hashlib.sha256(final.tobytes()).hexdigest(), "steps": 3, "median_ms": ms, "cores": cores,
                                     "sample": f"reference lawc (its CPU engine), OpenMP {cores} threads in the build "
                                     f"container, the valley law on {nx}x{ny}x{nz}, 3 steps, contracts every step, median "
                                     f"{ms:.1f} ms/step"}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bench_reference.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
