"""Writes the golden vectors under tests/golden/ by running the REFERENCE's own CPU engine.

The reference's `lawc` is built from its sources where they lie (make -f oracle/Makefile.ref, outputs in oracle/_ref/,
git-ignored) and run here, in the build container, once per case: a law (text written for this project), a seeded
initial state, N steps of dt. Each case's inputs and the reference's outputs are committed as one small .npz; the
tests hold the CPU restatement (oracle/) and the HIP backend to them. Nothing here runs on the GPU box.

    python tests/golden/make_golden.py [NAME...]    (from the repository root; no names: every case)
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HERE = os.path.dirname(os.path.abspath(__file__))
LAWC = os.path.join(ROOT, "oracle", "_ref", "lawc")
VIOLATIONS = os.path.join(ROOT, "oracle", "_ref", "ref_violations")
sys.path.insert(0, HERE)

from cases import CASES  # noqa: E402


def run_reference(case: dict) -> dict:
    nx, ny, nz, h = case["grid"]
    with tempfile.TemporaryDirectory() as tmp:
        law_path = os.path.join(tmp, "case.law")
        with open(law_path, "w") as f:
            f.write(case["law"])
        init_dir = os.path.join(tmp, "init")
        os.makedirs(init_dir)
        for name, values in case["init"].items():
            np.asarray(values, dtype=np.float32).tofile(os.path.join(init_dir, name + ".bin"))
        out_dir = os.path.join(tmp, "out")
        cmd = [LAWC, "--grid", str(nx), str(ny), str(nz), repr(float(h)), "--init", init_dir, "--steps",
               str(case["steps"]), "--dt", repr(float(case["dt"])), "--dump", out_dir]
        for name, value in case.get("params", {}).items():
            cmd += ["--param", name, repr(float(value))]
        cmd.append(law_path)
        # A law's nets load from assets/nets/NAME.net under the working directory (src/field/Law.cpp:1706).
        for name, text in case.get("nets", {}).items():
            os.makedirs(os.path.join(tmp, "assets", "nets"), exist_ok=True)
            with open(os.path.join(tmp, "assets", "nets", name + ".net"), "w") as f:
                f.write(text)
        res = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp, env=dict(os.environ, OMP_NUM_THREADS="4"))
        if res.returncode != 0:
            raise RuntimeError(f"lawc failed for {case['name']}: {res.stdout}\n{res.stderr}")
        final_dir = os.path.join(out_dir, f"step_{case['steps']}")
        meta = json.load(open(os.path.join(final_dir, "meta.json")))
        final = {name: np.fromfile(os.path.join(final_dir, name + ".bin"), dtype=np.float32) for name in meta["fields"]}
        # The last step's violations, from the same engine run the same way (oracle/ref_violations.cpp).
        viol = subprocess.run([VIOLATIONS, law_path, str(nx), str(ny), str(nz), repr(float(h)), init_dir, str(case["steps"]),
                               repr(float(case["dt"]))], capture_output=True, text=True, cwd=tmp,
                              env=dict(os.environ, OMP_NUM_THREADS="4"))
        if viol.returncode != 0:
            raise RuntimeError(f"ref_violations failed for {case['name']}: {viol.stderr}")
        last = [json.loads(line) for line in viol.stdout.splitlines() if line.strip()]
        return {"final": final, "stdout": res.stdout, "stderr": res.stderr, "final_violations": last}


def main() -> None:
    if not os.path.exists(LAWC):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    only = set(sys.argv[1:])
    for case in CASES:
        if only and case["name"] not in only:
            continue
        out = run_reference(case)
        arrays = {}
        for name, values in case["init"].items():
            arrays["init_" + name] = np.asarray(values, dtype=np.float32)
        for name, values in out["final"].items():
            arrays["final_" + name] = values
        meta = {k: case[k] for k in ("name", "law", "grid", "dt", "steps")}
        meta["params"] = case.get("params", {})
        for key in ("nets", "atol"):
            if key in case:
                meta[key] = case[key]
        if "atol_moved" in case:
            moved = max(float(np.max(np.abs(out["final"][f].astype(np.float64) - case["init"][f]))) for f in case["init"])
            meta["atol"] = case["atol_moved"] * moved + 1e-3
        meta["final_violations"] = out["final_violations"]
        meta["reference_stdout"] = out["stdout"]
        meta["reference_stderr"] = out["stderr"]
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        path = os.path.join(HERE, case["name"] + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"{case['name']}: {os.path.getsize(path)} bytes; {out['stdout'].strip().splitlines()[-1]}")


if __name__ == "__main__":
    main()
