"""Writes tests/golden/scene/pond.json: SHA-256 digests of every field the REFERENCE's own lawc makes of
tests/golden/scene/pond.scene, at its start (step_0) and after 6 steps of 0.05 s, on its CPU engine (built from its
sources: make -f oracle/Makefile.ref). The state-tooling tests hold this build's lawc (--scene, --dump, --init) to them.

    python tests/golden/make_scene_golden.py     (from the repository root, in the build container)
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
HERE = os.path.join(ROOT, "tests", "golden", "scene")
LAWC = os.path.join(ROOT, "oracle", "_ref", "lawc")


def digests(state_dir):
    meta = json.load(open(os.path.join(state_dir, "meta.json")))
    return {f: hashlib.sha256(np.fromfile(os.path.join(state_dir, f + ".bin"), np.float32).tobytes()).hexdigest()
            for f in meta["fields"]}


def main() -> None:
    if not os.path.exists(LAWC):
        sys.exit("build the reference first: make -f oracle/Makefile.ref REF=/root/reference")
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "out")
        subprocess.run([LAWC, "--scene", "pond.scene", "--steps", "6", "--dt", "0.05", "--dump", out, "--every", "3"],
                       cwd=HERE, check=True, capture_output=True, env=dict(os.environ, OMP_NUM_THREADS="4"))
        rec = {"steps": 6, "dt": 0.05, "every": 3,
               "step_0": digests(os.path.join(out, "step_0")), "step_3": digests(os.path.join(out, "step_3")),
               "step_6": digests(os.path.join(out, "step_6"))}
    with open(os.path.join(HERE, "pond.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
