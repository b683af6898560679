"""State tooling: scenes, dumped states and --init through this build's lawc, held to the reference's own lawc.

tests/golden/scene/pond.json holds SHA-256 digests of what the reference's lawc made of pond.scene (its start, and
its state after 3 and 6 steps on its CPU engine; tests/golden/make_scene_golden.py). The scene's start needs no card;
stepping does (-m gpu).
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LAWC = os.path.join(ROOT, "online-physics-cuda_amd", "lawc")
SCENE_DIR = os.path.join(ROOT, "tests", "golden", "scene")
REF = json.load(open(os.path.join(SCENE_DIR, "pond.json")))


def lawc(*args, cwd=SCENE_DIR):
    return subprocess.run([LAWC, *args], cwd=cwd, capture_output=True, text=True, timeout=600)


def digests(state_dir):
    meta = json.load(open(os.path.join(state_dir, "meta.json")))
    return {f: hashlib.sha256(np.fromfile(os.path.join(state_dir, f + ".bin"), np.float32).tobytes()).hexdigest()
            for f in meta["fields"]}


def test_scene_start_is_the_references(tmp_path):
    """fill, box, ball and hashed noise give the reference's start bit for bit (the derived field is taken on the card,
    so only the integrated fields are compared here)."""
    r = lawc("--scene", "pond.scene", "--steps", "0", "--dump", str(tmp_path))
    assert r.returncode == 0, r.stderr
    got = digests(os.path.join(tmp_path, "step_0"))
    for field in ("T", "water", "vapour"):
        assert got[field] == REF["step_0"][field], field
    meta = json.load(open(os.path.join(tmp_path, "step_0", "meta.json")))
    assert (meta["nx"], meta["ny"], meta["nz"], meta["h"]) == (20, 16, 10, 0.5)
    assert meta["fields"] == ["T", "water", "vapour", "wet"]


def test_dump_only_and_init_round_trip_on_the_host(tmp_path):
    a, b = os.path.join(tmp_path, "a"), os.path.join(tmp_path, "b")
    assert lawc("--scene", "pond.scene", "--steps", "0", "--dump", a).returncode == 0
    # A state read back with --init and dumped again is the same state; --dump-only keeps the fields named.
    r = lawc("--scene", "pond.scene", "--steps", "0", "--init", os.path.join(a, "step_0"), "--dump", b, "--dump-only", "water,T")
    assert r.returncode == 0, r.stderr
    assert json.load(open(os.path.join(b, "step_0", "meta.json")))["fields"] == ["water", "T"]
    da, db = digests(os.path.join(a, "step_0")), digests(os.path.join(b, "step_0"))
    assert da["water"] == db["water"] and da["T"] == db["T"]


@pytest.mark.parametrize("line,message", [
    ("grid 4 4\n", "expected 'grid NX NY NZ METRES'"),
    ("box T 1 0 0 0 2\n", "box needs 6 numbers after the value"),
    ("ball T\n", "expected 'ball FIELD VALUE ...'"),
    ("sphere T 1 2 3 4 5\n", "unknown line 'sphere'"),
])
def test_scene_errors_name_the_line(tmp_path, line, message):
    path = os.path.join(tmp_path, "bad.scene")
    open(path, "w").write("# a bad scene\n" + line)
    r = lawc("--scene", path, "--steps", "0", "--dump", str(tmp_path), "pond.law")
    assert r.returncode == 1
    assert f"bad.scene:2: {message}" in r.stderr


def test_unknown_field_and_wrong_state_size_are_refused(tmp_path):
    path = os.path.join(tmp_path, "x.scene")
    open(path, "w").write("grid 4 4 2 1\nfill nothing 1\n")
    r = lawc("--scene", path, "--steps", "0", "--dump", str(tmp_path), "pond.law")
    assert r.returncode == 1 and "the law has no field 'nothing'" in r.stderr
    state = os.path.join(tmp_path, "state")
    os.makedirs(state)
    np.zeros(7, np.float32).tofile(os.path.join(state, "T.bin"))
    r = lawc("--grid", "4", "4", "2", "1", "--init", state, "--steps", "0", "--dump", str(tmp_path), "pond.law")
    assert r.returncode == 1 and "holds 7 floats, the lattice 32" in r.stderr


@pytest.mark.gpu
def test_stepped_states_and_init_are_the_references(device, tmp_path):
    """6 steps on the card dumped every 3 equal the reference CPU engine's states; a run restarted from the dumped
    step 3 with --init reaches the same step 6."""
    out, again = os.path.join(tmp_path, "out"), os.path.join(tmp_path, "again")
    r = lawc("--scene", "pond.scene", "--steps", "6", "--dt", "0.05", "--dump", out, "--every", "3", "--json",
             os.path.join(tmp_path, "t.json"))
    assert r.returncode == 0, r.stdout + r.stderr
    for step in ("step_0", "step_3", "step_6"):
        assert digests(os.path.join(out, step)) == REF[step], step
    r = lawc("--scene", "pond.scene", "--init", os.path.join(out, "step_3"), "--steps", "3", "--dt", "0.05", "--dump", again)
    assert r.returncode == 0, r.stdout + r.stderr
    assert digests(os.path.join(again, "step_3")) == REF["step_6"]
    rec = json.load(open(os.path.join(tmp_path, "t.json")))
    assert rec["backend"] == "HIP" and rec["cells"] == 20 * 16 * 10 and rec["violations"] == 0
