"""The law front end and the --hip generator, on the CPU (no device needed)."""
import glob
import json
import os
import subprocess

import pytest

import golden_io

REFERENCE = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LAWC = os.path.join(ROOT, "oracle", "_ref", "lawc")


def reference_laws():
    laws = [[p] for p in sorted(glob.glob(os.path.join(REFERENCE, "bench", "laws", "*.law")))]
    bench = os.path.join(REFERENCE, "assets", "law", "bench")
    for name in ("heat.law", "front.law"):
        laws.append([os.path.join(bench, name)])
    laws.append([os.path.join(bench, "front.law"), os.path.join(bench, "front_fix.law")])  # an add-on to front.law
    assets = os.path.join(REFERENCE, "assets", "law")
    # people.law's creature blocks read the world's fields: it compiles only beside world.law (creatures are the demo's
    # agents, outside the lattice; this front end skips their blocks).
    for group in (["world.law", "magic.law"], ["world.law", "magic.law", "people.law"], ["life.law"], ["map.law"],
                  ["pours.law"]):
        laws.append([os.path.join(assets, g) for g in group])
    return [p for p in laws if all(os.path.exists(q) for q in p)]


@pytest.mark.skipif(not (os.path.isdir(REFERENCE) and os.path.exists(REF_LAWC)),
                    reason="the reference and its built lawc are only in the build container")
@pytest.mark.parametrize("paths", reference_laws(), ids=lambda p: "+".join(os.path.basename(q) for q in p))
def test_ir_json_is_the_references(tmp_path, paths):
    import field_hip

    ref = tmp_path / "ref.json"
    r = subprocess.run([REF_LAWC, "--ir-json", str(ref)] + paths, capture_output=True, cwd=tmp_path)
    if r.returncode != 0:
        with pytest.raises(field_hip.FieldHipError):  # refused there, refused here
            field_hip.Law.from_files(*paths)
        return
    mine = field_hip.Law.from_files(*paths).ir_json()
    assert mine == ref.read_text()


@pytest.mark.parametrize("name", golden_io.names())
def test_generated_source_compiles_for_gfx950(name):
    import field_hip

    meta, _, _ = golden_io.load(name)
    law = golden_io.make_law(field_hip, meta)
    src = law.hip_source()
    assert "void g_t0(K k" in src
    opts = ["-O3", "-std=c++17", "-ffp-contract=off", "-fgpu-flush-denormals-to-zero", "-Xclang", "-target-feature",
            "-Xclang", "-packed-fp32-ops"]
    blob = field_hip.compile_hip(src, "law.hip", opts)
    assert len(blob) > 1000
    # every window's LDS ring within 64 KB (a spilling window may reach it through flat addresses, which fault past
    # the first 64 KB)
    import re
    rings = [int(n) * 4 for n in re.findall(r"__shared__ __attribute__\(\(aligned\(16\)\)\) float L_\w+\[(\d+)\]", src)]
    assert all(b <= 64 * 1024 for b in rings), rings


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="the reference's laws are only in the build container")
@pytest.mark.parametrize("paths", reference_laws(), ids=lambda p: "+".join(os.path.basename(q) for q in p))
def test_shipped_laws_generate_and_compile(tmp_path, paths):
    # Every op of the shipped laws has a form on the card; the world law (world + magic, 43 terms) is also compiled
    # (the others' sources are ordinary ops the golden laws compile; compiling each would take minutes here).
    import field_hip

    try:
        law = field_hip.Law.from_files(*paths)
    except field_hip.FieldHipError:
        pytest.skip("the reference's front end refuses it too (checked by test_ir_json_is_the_references)")
    src = law.hip_source()
    assert "extern \"C\" __global__ void g_t0(K k" in src
    if [os.path.basename(p) for p in paths] == ["world.law", "magic.law"]:
        opts = ["-O3", "-std=c++17", "-ffp-contract=off", "-fgpu-flush-denormals-to-zero", "-Xclang", "-target-feature",
                "-Xclang", "-packed-fp32-ops"]
        assert len(field_hip.compile_hip(src, "law.hip", opts)) > 1000


ERRORS = [
    ("field T init 0\nterm t\n  d T = diffuse(k, T)\n", "unknown name 'k'"),
    ("field T init 0\nterm t\n  d T = lap(T) +\n", "expression ends too early"),
    ("field T init 0 static\nterm t\n  d T = lap(T)\n", "writes static field 'T'"),
    ("field T init 0\nterm t\n  d T = lap(T)\n  conserves U\n", "'U' is not a field"),
    ("field T init 0\nfield U init 0\nterm t\n  d T = lap(T)\n  conserves U\n", "conserves 'U' but never writes it"),
    ("field T init 0\nterm t\n  d T = shift(T, 0.5, 0, 0)\n", "whole numbers of cells"),
    ("field T init 0\nterm t\n  max_dt = T\n  d T = lap(T)\n", "max_dt must be a number"),
    ("field T init 0\nterm t\n  d T = lap(T, T)\n", "takes 1 argument(s), got 2"),
    ("field T init 0\nterm t\n", "has no equations"),
    ("field T init 0\nterm t\n  d T = foo(T)\n", "unknown function 'foo'"),
]


@pytest.mark.parametrize("text,message", ERRORS)
def test_broken_laws_are_refused_by_line(text, message):
    import field_hip

    with pytest.raises(field_hip.FieldHipError) as e:
        field_hip.Law(text)
    assert message in str(e.value)


def test_precedence_and_folding():
    import field_hip

    # -2^2 is -(2^2); 2^3^2 is 2^(3^2); constants fold as the CPU computes them
    law = field_hip.Law("field T init 0\nparam p = -2 ^ 2\nparam q = 2 ^ 3 ^ 2\nterm t\n  d T = 0 * lap(T) + (1 + 1) * T\n")
    ir = json.loads(law.ir_json())
    params = {p["name"]: p["value"] for p in ir["params"]}
    assert params == {"p": -4.0, "q": 512.0}
    ops = [i["op"] for i in ir["insts"]]
    assert "lap" not in ops or ir["terms"][0]["equations"][0][1] != ops.index("lap")
    assert any(i["op"] == "const" and i["number"] == 2.0 for i in ir["insts"])


def test_in_place_and_ping_pong_decision():
    import field_hip

    ir = json.loads(field_hip.Law("field A init 0\nfield B init 0\nterm s\n  d A = -A * B\nterm p\n  d A = lap(A)\n").ir_json())
    assert [t["in_place"] for t in ir["terms"]] == [True, False]


def test_every_op_has_a_form_on_the_card():
    import field_hip

    law = field_hip.Law("field A init 0\nfield B init 0\nterm t\n  d A = 0.1 * distance(B) + poisson(B, 3)\n"
                        "  d B = fall(B, A, 1) + spread(B, A, 0.2) + topple(B, A, 1)\n")
    src = law.hip_source()
    for kernel in ("_mark", "_scan", "_signed", "_sweep", "_copy"):
        assert kernel + "(K k" in src


def _round_trip(field_hip, law):
    text = law.text()
    again = field_hip.Law(text)
    assert again.ir_json() == law.ir_json(), text
    assert again.text() == text  # printing is a fixed point


@pytest.mark.parametrize("name", golden_io.names())
def test_law_text_compiles_back_to_the_same_ir(name, tmp_path):
    # Law -> text -> Law gives the same IR on every golden law (nets read from the same weight files).
    import field_hip

    meta, _, _ = golden_io.load(name)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        os.makedirs(os.path.join("assets", "nets"))
        for net, text in meta.get("nets", {}).items():
            with open(os.path.join("assets", "nets", net + ".net"), "w") as f:
                f.write(text)
        _round_trip(field_hip, field_hip.Law(meta["law"]))
    finally:
        os.chdir(cwd)


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="the reference's laws are only in the build container")
@pytest.mark.parametrize("paths", reference_laws(), ids=lambda p: "+".join(os.path.basename(q) for q in p))
def test_shipped_law_text_compiles_back_to_the_same_ir(tmp_path, paths):
    import field_hip

    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        try:
            law = field_hip.Law.from_files(*paths)
        except field_hip.FieldHipError:
            pytest.skip("refused (checked by test_ir_json_is_the_references)")
        _round_trip(field_hip, law)
    finally:
        os.chdir(cwd)
