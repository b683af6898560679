"""The front end against the reference's own compiler, law by law (CPU; the reference is only in the build container).

Every law of the corpus below goes through the reference's lawc (built from its sources by oracle/Makefile.ref) and
through this front end: where the reference accepts a law, the IR JSON must be byte-identical; where it refuses one,
the list of errors (source:line: message, in order) must be the same list. The corpus covers the grammar's corners
(precedence, signs, lets, learned schemes, derives with '+=', conservation groups, shifts, poisson sweeps, nets,
declarations with ranges and units, show lines, creature blocks) and one law for every refusal the parser has.
"""
import os
import subprocess

import pytest

REFERENCE = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LAWC = os.path.join(ROOT, "oracle", "_ref", "lawc")

pytestmark = pytest.mark.skipif(not (os.path.isdir(REFERENCE) and os.path.exists(REF_LAWC)),
                                reason="the reference and its built lawc are only in the build container")

HEAD = "field T init 0 range -10 10\nfield U init 1\nparam k = 0.1\n"

CORPUS = [
    # precedence, associativity and signs
    HEAD + "term t\n  d T = -T ^ 2 + 2 ^ -1 * U - -U / 3 ^ 2 ^ 0.5\n",
    HEAD + "term t\n  d T = ((T)) * -(U - 1) / +k - 2 * 3 + 4 - 5\n",
    HEAD + "param a = -2 ^ 2\nparam b = 2 ^ 3 ^ 2\nparam c = (1 + 2) * 3 - 4 / 8\nterm t\n  d T = a * b * c * T\n",
    HEAD + "param a = 1.5e3\nparam b = .25\nparam c = 2E-2\nparam d = 1.2.3\nterm t\n  d T = a + b + c + d\n",
    HEAD + "param a = pi\nparam b = h\nparam c = dt\nparam d = rim\nparam e = maxof(3)\nparam f = sumof(2)\n"
           "param g = lap(2)\nparam i = 0 * rim\nparam j = exp(100)\nparam l = log(0)\nparam m = sqrt(-1)\n"
           "param n = clamp(5, 3, 1)\nparam o = min(1, 2) + max(3, 4) + abs(-2) + step(1) + pos(-3)\n"
           "term t\n  d T = a + b + c + d + e + f + g + i + j + l + m + n + o\n",
    # folding and sharing
    HEAD + "term t\n  d T = 0 * lap(T) + 1 * T + T * 1 + 0 + T - 0 + T / 1 + 0 / T + --T\n  d U = shift(T, 0, 0, 0) + shift(k, 1, 0, 0)\n",
    HEAD + "term a\n  d T = lap(T) + gx(T) * gy(U)\nterm b\n  d U = lap(T) + gx(T) * gy(U)\n",
    HEAD + "term t\n  d T = lap(k) + gx(2) + diffuse(k, 3) + adv(1, T, U, T)\n",
    # stencils, stages and the in-place decision
    HEAD + "term t\n  d T = diffuse(k, T) + adv(T, U, U, U) + shift(U, -2, 1, 0) + shift_t(T, 1, -1, 3)\n  d U = gx_t(T) + gy_t(U) + gz_t(T)\n",
    HEAD + "term t\n  d T = lap(lap(lap(T * U + exp(T) * sqrt(U) + T / (U + 1))))\n",
    HEAD + "term t\n  d T = tridiag_x(-k, 1 + 2 * k, -k, T) + tridiag_z(T, U, T, U * T)\n  d U = tridiag_y(1, 2, 1, U)\n",
    HEAD + "term t\n  d T = distance(U) + poisson(T) + poisson(U, 7) + maxof(T) + minof(U) * sumof(T)\n  max_dt = 1 / (1 + maxof(U))\n",
    HEAD + "field B init 0 static\nterm t\n  d T = fall(T, B, 1) + spread(T, B, 0.2) + topple(T, B, 1)\n  conserves T\n",
    HEAD + "term t\n  d T = T * U\nterm s\n  d U = -U\n  d T = U\n  conserves T + U\n",
    # lets, learned schemes, max_dt, conservation groups
    HEAD + "term t\n  let a = lap(T)\n  let b = a * a + U\n  d T = a + b\n  d U = b - a\n  conserves T U\n  max_dt = 0.15 * h * h / k\n",
    HEAD + "field V init 0\nterm t\n  d T = lap(T)\n  d U = lap(U)\n  conserves T + U V\n  conserves U + T\n",
    HEAD + "term t\n  let a = lap(T)\n  d T = a\n  learned d T = a * 2\n  learned max_dt = 2 * h * h / k\n  max_dt = h * h / k\n",
    # derives
    HEAD + "derive D = T * 2\nderive D += U\nderive E = D + 1\nderive D += 3\nterm t\n  d T = D + E\n",
    HEAD + "field S init 0\nderive S += T\nderive S = U\nterm t\n  d T = S\n",
    # declarations, show lines, creature blocks
    "field A init -1 range -inf inf unit K\nfield B init 2 range 0 4 clamp\nfield C init 0 static\n"
    "show A colour 1 0.5 0.2 above 0.1 full 2 alpha 0.3\nterm t\n  d A = B - A\n  d B = -B\n",
    HEAD + "creature grazer\n  radius 0.4 speed 2\n  state hunger init 0.3 range 0 1\n  d hunger = 0.1\n  seek = T\n"
           "term t\n  d T = lap(T)\n",
    HEAD + "term t\n  d T = lap(T)\ncreature wolf\n  hp 80\n  eat T = 0.1\nfield Late init 3\nterm s\n  d Late = T\n",
    # refusals: one law for each
    HEAD + "term t\n  d T = diffuse(q, T)\n",
    HEAD + "term t\n  d T = lap(T) +\n",
    HEAD + "term t\n  d T = lap(T\n",
    HEAD + "term t\n  d T = lap(T))\n",
    HEAD + "term t\n  d T = * T\n",
    HEAD + "term t\n  d T = T T\n",
    HEAD + "term t\n  d T = lap(T, T)\n",
    HEAD + "term t\n  d T = min(T)\n",
    HEAD + "term t\n  d T = poisson(T, 0)\n",
    HEAD + "term t\n  d T = poisson(T, T)\n",
    HEAD + "term t\n  d T = shift(T, 0.5, 0, 0)\n",
    HEAD + "term t\n  d T = shift(T, 9, 0, 0)\n",
    HEAD + "term t\n  d T = shift(T, -1, 0, U)\n",
    HEAD + "term t\n  d T = foo(T)\n",
    HEAD + "term t\n  d T = foo(q)\n",
    HEAD + "term t\n  d T = mlp(nope, T)\n",
    HEAD + "term t\n  d T = h(T)\n",
    HEAD + "term t\n  d T = T @ U\n",
    HEAD + "term t\n  d T = 1e999 * T\n",
    HEAD + "term t\n  max_dt = T\n  d T = lap(T)\n",
    HEAD + "term t\n  max_dt = dt\n  d T = lap(T)\n",
    HEAD + "term t\n  max_dt 3\n  d T = lap(T)\n",
    HEAD + "term t\n  d T = lap(T)\n  learned max_dt = T\n",
    HEAD + "term t\n  d T = lap(T)\n  learned max_dt = 1 2\n",
    HEAD + "term t\n  d T = lap(T)\n  learned d U = T\n",
    HEAD + "term t\n  d T = lap(T)\n  learned T = 1\n",
    HEAD + "term t\n",
    HEAD + "term t extra\n  d T = T\n",
    HEAD + "term t\n  d T = T\nterm t\n  d U = U\n",
    HEAD + "d T = T\n",
    HEAD + "let a = 1\n",
    HEAD + "conserves T\n",
    HEAD + "learned d T = T\n",
    HEAD + "term t\n  let a = 1\n  let a = 2\n  d T = a\n",
    HEAD + "term t\n  let k = 1\n  d T = k\n",
    HEAD + "term t\n  let lap = 1\n  d T = T\n",
    HEAD + "term t\n  let = 1\n  d T = T\n",
    HEAD + "term t\n  d Q = 1\n",
    HEAD + "term t\n  d T 1\n",
    HEAD + "term t\n  d T = lap(T)\n  conserves Q\n",
    HEAD + "term t\n  d T = lap(T)\n  conserves U\n",
    HEAD + "term t\n  d T = lap(T)\n  conserves 3\n",
    "field T init 0 static\nterm t\n  d T = lap(T)\n",
    HEAD + "derive T = 1\nterm t\n  d T = lap(T)\n",
    HEAD + "derive D = 1\nderive D = 2\n",
    HEAD + "derive Q + 1\n",
    HEAD + "derive 3 = 1\n",
    HEAD + "param p = T\n",
    HEAD + "param p 3\n",
    HEAD + "param k = 2\n",
    HEAD + "param T = 2\n",
    HEAD + "param field = 2\n",
    HEAD + "param p = 1 +\n",
    "field T init 0\nfield T init 1\n",
    "field lap init 0\n",
    "field T init x\n",
    "field T range 3 1\n",
    "field T unit 3\n",
    "field T wibble\n",
    "field\n",
    HEAD + "show T colour 1 1\n",
    HEAD + "show T colour 1 1 1 above 2 full 1\n",
    HEAD + "show Q colour 1 1 1 above 0 full 1\n",
    HEAD + "show T colour 1 1 1 above 0 full 1 glow 3\n",
    HEAD + "creature\n",
    HEAD + "creature a b\n",
    HEAD + "creature lap\n",
    HEAD + "creature a\ncreature a\n",
    HEAD + "net\n",
    HEAD + "net lap in 2 out 1 act tanh\n",
    HEAD + "net g in 2 out 1 act tanh\nnet g in 2 out 1 act tanh\n",
    HEAD + "net g in 0 out 1 act tanh\n",
    HEAD + "net g in 2 hidden 200 out 1 act tanh\n",
    HEAD + "net g in 2 out 1 act wibble\n",
    HEAD + "net g 3 in 2 out 1\n",
    HEAD + "net g in 2 out 1 act tanh file\n",
    HEAD + "net g in 2 out 1 act tanh\nterm t\n  d T = mlp(g, T)\n",
    HEAD + "net g in 2 out 1 act tanh\nterm t\n  d T = mlp(g T, U)\n",
    HEAD + "net g in 2 out 1 act tanh\nterm t\n  d T = mlp(g)\n",
    HEAD + "net g in 2 out 1 act tanh\nterm t\n  d T = mlp(g, )\n",
    HEAD + "net g in 2 out 1 act tanh\nterm t\n  d T = flux(g, T + 1)\n",
    HEAD + "net g in 5 out 3 act tanh\nterm t\n  d T = ldiffuse(g, k, T)\n",
    HEAD + "net g in 5 out 4 act tanh\nterm t\n  d T = ldiffuse(g, T, k)\n",
    HEAD + "3 + 4\n",
    HEAD + "wibble T\n",
    HEAD + "state x init 0\n",
    HEAD + "term t\n  d T = lap(T)\n# a comment\n  # another\n\n  d U = U # trailing\n",
    HEAD + "term t\n  d T = T $ 2\n",
    # nets that compile (untrained: no weight file in the working directory)
    HEAD + "net g in 2 hidden 8 8 out 3 act silu half\nnet w in 5 hidden 4 out 4 act relu\n"
           "term t\n  d T = mlp(g, T, U) + flux(g, T) + ldiffuse(w, k, U)\n",
    HEAD + "net g in 4 hidden 8 out 1 act linear\nterm t\n  d T = flux(g, T, U)\n  learned d T = flux(g, T, U) * 2\n",
]


def reference(tmp_path, text):
    path = tmp_path / "case.law"
    path.write_text(text)
    out = tmp_path / "ref.json"
    r = subprocess.run([REF_LAWC, "--ir-json", str(out), str(path)], capture_output=True, cwd=tmp_path, text=True)
    if r.returncode == 0:
        return True, out.read_text()
    lines = [ln for ln in (r.stderr + r.stdout).splitlines() if ln.startswith(str(path) + ":")]
    return False, lines


@pytest.mark.parametrize("index", range(len(CORPUS)))
def test_front_end_agrees_with_the_reference(tmp_path, index):
    import field_hip

    text = CORPUS[index]
    ok, ref = reference(tmp_path, text)
    path = str(tmp_path / "case.law")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        if ok:
            assert field_hip.Law(text, names=[path]).ir_json() == ref
        else:
            with pytest.raises(field_hip.FieldHipError) as e:
                field_hip.Law(text, names=[path])
            mine = [ln for ln in str(e.value).splitlines() if ln.startswith(path + ":")]
            assert mine == ref
    finally:
        os.chdir(cwd)
