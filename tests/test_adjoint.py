"""The adjoint law generator (CPU) and the checkpointed reverse mode on the card (GPU).

CPU: the generated adjoint laws compile and generate gfx950 code; an op without an adjoint is refused by name.
GPU: dL/d(initial fields) and dL/d(diffusivities) of L = sum c^2 after 12 steps of the three-field transport law,
checkpointed every 4 steps (three segments recomputed), against central differences of the FP64 oracle run; the
checkpoint interval does not change the result.
"""
import numpy as np
import pytest

from oracle import transport

SHAPE, H, DT, STEPS = (10, 9, 8), 0.125, 0.01, 12
PARAMS = {"Da": 0.05, "Db": 0.02, "Dc": 0.01}


def test_adjoint_laws_compile_and_generate():
    import field_hip

    for text in (transport.LAW, "field T init 0\nparam k = 0.3\nterm t\n  d T = diffuse(k, T) + 0.5 * gx(T) - T * T\n",
                 "field u init 0\nfield w init 0\nparam r = 2\nterm s\n  d u = lap(u) + exp(w) - r * sqrt(u)\n  d w = max(u, w) - clamp(w, 0, r)\n"):
        adj = field_hip.Law(text).adjoint()
        assert "term adjoint" in adj
        assert "void g_t0(K k" in field_hip.Law(adj).hip_source()


def test_an_op_without_an_adjoint_is_refused():
    import field_hip

    with pytest.raises(field_hip.FieldHipError, match="distance"):
        field_hip.Law("field A init 0\nfield B init 0\nterm t\n  d A = 0.1 * distance(B)\n").adjoint()


def start(seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(SHAPE))
    return {f: rng.uniform(0.2, 1.0, n) for f in ("a", "b", "c")}


@pytest.mark.gpu
def test_checkpointed_adjoint_matches_central_differences(device):
    import adjoint

    init = start(3)
    res = adjoint.vjp(transport.LAW, SHAPE, H, PARAMS, init, DT, STEPS, "c", every=4)
    D = [PARAMS["Da"], PARAMS["Db"], PARAMS["Dc"]]
    ref = transport.loss(SHAPE, H, DT, STEPS, D, init["a"], init["b"], init["c"])
    assert abs(res.loss - ref) < 1e-5 * ref
    for i, name in enumerate(("Da", "Db", "Dc")):
        e = 1e-4 * D[i]  # central: truncation ~(e/D)^2, rounding ~ulp(L)/e
        up, dn = list(D), list(D)
        up[i] += e
        dn[i] -= e
        fd = (transport.loss(SHAPE, H, DT, STEPS, up, init["a"], init["b"], init["c"])
              - transport.loss(SHAPE, H, DT, STEPS, dn, init["a"], init["b"], init["c"])) / (2 * e)
        print(name, res.param_grads[name], fd, flush=True)
        assert abs(res.param_grads[name] - fd) <= 1e-5 * abs(fd) + 1e-7
    rng = np.random.default_rng(4)
    for f in ("a", "b", "c"):
        for cell in rng.choice(int(np.prod(SHAPE)), 4, replace=False):
            e = 1e-4
            up = {k: v.copy() for k, v in init.items()}
            dn = {k: v.copy() for k, v in init.items()}
            up[f][cell] += e
            dn[f][cell] -= e
            fd = (transport.loss(SHAPE, H, DT, STEPS, D, up["a"], up["b"], up["c"])
                  - transport.loss(SHAPE, H, DT, STEPS, D, dn["a"], dn["b"], dn["c"])) / (2 * e)
            got = res.field_grads[f][cell]
            assert abs(got - fd) <= 1e-4 * abs(fd) + 1e-6, (f, cell, got, fd)


@pytest.mark.gpu
def test_the_checkpoint_interval_changes_nothing(device):
    import adjoint

    init = start(5)
    a = adjoint.vjp(transport.LAW, SHAPE, H, PARAMS, init, DT, STEPS, "c", every=12)
    b = adjoint.vjp(transport.LAW, SHAPE, H, PARAMS, init, DT, STEPS, "c", every=5)
    for f in ("a", "b", "c"):
        assert np.array_equal(a.field_grads[f], b.field_grads[f])
    assert a.param_grads == b.param_grads
