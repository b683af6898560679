"""`half` law nets on the matrix cores (fe_net_half_forward): fp16 operands, fp32 accumulation, held to an fp32 NumPy
forward pass of the same net (the reference's CPU evaluates `half` nets in fp32, src/field/nn/Net.cpp:132-152).

The bound is fp16's: inputs, weights and hidden activations each round to 11 bits, so outputs agree to about 1e-2 of
their spread; a wrong fragment layout (rows or the k chunk misplaced) is off by the outputs' whole size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ACTS = {"tanh": 0, "relu": 1, "silu": 2, "linear": 3, "gelu": 4}


def forward32(sizes, act, out_scale, w, x):
    a = x.astype(np.float64)
    off = 0
    for l, (ni, no) in enumerate(zip(sizes[:-1], sizes[1:])):
        W = w[off:off + ni * no].reshape(no, ni).astype(np.float64)
        b = w[off + ni * no:off + ni * no + no].astype(np.float64)
        off += ni * no + no
        a = a @ W.T + b
        if l + 1 < len(sizes) - 1:
            if act == "tanh":
                a = np.tanh(a)
            elif act == "relu":
                a = np.maximum(a, 0)
            elif act == "silu":
                a = a / (1 + np.exp(-a))
            elif act == "gelu":
                a = 0.5 * a * (1 + np.tanh(0.7978845608028654 * (a + 0.044715 * a ** 3)))
    return a * out_scale


def weights(sizes, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for ni, no in zip(sizes[:-1], sizes[1:]):
        lim = np.sqrt(6.0 / (ni + no))
        parts += [rng.uniform(-lim, lim, ni * no), rng.uniform(-0.2, 0.2, no)]
    return np.concatenate(parts).astype(np.float32)


CASES = [([4, 8, 1], "relu", 37), ([5, 32, 24, 4], "tanh", 1000), ([18, 64, 64, 2], "gelu", 4096 + 5),
         ([2, 16, 1], "silu", 16), ([7, 128, 128, 9], "tanh", 300), ([33, 40, 3], "linear", 71)]


@pytest.mark.parametrize("sizes,act,rows", CASES, ids=lambda v: str(v))
def test_half_net_matches_fp32_to_fp16(device, sizes, act, rows):
    import field_hip

    w = weights(sizes, len(sizes) * 7 + rows)
    x = np.random.default_rng(rows).uniform(-1.5, 1.5, (rows, sizes[0])).astype(np.float32)
    got = field_hip.net_half_forward(sizes, ACTS[act], 0.8, w, x)
    want = forward32(sizes, act, 0.8, w, x)
    spread = np.max(np.abs(want)) + 1e-3
    err = np.max(np.abs(got - want))
    print(f"{sizes} {act} rows {rows}: max |diff| {err:.3g} of {spread:.3g}", flush=True)
    assert np.all(np.isfinite(got))
    assert err <= 1.5e-2 * spread + 1e-3
    # the same rows twice give the same bits
    again = field_hip.net_half_forward(sizes, ACTS[act], 0.8, w, x)
    assert np.array_equal(got.view(np.uint32), again.view(np.uint32))
