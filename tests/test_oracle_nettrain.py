"""The net trainer's oracle (oracle/nettrain.py) pinned: a hand-computed known answer for the forward pass and the loss,
the analytic gradient against central differences of the loss for every activation (a dropped term, a transposed
operand or a wrong derivative fails it), the activations against their textbook definitions, Adam's first step
(lr sign(g) whatever g's size) and clipping, a linear regression that Adam fits, and the reference's own CPU net code
(net_forward_taped / net_backward) on five nets (tests/golden/nettrain/ref_cases.npz)."""
import math

import golden_io
import numpy as np
import pytest
from scipy.special import erf, expit

from oracle import nettrain


def test_known_answer_of_a_one_one_one_net():
    # x 0.5, shift 0.1, scale 2 -> 0.8; W1 0.5, b1 0.1 -> 0.5; tanh -> 0.46211715726000974; W2 2, b2 -0.3 ->
    # 0.6242343145200195; out_scale 0.8 -> 0.4993874516160156; t 0.25 -> e 0.2493874516160156. The image: shift,
    # scale, out_scale, one float of padding (the first weight at 4), W1, b1, W2, b2.
    image = np.array([0.1, 2.0, 0.8, 0.0, 0.5, 0.1, 2.0, -0.3])
    y = nettrain.forward([1, 1, 1], "tanh", image, np.array([[0.5]]))
    assert y[0, 0] == pytest.approx(0.4993874516160156, rel=1e-14)
    loss, g = nettrain.gradient([1, 1, 1], "tanh", image, np.array([[0.5]]), np.array([[0.25]]), np.array([2.0]))
    e = 0.2493874516160156
    assert loss == pytest.approx(0.5 * 2.0 * e * e, rel=1e-13)
    d2 = 2.0 * e * 0.8  # m e out_scale
    d1 = d2 * 2.0 * (1.0 - 0.46211715726000974**2)  # through W2 and tanh'
    np.testing.assert_allclose(g, [d1 * 0.8, d1, d2 * 0.46211715726000974, d2], rtol=1e-13)


@pytest.mark.parametrize("act", nettrain.ACTS)
def test_gradient_is_the_loss_derivative(act):
    sizes = [5, 7, 6, 3]
    image = nettrain.make_image(sizes, 11)
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, (9, 5))
    t = rng.uniform(-1, 1, (9, 3))
    m = rng.uniform(0.2, 2.0, 9)
    _, g = nettrain.gradient(sizes, act, image, x, t, m)
    base = nettrain.first_weight(sizes)
    h = 1e-6
    for p in range(base, image.size):
        up, dn = image.copy(), image.copy()
        up[p] += h
        dn[p] -= h
        lu, _ = nettrain.gradient(sizes, act, up, x, t, m)
        ld, _ = nettrain.gradient(sizes, act, dn, x, t, m)
        fd = (lu - ld) / (2 * h)
        assert g[p - base] == pytest.approx(fd, rel=1e-5, abs=1e-9), (act, p)


def test_activations_are_the_textbook_ones():
    x = np.linspace(-4, 4, 81)
    np.testing.assert_allclose(nettrain.act("silu", x), x * expit(x), rtol=1e-14)
    np.testing.assert_allclose(nettrain.act("relu", x), np.where(x > 0, x, 0.0))
    # the tanh form of GELU stays within 1e-3 of the erf form
    assert np.max(np.abs(nettrain.act("gelu", x) - 0.5 * x * (1 + erf(x / math.sqrt(2))))) < 1e-3
    for a in nettrain.ACTS:
        xs = x + 0.013  # off relu's kink
        fd = (nettrain.act(a, xs + 1e-6) - nettrain.act(a, xs - 1e-6)) / 2e-6
        np.testing.assert_allclose(nettrain.act_grad(a, xs), fd, rtol=1e-6, atol=1e-8)


def test_adams_first_step_is_lr_times_the_sign_and_clips():
    sizes = [2, 3]
    image = nettrain.make_image(sizes, 3)
    base = nettrain.first_weight(sizes)
    n = image.size - base
    g = np.linspace(-50, 50, n)
    g[n // 2] = 1e-3
    before = image.copy()
    m1, m2 = np.zeros(n), np.zeros(n)
    nettrain.adam(image, g, m1, m2, 1, sizes, lr=0.01, clip=5.0)
    step = before[base:] - image[base:]
    np.testing.assert_allclose(step, 0.01 * np.sign(g) * np.abs(np.clip(g, -5, 5)) / (np.abs(np.clip(g, -5, 5)) + 1e-8), rtol=1e-9)
    np.testing.assert_allclose(np.abs(step[np.abs(g) > 1e-2]), 0.01, rtol=1e-6)
    assert np.array_equal(image[:base], before[:base])  # the normalisation is not trained
    np.testing.assert_allclose(m1, 0.1 * np.clip(g, -5, 5), rtol=1e-12)


def test_adam_fits_a_linear_map():
    sizes = [3, 2]
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (2, 3))
    c = rng.uniform(-0.5, 0.5, 2)
    x = rng.uniform(-1, 1, (256, 3))
    t = x @ A.T + c
    image = np.zeros(nettrain.layout(sizes)[1])
    image[3:6] = 1.0
    image[6] = 1.0
    base = 8  # 2 x 3 + 1 floats of normalisation, padded to 8
    m1, m2 = np.zeros(image.size - base), np.zeros(image.size - base)
    first, _ = nettrain.gradient(sizes, "linear", image, x, t)
    for s in range(1, 3001):
        _, g = nettrain.gradient(sizes, "linear", image, x, t)
        nettrain.adam(image, g, m1, m2, s, sizes, lr=0.01, clip=5.0)
    last, _ = nettrain.gradient(sizes, "linear", image, x, t)
    assert last < 1e-8 * first
    W = image[base:base + 6].reshape(2, 3)
    np.testing.assert_allclose(W, A, atol=1e-4)
    np.testing.assert_allclose(image[base + 6:], c, atol=1e-4)


def test_the_image_is_the_reference_gpu_layout():
    """CudaNet.cpp:97-110: the first weight at 2 in + 1 padded to a multiple of 4, then per layer (in + 1) x out floats."""
    for sizes, first, total in [([1, 1, 1], 4, 4 + 2 + 2), ([18, 64, 64, 2], 40, 40 + 19 * 64 + 65 * 64 + 65 * 2),
                                ([24, 64, 64, 22], 52, 52 + 25 * 64 + 65 * 64 + 65 * 22)]:
        assert nettrain.first_weight(sizes) == first
        assert nettrain.layout(sizes)[1] == total
    # the padding is not read: changing it changes nothing
    sizes = [5, 7, 3]
    image = nettrain.make_image(sizes, 9)
    x = np.random.default_rng(1).uniform(-1, 1, (4, 5))
    other = image.copy()
    other[11] = 123.0
    np.testing.assert_array_equal(nettrain.forward(sizes, "tanh", image, x), nettrain.forward(sizes, "tanh", other, x))


def test_matches_the_reference_cpu_net_code():
    """The oracle (float64) against the reference's net_forward_taped / net_backward (fp32): loss to 1e-6 relative,
    gradient to 1e-5 of its largest element."""
    seen = 0
    for name, sizes, act, image, x, t, m, loss, grad in golden_io.nettrain_reference_cases():
        wl, wg = nettrain.gradient(sizes, act, image, x.astype(np.float64), t.astype(np.float64), m.astype(np.float64))
        assert wl == pytest.approx(loss, rel=1e-6), name
        assert np.max(np.abs(wg - grad)) <= 1e-5 * np.max(np.abs(wg)), name
        seen += 1
    assert seen == 5
