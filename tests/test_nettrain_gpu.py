"""A law net's trainer on the card (include/fe_nettrain.h) against its oracle (oracle/nettrain.py, float64, pinned by a
known answer, central differences and Adam's closed forms).

The card computes in fp32 on the matrix cores, as the reference's trainer computes in fp32; against the float64 oracle
the loss agrees to 1e-4 relative and the gradient to 2e-4 of its largest element. The rows are summed in a fixed
order, so the same rows give the same bits; the Adam step is the oracle's applied to the card's gradient, to 1e-3 of
the learning rate. Shapes: the world model's dynamics net (F + 2 -> 64 -> 64 -> F, tanh, 1024 rows), every
activation on a small net with a ragged last tile, 128-wide layers, more rows than one pass of blocks holds."""
import numpy as np
import pytest

from oracle import nettrain

pytestmark = pytest.mark.gpu


def rows_for(sizes, n, seed, masked=True):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, sizes[0])).astype(np.float32)
    t = rng.uniform(-1, 1, (n, sizes[-1])).astype(np.float32)
    m = rng.uniform(0.0, 2.0, n).astype(np.float32) if masked else None
    return x, t, m


def check_gradient(field_hip, sizes, act, n, seed, masked=True):
    image = nettrain.make_image(sizes, seed).astype(np.float32)
    tr = field_hip.NetTrainer(sizes, act)
    tr.set_image(image)
    x, t, m = rows_for(sizes, n, seed + 1, masked)
    loss, g = tr.gradient(x, t, m)
    wl, wg = nettrain.gradient(sizes, act, image.astype(np.float64), x, t, m)
    assert loss == pytest.approx(wl, rel=1e-4), (sizes, act, n)
    err = np.max(np.abs(g - wg)) / np.max(np.abs(wg))
    assert err < 2e-4, (sizes, act, n, err)
    again, g2 = tr.gradient(x, t, m)
    assert again == loss and np.array_equal(g.view(np.uint32), g2.view(np.uint32))
    return tr, image, x, t, m, g


def test_world_model_net_gradient(device):
    import field_hip

    check_gradient(field_hip, [24, 64, 64, 22], "tanh", 1024, 3, masked=False)


@pytest.mark.parametrize("act", nettrain.ACTS)
def test_every_activation_and_a_ragged_tile(device, act):
    import field_hip

    check_gradient(field_hip, [5, 7, 6, 3], act, 37, 21)


@pytest.mark.parametrize("sizes,n", [([128, 128, 128], 300), ([3, 40, 17, 1], 20000), ([1, 1], 1), ([9, 33, 65, 2], 4096)])
def test_wide_deep_and_many_rows(device, sizes, n):
    import field_hip

    check_gradient(field_hip, sizes, "silu", n, 31)


def test_image_padding_is_not_read_nor_written(device):
    """The reference's GPU image pads the normalisation to a multiple of 4 floats (CudaNet.cpp:103): what lies there
    changes nothing, and a step leaves it and the normalisation as they were."""
    import field_hip

    sizes = [5, 9, 4]
    image = nettrain.make_image(sizes, 51).astype(np.float32)
    padded = image.copy()
    padded[11] = 1e30
    x, t, m = rows_for(sizes, 50, 52)
    a, b = field_hip.NetTrainer(sizes, "relu"), field_hip.NetTrainer(sizes, "relu")
    a.set_image(image)
    b.set_image(padded)
    la, ga = a.gradient(x, t, m)
    lb, gb = b.gradient(x, t, m)
    assert la == lb and np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
    b.step(x, t, m)
    assert np.array_equal(b.image()[:12].view(np.uint32), padded[:12].view(np.uint32))


def test_adam_steps_follow_the_oracle(device):
    import field_hip

    sizes = [6, 32, 32, 4]
    tr, image, x, t, m, _ = check_gradient(field_hip, sizes, "tanh", 200, 41)
    base = nettrain.first_weight(sizes)
    ref = image.astype(np.float64)
    m1, m2 = np.zeros(ref.size - base), np.zeros(ref.size - base)
    lr = 1e-3
    for s in range(1, 4):
        _, g = tr.gradient(x, t, m)
        loss = tr.step(x, t, m, lr=lr, clip=0.05)
        nettrain.adam(ref, g.astype(np.float64), m1, m2, s, sizes, lr=lr, clip=0.05)
        got = tr.image()
        assert np.max(np.abs(got - ref)) < 1e-3 * lr, s
        assert np.array_equal(got[:base], image[:base])  # the normalisation stays
        ref = got.astype(np.float64)  # carry the card's weights (fp32) to the next step
        assert loss > 0


def test_learning_lowers_the_prediction_loss(device):
    """The reference's own check of its world model (loss after learning below half the first): a dynamics-shaped net
    fitted to a smooth map of its inputs."""
    import field_hip

    F = 10
    sizes = [F + 2, 64, 64, F]
    rng = np.random.default_rng(7)
    image = nettrain.make_image(sizes, 8, shift=0.0, scale=1.0, out_scale=1.0).astype(np.float32)
    tr = field_hip.NetTrainer(sizes, "tanh")
    tr.set_image(image)
    A = rng.uniform(-1, 1, (F, F + 2)).astype(np.float32) / np.sqrt(F + 2)
    x = rng.uniform(-1, 1, (4096, F + 2)).astype(np.float32)
    t = np.sin(x @ A.T).astype(np.float32)
    first = tr.step(x[:1024], t[:1024])
    for it in range(1, 300):
        pick = (it * 1024) % 4096
        tr.step(x[pick:pick + 1024], t[pick:pick + 1024])
    last, _ = tr.gradient(x, t)
    assert last < 0.5 * first, (first, last)


def test_refusals(device):
    import field_hip

    for sizes, act, what in [([4], "tanh", "2..9 sizes"), ([4, 0, 2], "tanh", "every size 1..128"), ([4, 129], "tanh", "every size 1..128"),
                             ([128] * 9, "relu", "bytes of LDS")]:
        with pytest.raises(field_hip.FieldHipError, match=what):
            field_hip.NetTrainer(sizes, act)
    with pytest.raises(field_hip.FieldHipError, match="activation"):
        field_hip.NetTrainer([2, 2], "softplus")
    tr = field_hip.NetTrainer([3, 8, 2], "relu")
    with pytest.raises(field_hip.FieldHipError, match="image of"):
        tr.set_image(np.zeros(5, np.float32))
    with pytest.raises(field_hip.FieldHipError, match="rows"):
        tr.gradient(np.zeros((4, 2), np.float32), np.zeros((4, 2), np.float32))
    with pytest.raises(field_hip.FieldHipError, match="no rows"):
        tr.gradient(np.zeros((0, 3), np.float32), np.zeros((0, 2), np.float32))


def test_matches_the_reference_cpu_net_code(device):
    """The card against the vectors the reference's own CPU net code wrote (net_forward_taped / net_backward, fp32,
    tests/golden/nettrain/ref_cases.npz): loss within 1e-4 relative, gradient within 2e-4 of its largest element."""
    import field_hip
    import golden_io

    seen = 0
    for name, sizes, act, image, x, t, m, loss, grad in golden_io.nettrain_reference_cases():
        tr = field_hip.NetTrainer(sizes, act)
        tr.set_image(image.astype(np.float32))
        got_loss, got = tr.gradient(x, t, m)
        assert got_loss == pytest.approx(loss, rel=1e-4), name
        assert np.max(np.abs(got - grad)) <= 2e-4 * np.max(np.abs(grad)), name
        seen += 1
    assert seen == 5


def test_rows_weighted_zero_give_zero_and_adam_leaves_the_weights(device):
    """Every row weighted 0: the loss and every gradient are exactly 0, and a step moves nothing (Adam's update of a
    zero gradient is 0 / (0 + eps))."""
    import field_hip

    sizes = [7, 20, 5]
    image = nettrain.make_image(sizes, 61).astype(np.float32)
    tr = field_hip.NetTrainer(sizes, "gelu")
    tr.set_image(image)
    x, t, _ = rows_for(sizes, 333, 62)
    m = np.zeros(333, np.float32)
    loss, g = tr.gradient(x, t, m)
    assert loss == 0.0 and not np.any(g)
    tr.step(x, t, m)
    assert np.array_equal(tr.image().view(np.uint32), image.view(np.uint32))
