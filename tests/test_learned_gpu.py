"""The learned flux correction on the card (include/fe_learned.h) against its ORACLE (oracle/learned.py).

The oracle rounds to bf16 where the card does, so the two differ by fp32 accumulation and by the odd activation that
lands on the other side of a bf16 rounding boundary (fp32 tanh on the card, FP64 in the oracle): face fluxes agree
to 5e-3 of the largest at worst and 2e-4 on average, gradients to 1e-2 in norm.
"""
import numpy as np
import pytest

from oracle import learned as L

pytestmark = pytest.mark.gpu

NX, NY, COUNT, H = 40, 36, 3, 1.0 / 32
DU, DV, K, DT, SCALE = 1e-3, 5e-3, 5e-3, 2e-3, 0.05


def states(seed):
    rng = np.random.default_rng(seed)
    u0 = rng.standard_normal(NX * NY * COUNT).astype(np.float32)
    v0 = rng.standard_normal(NX * NY * COUNT).astype(np.float32)
    u1 = (u0 + DT * rng.standard_normal(u0.size)).astype(np.float32)
    v1 = (v0 + DT * rng.standard_normal(v0.size)).astype(np.float32)
    return u0, v0, u1, v1


def dev(a):
    import field_hip
    return field_hip.DeviceArray.from_host(np.ascontiguousarray(a, np.float32))


def test_face_fluxes_match_the_oracle(device):
    import field_hip

    p = L.init_params(1, last_scale=1.0)
    u, v, _, _ = states(2)
    net = field_hip.LearnedNet(p, SCALE)
    out = field_hip.DeviceArray(4 * NX * NY * COUNT)
    net.faces(dev(u), dev(v), NX, NY, COUNT, H, out)
    got = out.download().reshape(-1, 2)
    ref = L.faces(p, u, v, NX, NY, COUNT, SCALE)
    top = np.max(np.abs(ref))
    assert top > 1e-3
    err = np.abs(got - ref)
    assert err.max() < 5e-3 * top and err.mean() < 2e-4 * top, (err.max() / top, err.mean() / top)


def test_learned_and_plain_steps_match_the_oracle(device):
    import field_hip

    p = L.init_params(3, last_scale=1.0)
    u, v, _, _ = states(4)
    net = field_hip.LearnedNet(p, SCALE)
    u2, v2 = field_hip.DeviceArray(u.size), field_hip.DeviceArray(u.size)
    for learned in (False, True):
        net.step(dev(u), dev(v), NX, NY, COUNT, H, DU, DV, K, DT, u2, v2, learned=learned)
        ru, rv = L.step(p, u, v, NX, NY, COUNT, H, DU, DV, K, DT, SCALE, learned=learned)
        assert np.max(np.abs(u2.download() - ru)) < 1e-4 and np.max(np.abs(v2.download() - rv)) < 1e-4
    assert net.fallbacks() == 0


def test_the_correction_conserves_each_field(device):
    import field_hip

    p = L.init_params(5, last_scale=1.0)
    u, v, _, _ = states(6)
    net = field_hip.LearnedNet(p, SCALE)
    plain_u, plain_v = field_hip.DeviceArray(u.size), field_hip.DeviceArray(u.size)
    lu, lv = field_hip.DeviceArray(u.size), field_hip.DeviceArray(u.size)
    du_, dv_ = dev(u), dev(v)
    net.step(du_, dv_, NX, NY, COUNT, H, DU, DV, K, DT, plain_u, plain_v, learned=False)
    net.step(du_, dv_, NX, NY, COUNT, H, DU, DV, K, DT, lu, lv, learned=True)
    for a, b in ((lu, plain_u), (lv, plain_v)):
        diff = (a.download().astype(np.float64) - b.download()).reshape(COUNT, -1)
        assert np.max(np.abs(diff)) > 1e-5                          # the net does something
        assert np.max(np.abs(diff.sum(1))) < 1e-6 * NX * NY         # and moves nothing in or out


def test_gradient_matches_the_oracle_and_repeats_to_the_bit(device):
    import field_hip

    p = L.init_params(7, last_scale=1.0)
    u0, v0, u1, v1 = states(8)
    net = field_hip.LearnedNet(p, SCALE)
    d = [dev(a) for a in (u0, v0, u1, v1)]
    loss, g = net.gradient(*d, NX, NY, COUNT, H, DU, DV, K, DT)
    loss2, g2 = net.gradient(*d, NX, NY, COUNT, H, DU, DV, K, DT)
    assert loss == loss2 and np.array_equal(g.view(np.uint32), g2.view(np.uint32))
    ref_loss, ref_g = L.loss_and_grad(p, u0, v0, u1, v1, NX, NY, COUNT, H, DU, DV, K, DT, SCALE)
    assert abs(loss - ref_loss) < 1e-4 * ref_loss
    rel = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    assert rel < 1e-2, rel
    # each block of parameters on its own
    o = 0
    for name, size in (("W1", 1152), ("b1", 64), ("W2", 4096), ("b2", 64), ("W3", 128), ("b3", 2)):
        a, b = g[o:o + size], ref_g[o:o + size]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-9, name
        o += size


def test_training_on_the_card_recovers_a_teacher(device):
    import field_hip

    teacher = L.init_params(9, last_scale=1.0)
    u0, v0, _, _ = states(10)
    u1, v1 = L.step(teacher, u0, v0, NX, NY, COUNT, H, DU, DV, K, DT, SCALE)   # what the teacher's correction does
    d = [dev(a) for a in (u0, v0, u1, v1)]
    net = field_hip.LearnedNet(L.init_params(11, last_scale=0.0), SCALE)
    first, _ = net.gradient(*d, NX, NY, COUNT, H, DU, DV, K, DT)
    for _ in range(300):
        net.gradient(*d, NX, NY, COUNT, H, DU, DV, K, DT, want=False)
        net.adam(2e-3)
    last, _ = net.gradient(*d, NX, NY, COUNT, H, DU, DV, K, DT)
    print(f"loss {first:.4g} -> {last:.4g}", flush=True)
    assert last < 0.1 * first


def test_a_breach_falls_back_to_the_plain_step(device):
    import field_hip

    p = L.init_params(12, last_scale=1.0)
    u, v, _, _ = states(13)
    net = field_hip.LearnedNet(p, SCALE)
    u2, v2 = field_hip.DeviceArray(u.size), field_hip.DeviceArray(u.size)
    net.step(dev(u), dev(v), NX, NY, COUNT, H, DU, DV, K, DT, u2, v2, learned=True, bound=0.5)  # every state breaches
    ru, rv = L.step(p, u, v, NX, NY, COUNT, H, DU, DV, K, DT, SCALE, learned=False)
    assert net.fallbacks() == 1
    assert np.max(np.abs(u2.download() - ru)) < 1e-4 and np.max(np.abs(v2.download() - rv)) < 1e-4


def test_coarsen_is_the_block_average(device):
    import field_hip

    fine = np.random.default_rng(14).standard_normal(2 * 32 * 24).astype(np.float32)
    out = field_hip.DeviceArray(2 * 8 * 6)
    field_hip.coarsen(dev(fine), 32, 24, 2, 4, out)
    assert np.allclose(out.download(), L.coarsen(fine, 32, 24, 2, 4), atol=1e-6)
