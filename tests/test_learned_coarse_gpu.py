"""The learned coarse lattice end to end on the card, small: the engine's fine runs (128^2, one seed a plane) feed a
32^2 dataset, the net trains on 32 seeds and is judged on 8 it never saw, against the coarsened fine run. The test
asks what the trained net gains over the plain coarse law: a smaller error on the held-out seeds, everything
conserved, no fallbacks; and that the engine's fine data is the FP64 oracle's run to fp32 rounding."""
import numpy as np
import pytest

from oracle import diffreact

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    import field_hip
    import learned_coarse as lc
    from oracle import learned as L

    field_hip.require_device()
    s = lc.Setup(fine=128, factor=4, snapshots=30, every=4, dt_fine=2e-3)
    print("generating", flush=True)
    data = lc.generate(s, range(40), base_seed=100)
    scale = lc.target_scale(data)
    net = field_hip.LearnedNet(L.init_params(1, last_scale=0.0), scale)
    print(f"training, scale {scale:.3g}", flush=True)
    before = lc.eval_loss(net, data, slice(0, 8))
    lc.train(net, data, steps=600, batch=8, lr=3e-3, train_seeds=32)
    after = lc.eval_loss(net, data, slice(0, 8))
    return s, data, net, (before, after)


def test_the_engine_fine_data_is_the_fp64_run(trained):
    import learned_coarse as lc
    s, data, _, _ = trained
    u0, v0 = lc.initial(s, [32, 33], base_seed=100)
    U, V = diffreact.run64(lc.upsample(u0, s.factor), lc.upsample(v0, s.factor), s.fine, s.fine, 2, s.h_fine, s.du, s.dv, s.k,
                           s.dt_fine, s.every * 10)
    c = s.coarse
    Uc = U.reshape(2, c, s.factor, c, s.factor).mean(axis=(2, 4)).reshape(-1)
    got_u, _ = data.snapshot(10)
    got = got_u[32 * c * c:34 * c * c]
    assert np.max(np.abs(got - Uc)) < 1e-4 * max(1.0, np.max(np.abs(Uc)))


def test_training_lowers_the_loss(trained):
    _, _, _, (before, after) = trained
    print(f"loss on fixed training pairs: {before:.4g} -> {after:.4g}", flush=True)
    assert after < 0.7 * before


def test_the_trained_net_beats_the_plain_coarse_law_on_unseen_seeds(trained):
    import learned_coarse as lc
    s, data, net, _ = trained
    steps = s.snapshots
    ref_u, ref_v = data.snapshot(steps)
    c = s.coarse
    test = slice(32, 40)
    ref_u, ref_v = ref_u[32 * c * c:40 * c * c], ref_v[32 * c * c:40 * c * c]
    pu, pv = lc.rollout(net, data, test, steps, learned=False)
    lu, lv = lc.rollout(net, data, test, steps, learned=True)
    e_plain = np.sqrt(np.mean((pu - ref_u) ** 2) + np.mean((pv - ref_v) ** 2))
    e_learned = np.sqrt(np.mean((lu - ref_u) ** 2) + np.mean((lv - ref_v) ** 2))
    print(f"plain {e_plain:.4g} learned {e_learned:.4g}", flush=True)
    assert e_learned < 0.9 * e_plain
    assert net.fallbacks() == 0
