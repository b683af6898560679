"""The ORACLE of the learned flux correction (oracle/learned.py), pinned before it judges the card:
its backward pass against central finite differences of its own loss (bf16 rounding off, FP64), and the properties
the construction promises: a mirrored field flips every face flux, a flat field has none, the correction sums to zero
over the lattice (conservation), and GELU against its erf form and its derivative against a difference quotient."""
import numpy as np
import pytest
from scipy.special import erf

from oracle import learned as L

NX, NY, COUNT, H = 7, 6, 2, 1.0 / 8
DU, DV, K, DT, SCALE = 1e-3, 5e-3, 5e-3, 1e-3, 0.05


def fields(seed):
    """s0 = (u0, v0) seeded N(0, 1); s1 a step of DT away (targets of moderate size: the loss stays O(1), so central
    differences are not swamped by its rounding)."""
    rng = np.random.default_rng(seed)
    u0, v0 = rng.standard_normal(NX * NY * COUNT), rng.standard_normal(NX * NY * COUNT)
    return [u0, v0, u0 + DT * rng.standard_normal(u0.size), v0 + DT * rng.standard_normal(v0.size)]


@pytest.fixture
def no_rounding():
    L.ROUND = False
    yield
    L.ROUND = True


def test_backward_matches_finite_differences(no_rounding):
    p = L.init_params(3, last_scale=1.0).astype(np.float64)
    u0, v0, u1, v1 = fields(4)
    loss, g = L.loss_and_grad(p, u0, v0, u1, v1, NX, NY, COUNT, H, DU, DV, K, DT, SCALE)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(1152, 6, replace=False), 1152 + rng.choice(64, 3, replace=False),
                          1216 + rng.choice(4096, 6, replace=False), 5312 + rng.choice(64, 3, replace=False),
                          5376 + rng.choice(128, 4, replace=False), [5504, 5505]])
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (L.loss_and_grad(pp, u0, v0, u1, v1, NX, NY, COUNT, H, DU, DV, K, DT, SCALE)[0]
              - L.loss_and_grad(pm, u0, v0, u1, v1, NX, NY, COUNT, H, DU, DV, K, DT, SCALE)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (i, fd, g[i])


def test_a_mirrored_field_flips_the_fluxes():
    p = L.init_params(6, last_scale=1.0)
    u, v, _, _ = fields(7)
    F = L.faces(p, u, v, NX, NY, COUNT, SCALE).reshape(COUNT, NY, NX, 2, 2)
    um = u.reshape(COUNT, NY, NX)[:, :, ::-1].reshape(-1)
    vm = v.reshape(COUNT, NY, NX)[:, :, ::-1].reshape(-1)
    Fm = L.faces(p, um, vm, NX, NY, COUNT, SCALE).reshape(COUNT, NY, NX, 2, 2)
    # the +x face of column x is the +x face of column NX - 2 - x in the mirror, with the sign flipped
    assert np.allclose(F[:, :, :NX - 1, 0], -Fm[:, :, NX - 2::-1, 0], atol=1e-12)


def test_flat_fields_have_no_flux_and_the_correction_is_conservative():
    p = L.init_params(8, last_scale=1.0)
    flat = np.full(NX * NY * COUNT, 0.7)
    assert np.max(np.abs(L.faces(p, flat, -flat, NX, NY, COUNT, SCALE))) < 1e-12
    u, v, _, _ = fields(9)
    c = L.correction(L.faces(p, u, v, NX, NY, COUNT, SCALE), NX, NY, COUNT, H)
    assert np.max(np.abs(c.sum(axis=(1, 2)))) < 1e-9 * np.max(np.abs(c))


def test_gelu_against_its_erf_form_and_its_derivative():
    z = np.linspace(-6, 6, 241)
    assert np.max(np.abs(L.gelu(z) - 0.5 * z * (1 + erf(z / np.sqrt(2))))) < 1e-3  # the tanh form's known accuracy
    e = 1e-6
    assert np.allclose(L.gelu_grad(z), (L.gelu(z + e) - L.gelu(z - e)) / (2 * e), atol=1e-8)


def test_bf16_rounds_to_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -3.14159, 65504.0], np.float32)  # 1 + 2^-8 is a tie: stays at 1
    got = L.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.015625
    assert abs(got[3] + 3.140625) < 1e-12
