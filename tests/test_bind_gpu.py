"""A static field bound to the caller's device buffer (fe_solver_bind_field_device): the law reads the buffer where it
lies, the same bits as copying it in; a field the law writes cannot be bound; unbinding returns to the solver's own."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAW = """field T init 0 range -1000000 1000000
field w init 1 range 0 10 static
term t
  d T = w * diffuse(0.1, T) - 0.01 * w * T
"""


def make(field_hip, t0, w0):
    s = field_hip.Solver(field_hip.Law(LAW), (40, 24, 12), 0.25)
    s.set_field("T", t0)
    s.set_field("w", w0)
    s.refresh_derived()
    return s


def test_bound_field_equals_a_copied_one(device):
    import field_hip

    rng = np.random.default_rng(4)
    t0 = rng.random(40 * 24 * 12, dtype=np.float32)
    w0 = (0.5 + rng.random(40 * 24 * 12)).astype(np.float32)
    w1 = (0.2 + rng.random(40 * 24 * 12)).astype(np.float32)
    copied, bound = make(field_hip, t0, np.ones_like(w0)), make(field_hip, t0, np.ones_like(w0))
    buf = field_hip.DeviceArray.from_host(w0)
    copied.set_field_device("w", buf)
    bound.bind_field_device("w", buf)
    for _ in range(3):
        copied.step(0.05)
        bound.step(0.05)
    assert np.array_equal(copied.get_field("T").view(np.uint32), bound.get_field("T").view(np.uint32))
    # the buffer's new contents are what the bound solver reads next
    buf.upload(w1)
    copied.set_field_device("w", buf)
    copied.step(0.05)
    bound.step(0.05)
    assert np.array_equal(copied.get_field("T").view(np.uint32), bound.get_field("T").view(np.uint32))
    # back to its own array (which still holds ones)
    bound.bind_field_device("w", None)
    assert np.array_equal(bound.get_field("w"), np.ones_like(w0))


def test_a_written_field_cannot_be_bound(device):
    import field_hip

    s = make(field_hip, np.zeros(40 * 24 * 12, np.float32), np.ones(40 * 24 * 12, np.float32))
    buf = field_hip.DeviceArray(40 * 24 * 12)
    with pytest.raises(field_hip.FieldHipError, match="only a static field can be bound"):
        s.bind_field_device("T", buf)
