"""fe_lnet_step_noalloc, the learned step that allocates nothing, against fe_lnet_step through the C ABI.

Both run the same kernels; the no-allocation entry point keeps the face fluxes and the breach counter in the library's
static device array instead of the net's buffers. Its results must equal fe_lnet_step's bit for bit, on two calls in a
row (the second on a smaller batch, so stale scratch from the first would show), on the fallback path too, and it must
refuse a batch beyond its scratch.
"""
import ctypes

import numpy as np
import pytest

from oracle import learned as L

pytestmark = pytest.mark.gpu

DU, DV, K, DT, SCALE = 1e-3, 5e-3, 5e-3, 2e-3, 0.05


def lattices(seed, nx, ny, count):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(nx * ny * count).astype(np.float32),
            rng.standard_normal(nx * ny * count).astype(np.float32))


def both(field_hip, net, u, v, nx, ny, count, h, bound):
    """(u2, v2, fallback count) from fe_lnet_step (learned) and from fe_lnet_step_noalloc, called through the C ABI."""
    lib = field_hip.lib()
    du, dv = field_hip.DeviceArray.from_host(u), field_hip.DeviceArray.from_host(v)
    out = []
    for noalloc in (False, True):
        u2, v2 = field_hip.DeviceArray(u.size), field_hip.DeviceArray(u.size)
        before = net.fallbacks()
        if noalloc:
            rc = lib.fe_lnet_step_noalloc(net.ptr, du.ptr, dv.ptr, nx, ny, count, h, DU, DV, K, DT, bound, u2.ptr, v2.ptr)
        else:
            rc = lib.fe_lnet_step(net.ptr, du.ptr, dv.ptr, nx, ny, count, h, DU, DV, K, DT, 1, bound, u2.ptr, v2.ptr)
        assert rc == 0, lib.fe_lnet_error(net.ptr)
        field_hip.sync()
        out.append((u2.download(), v2.download(), net.fallbacks() - before))
    return out


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_step_equals_the_step_on_two_calls_in_a_row(device):
    import field_hip

    net = field_hip.LearnedNet(L.init_params(7, last_scale=1.0), SCALE)
    # First a larger batch, then a smaller one: the second call reads only what it wrote into the shared scratch.
    for seed, (nx, ny, count) in ((11, (128, 128, 6)), (12, (40, 36, 3))):
        u, v = lattices(seed, nx, ny, count)
        print(f"noalloc: {nx}x{ny}x{count}", flush=True)
        (au, av, af), (bu, bv, bf) = both(field_hip, net, u, v, nx, ny, count, 1.0 / nx, 1e3)
        assert same_bits(au, bu) and same_bits(av, bv)
        assert af == bf == 0


def test_noalloc_fallback_equals_the_step_fallback(device):
    """A bound every learned step breaks: both entry points replace the step by the plain law's and count it."""
    import field_hip

    net = field_hip.LearnedNet(L.init_params(8, last_scale=1.0), SCALE)
    for seed, (nx, ny, count) in ((21, (64, 64, 4)), (22, (32, 16, 2))):
        u, v = lattices(seed, nx, ny, count)
        (au, av, af), (bu, bv, bf) = both(field_hip, net, u, v, nx, ny, count, 1.0 / nx, 1e-3)
        assert af == bf == 1
        assert same_bits(au, bu) and same_bits(av, bv)


def test_noalloc_refuses_a_batch_beyond_its_scratch(device):
    import field_hip

    net = field_hip.LearnedNet(L.init_params(9, last_scale=1.0), SCALE)
    lib = field_hip.lib()
    nx, ny, count = 1024, 1024, 2  # 2^21 cells, twice the scratch; the pointers are never read
    rc = lib.fe_lnet_step_noalloc(net.ptr, ctypes.c_void_p(256), ctypes.c_void_p(256), nx, ny, count, 1.0, DU, DV, K, DT, 1e3,
                                  ctypes.c_void_p(256), ctypes.c_void_p(256))
    assert rc != 0
    assert b"scratch" in lib.fe_lnet_error(net.ptr)
