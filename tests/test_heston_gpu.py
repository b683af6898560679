"""The card's batched Heston ADI (include/fe_adi.h) against the FP64 oracle of the same scheme (oracle/heston_ref.c).

fp32 on the card, FP64 in the oracle: lattices agree to a relative 2e-5 of the largest value and prices to 2e-4
absolute after 32 steps, through every launch shape (lanes a row, waves a block, threads a column block), both
schemes, uneven batches and graph capture.
"""
import numpy as np
import pytest

from oracle import heston

pytestmark = pytest.mark.gpu

LATTICE_RTOL = 2e-5
PRICE_ATOL = 2e-4


def draw(n, seed, shapes=((128, 64), (256, 128), (512, 256))):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(shapes), n)
    nx = np.array([shapes[i][0] for i in idx], np.int32)
    ny = np.array([shapes[i][1] for i in idx], np.int32)
    opts = np.stack([rng.uniform(0.5, 5, n), rng.uniform(0.01, 0.1, n), rng.uniform(0.1, 0.8, n), rng.uniform(-0.9, -0.1, n),
                     np.full(n, 0.02), rng.uniform(80, 120, n), rng.uniform(0.25, 2, n), rng.uniform(0.01, 0.1, n)], 1)
    return nx, ny, opts


def compare(batch, nx, ny, opts, steps, scheme, theta):
    prices, V, offsets = heston.batch64(nx, ny, opts, steps, scheme, theta)
    got = batch.prices()
    worst = 0.0
    for o in range(len(nx)):
        ref = V[offsets[o]:offsets[o] + nx[o] * ny[o]]
        lat = batch.lattice(o)
        worst = max(worst, float(np.max(np.abs(lat - ref)) / np.max(np.abs(ref))))
    return worst, float(np.max(np.abs(got - prices)))


@pytest.mark.parametrize("scheme,theta", [(0, 0.5), (1, 0.5 + np.sqrt(3) / 6)], ids=["cs", "hv"])
def test_batch_matches_the_fp64_oracle(device, scheme, theta):
    import field_hip

    nx, ny, opts = draw(9, 11)
    print(f"batch of {len(nx)} scheme {scheme}", flush=True)
    b = field_hip.HestonBatch(nx, ny, opts, 32, scheme, theta)
    b.step(32)
    lat, price = compare(b, nx, ny, opts, 32, scheme, theta)
    assert lat < LATTICE_RTOL and price < PRICE_ATOL, (lat, price)


def test_every_launch_shape_matches(device):
    import field_hip

    shapes = ((64, 32), (128, 64), (256, 128), (512, 256), (1024, 128))
    nx, ny, opts = draw(10, 12, shapes)
    print("shapes", list(zip(nx, ny)), flush=True)
    ref_prices, V, offsets = heston.batch64(nx, ny, opts, 16)
    b = field_hip.HestonBatch(nx, ny, opts, 16)
    for cls, shape in enumerate(b.shapes()):
        for lanes in (16, 32, 64):
            k = shape["nx"] // lanes
            if k not in (2, 4, 8, 16):
                continue
            ny_max = int(ny[nx == shape["nx"]].max())
            # -64: d' in LDS; -8, -16: partitioned columns, that many rows a thread (at most 16 waves a block)
            cols = [(1, 64), (2, 128), (4, 256), (2, -64), (4, -16)] + ([(1, -8)] if ny_max <= 128 else [])
            for row_waves, col_threads in cols:
                print(f"class {shape['nx']} lanes {lanes} waves {row_waves} cols {col_threads}", flush=True)
                b.set_shape(cls, lanes, row_waves, col_threads)
                b.reset()
                b.step(16)
                got = b.prices()
                assert np.max(np.abs(got - ref_prices)) < PRICE_ATOL
                for o in np.flatnonzero(nx == shape["nx"]):
                    ref = V[offsets[o]:offsets[o] + nx[o] * ny[o]]
                    assert np.max(np.abs(b.lattice(o) - ref)) / np.max(np.abs(ref)) < LATTICE_RTOL


@pytest.mark.parametrize("scheme,theta", [(0, 0.5), (1, 0.5 + np.sqrt(3) / 6)], ids=["cs", "hv"])
def test_partitioned_columns_on_uneven_heights(device, scheme, theta):
    """Column passes split over waves (rows a thread in registers, the chunk equations through LDS): heights that are
    no multiple of the chunk, options of one class with different heights, a class needing 16 waves."""
    import field_hip

    rng = np.random.default_rng(31)
    nx = np.array([128, 128, 128, 256, 256, 64, 64], np.int32)
    ny = np.array([21, 100, 64, 256, 131, 5, 37], np.int32)
    _, _, opts = draw(len(nx), 32)
    opts[:, 6] = rng.uniform(0.25, 1.0, len(nx))
    print("lattices", list(zip(nx, ny)), flush=True)
    ref_prices, V, offsets = heston.batch64(nx, ny, opts, 16, scheme, theta)
    b = field_hip.HestonBatch(nx, ny, opts, 16, scheme, theta)
    for rows_a_thread in (8, 16):
        for cls, shape in enumerate(b.shapes()):
            ny_max = int(ny[nx == shape["nx"]].max())
            if -(-ny_max // rows_a_thread) <= 16:
                b.set_shape(cls, shape["lanes"], shape["row_waves"], -rows_a_thread)
        print(f"rows a thread {rows_a_thread}: {b.shapes()}", flush=True)
        b.reset()
        b.step(16)
        assert np.max(np.abs(b.prices() - ref_prices)) < PRICE_ATOL
        for o in range(len(nx)):
            ref = V[offsets[o]:offsets[o] + nx[o] * ny[o]]
            assert np.max(np.abs(b.lattice(o) - ref)) / np.max(np.abs(ref)) < LATTICE_RTOL, (o, nx[o], ny[o])


def test_captured_graph_equals_queued_steps_to_the_bit(device):
    import field_hip

    nx, ny, opts = draw(6, 13)
    a = field_hip.HestonBatch(nx, ny, opts, 24)
    a.step(24)
    g = field_hip.HestonBatch(nx, ny, opts, 24)
    g.capture(24)
    for replay in range(3):   # replays start from the payoff each time, with eager work between them
        g.reset()
        g.launch()
        g.sync()
        assert np.array_equal(a.prices(), g.prices())
        for o in range(len(nx)):
            assert np.array_equal(a.lattice(o).view(np.uint32), g.lattice(o).view(np.uint32))


def test_refuses_lattices_it_cannot_run(device):
    import field_hip

    with pytest.raises(field_hip.FieldHipError, match="lattice"):
        field_hip.HestonBatch([96], [64], [[1.5, 0.04, 0.3, -0.9, 0.025, 100, 0.5, 0.04]], 8)
    with pytest.raises(field_hip.FieldHipError, match="positive"):
        field_hip.HestonBatch([128], [64], [[1.5, 0.04, 0.0, -0.9, 0.025, 100, 0.5, 0.04]], 8)


def test_tuner_picks_valid_shapes_and_results_hold(device):
    import field_hip

    nx, ny, opts = draw(12, 14)
    ref_prices, _, _ = heston.batch64(nx, ny, opts, 16)
    b = field_hip.HestonBatch(nx, ny, opts, 16)
    report = b.tune(1)
    print(report, flush=True)
    assert sorted(r["nx"] for r in report) == sorted(set(int(v) for v in nx))
    for r in report:
        assert r["lanes"] in (16, 32, 64) and r["row_waves"] in (1, 2, 4) and r["col_threads"] in (64, 128, 256)
    b.step(16)
    assert np.max(np.abs(b.prices() - ref_prices)) < PRICE_ATOL


def test_column_sweeps_in_lds_give_the_same_bits(device):
    import field_hip

    shapes = ((128, 64), (256, 128), (512, 256))
    nx, ny, opts = draw(9, 21, shapes)
    runs = []
    for cols in (64, -64):
        b = field_hip.HestonBatch(nx, ny, opts, 12)
        for cls, shape in enumerate(b.shapes()):
            b.set_shape(cls, shape["lanes"], shape["row_waves"], cols)
        b.reset()
        b.step(12)
        runs.append(np.concatenate([b.lattice(o) for o in range(len(nx))]))
    assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))
