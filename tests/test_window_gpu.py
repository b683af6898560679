"""A moving window (Solver::shift_window, src/field/Solver.hpp:40-42): the card slides its copy of every field and
receives only the arriving boxes. Held to the bits of a solver that is handed the same slid lattice whole, as the
reference's ring gate holds its ring to a plain solver (src/tests/cuda_parity_smoke.cpp:163-233)."""
import numpy as np
import pytest

import golden_io

pytestmark = pytest.mark.gpu

# (dx, dy, dz): each move's arriving boxes are the cells whose source lies outside the lattice
MOVES = [(3, 0, 0), (0, -2, 1), (-1, 1, -2), (0, 0, 0)]


def arriving_boxes(shape, d):
    nx, ny, nz = shape
    boxes = []
    for axis, n in enumerate((nx, ny, nz)):
        if d[axis] == 0:
            continue
        lo, hi = (n - d[axis], n) if d[axis] > 0 else (0, -d[axis])
        box = [0, 0, 0, nx, ny, nz]
        box[axis], box[axis + 3] = lo, hi
        boxes.append(box)
    return boxes


def slide(a, d, shape):
    """What is now at c was at c + d, where that lies inside; the rest keeps its value (a box will fill it)."""
    nx, ny, nz = shape
    v = a.reshape(nz, ny, nx)
    out = v.copy()
    dx, dy, dz = d
    xs = slice(max(0, -dx), min(nx, nx - dx)), slice(max(0, dx), min(nx, nx + dx))
    ys = slice(max(0, -dy), min(ny, ny - dy)), slice(max(0, dy), min(ny, ny + dy))
    zs = slice(max(0, -dz), min(nz, nz - dz)), slice(max(0, dz), min(nz, nz + dz))
    out[zs[0], ys[0], xs[0]] = v[zs[1], ys[1], xs[1]]
    return out.ravel()


def fill(a, box, values, shape):
    nx, ny, nz = shape
    v = a.reshape(nz, ny, nx)
    x0, y0, z0, x1, y1, z1 = box
    v[z0:z1, y0:y1, x0:x1] = values.reshape(z1 - z0, y1 - y0, x1 - x0)
    return v.ravel()


@pytest.mark.parametrize("name", ["heat_20x14x9", "multi", "kitchen", "op_poisson", "op_distance"])
@pytest.mark.parametrize("lazy", [False, True])
def test_moved_window_equals_a_whole_upload(device, name, lazy):
    import field_hip

    meta, init, _ = golden_io.load(name)
    nx, ny, nz, h = meta["grid"]
    shape = (nx, ny, nz)
    law = golden_io.make_law(field_hip, meta)
    names = law.field_names()
    ring = field_hip.Solver(law, shape, h)
    plain = field_hip.Solver(golden_io.make_law(field_hip, meta), shape, h)
    for s in (ring, plain):
        for p, value in meta["params"].items():
            s.set_param(p, value)
        for f, values in init.items():
            s.set_field(f, values)
        s.refresh_derived()
        s.set_lazy_readback(lazy)
    rng = np.random.default_rng(5)
    for move, d in enumerate(MOVES):
        for _ in range(2):
            ring.step(meta["dt"])
            plain.step(meta["dt"])
        boxes = arriving_boxes(shape, d)
        arriving, host = [], {}
        for f in names:
            cur = slide(plain.get_field(f), d, shape)
            lo, hi = float(np.min(init[f])) if f in init else 0.0, float(np.max(init[f])) if f in init else 1.0
            for box in boxes:
                cells = (box[3] - box[0]) * (box[4] - box[1]) * (box[5] - box[2])
                vals = rng.uniform(lo, hi, cells).astype(np.float32)
                arriving.append(vals)
                cur = fill(cur, box, vals, shape)
            host[f] = cur
        print(f"{name} lazy {lazy} move {move} {d}: {len(boxes)} boxes", flush=True)
        ring.shift_window(*d, boxes, np.concatenate(arriving) if arriving else np.zeros(0, np.float32))
        for f in names:
            plain.set_field(f, host[f])
        for _ in range(2):
            ring.step(meta["dt"])
            plain.step(meta["dt"])
        for f in names:
            a, b = ring.get_field(f), plain.get_field(f)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{name} move {move} field {f}"
