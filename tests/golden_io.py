"""Reads the golden vectors (tests/golden/*.npz) the reference's CPU engine wrote."""
from __future__ import annotations

import contextlib
import json
import os
import tempfile

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    init = {k[5:]: z[k] for k in z.files if k.startswith("init_")}
    final = {k[6:]: z[k] for k in z.files if k.startswith("final_")}
    return meta, init, final


def names(prefix: str = ""):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz") and f.startswith(prefix))


def make_law(field_hip, meta):
    """The case's law, its nets read from NAME.net files under assets/nets/ of the working directory, as the
    reference's lawc read them when it wrote the case."""
    nets = meta.get("nets", {})
    if not nets:
        return field_hip.Law(meta["law"])
    with tempfile.TemporaryDirectory() as tmp, _cwd(tmp):
        os.makedirs(os.path.join("assets", "nets"))
        for name, text in nets.items():
            with open(os.path.join("assets", "nets", name + ".net"), "w") as f:
                f.write(text)
        return field_hip.Law(meta["law"])


@contextlib.contextmanager
def _cwd(path):
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


def same(meta, got, want) -> bool:
    """The reference's bits, or within the case's atol where it names one (nets through the card's tanhf / expf)."""
    atol = meta.get("atol")
    if atol is None:
        return got.shape == want.shape and bool(np.array_equal(got.view(np.uint32), want.view(np.uint32)))
    return got.shape == want.shape and bool(np.allclose(got, want, rtol=0.0, atol=atol))


def nettrain_reference_cases():
    """The vectors the reference's own CPU net code wrote for the net trainer (tests/golden/make_nettrain_golden.py): per
    case (name, sizes, activation, the image in the reference's GPU layout as float64, x, t, m, the rows' mean loss,
    the mean gradient)."""
    from oracle import nettrain

    z = np.load(os.path.join(GOLDEN, "nettrain", "ref_cases.npz"), allow_pickle=False)
    names = sorted({k[: -len("_sizes")] for k in z.files if k.endswith("_sizes")})
    for name in names:
        sizes = [int(v) for v in z[f"{name}_sizes"]]
        image = np.zeros(nettrain.layout(sizes)[1])
        n = sizes[0]
        image[:n] = z[f"{name}_shift"]
        image[n:2 * n] = z[f"{name}_scale"]
        image[2 * n] = z[f"{name}_out_scale"]
        image[nettrain.first_weight(sizes):] = z[f"{name}_weights"]
        yield (name, sizes, nettrain.ACTS[int(z[f"{name}_act"])], image, z[f"{name}_x"], z[f"{name}_t"], z[f"{name}_m"],
               float(z[f"{name}_loss"]), z[f"{name}_grad"])
