"""The learned coarse lattice for 2-D diffusion-reaction: data from the engine's fine runs, training and inference on
the card, evaluation against a fine reference.

    du/dt = Du lap u + u - u^3 - k - v,   dv/dt = Dv lap v + u - v       on [-1, 1]^2, zero flux at the edges

1. `generate`: the engine (this package's law compiler and HIP backend) runs the fine law, one seed a z-plane of one
   lattice; every `every` fine steps each plane is block-averaged to the coarse lattice on the card, into a dataset
   resident in HBM (snapshot t of seed s: U[t, s], V[t, s]).
2. `train`: the net (include/fe_learned.h) is fitted on pairs (snapshot t, snapshot t + 1) of the training seeds: its
   conservative face fluxes should make the coarse law's step land on the next coarsened fine snapshot.
3. `rollout`: the coarse law, with or without the net, from coarse states, many seeds at once, snapshot by snapshot.

The reference trains its law nets outside the engine (bench/model2_front.py:18,136-151, in JAX); here the fine data,
the coarsening, the training and the inference all run on the card.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

import field_hip

DIFFREACT_LAW = """# 2-D diffusion-reaction, one seed a z-plane (the Laplacian reaches along x and y only)
field u init 0
field v init 0
param Du = 0.001
param Dv = 0.005
param kr = 0.005
term react
  d u = Du * (shift(u, 1, 0, 0) + shift(u, -1, 0, 0) + shift(u, 0, 1, 0) + shift(u, 0, -1, 0) - 4 * u) / (h * h) + u - u * u * u - kr - v
  d v = Dv * (shift(v, 1, 0, 0) + shift(v, -1, 0, 0) + shift(v, 0, 1, 0) + shift(v, 0, -1, 0) - 4 * v) / (h * h) + u - v
"""


@dataclass
class Setup:
    fine: int = 512          # fine cells a side
    factor: int = 4          # fine cells a coarse cell, a side
    snapshots: int = 100     # coarse snapshots a seed after the start
    every: int = 16          # fine steps a snapshot
    dt_fine: float = 5e-4
    du: float = 1e-3
    dv: float = 5e-3
    k: float = 5e-3
    width: float = 2.0       # the square's side

    @property
    def coarse(self) -> int:
        return self.fine // self.factor

    @property
    def h_fine(self) -> float:
        return self.width / self.fine

    @property
    def h_coarse(self) -> float:
        return self.width / self.coarse

    @property
    def dt_coarse(self) -> float:
        return self.every * self.dt_fine


def initial(setup: Setup, seeds, base_seed: int = 0):
    """u, v ~ N(0, 1) on the coarse lattice, one draw a seed; the fine start is each coarse cell repeated factor^2
    times (so the coarse start is exactly the fine one block-averaged)."""
    c = setup.coarse
    us, vs = [], []
    for s in seeds:
        rng = np.random.default_rng(base_seed + int(s))
        us.append(rng.standard_normal((c, c)))
        vs.append(rng.standard_normal((c, c)))
    return np.array(us, np.float32), np.array(vs, np.float32)


def upsample(a: np.ndarray, factor: int) -> np.ndarray:
    return np.repeat(np.repeat(a, factor, axis=-2), factor, axis=-1)


class Dataset:
    """U[t][s], V[t][s]: coarse snapshots t = 0..snapshots of seeds s, resident on the card."""

    def __init__(self, setup: Setup, seeds: int):
        self.setup = setup
        self.seeds = seeds
        self.cells = setup.coarse ** 2
        self.U = field_hip.DeviceArray((setup.snapshots + 1) * seeds * self.cells)
        self.V = field_hip.DeviceArray((setup.snapshots + 1) * seeds * self.cells)

    def at(self, field: str, t: int, s: int = 0) -> int:
        arr = self.U if field == "u" else self.V
        return arr.ptr + 4 * ((t * self.seeds + s) * self.cells)

    def snapshot(self, t: int):
        n = self.seeds * self.cells
        out = []
        for arr in (self.U, self.V):
            host = np.empty(n, np.float32)
            field_hip.check(field_hip.lib().fe_hip_memcpy_dtoh_async(host.ctypes.data, arr.ptr + 4 * t * n, 4 * n), "download")
            field_hip.sync()
            out.append(host)
        return out


def generate(setup: Setup, seeds, base_seed: int = 0, stats: dict | None = None) -> Dataset:
    """Fine runs of the law on the engine, one seed a plane, block-averaged into the dataset every `every` steps."""
    seeds = list(seeds)
    n = setup.fine
    law = field_hip.Law(DIFFREACT_LAW)
    solver = field_hip.Solver(law, (n, n, len(seeds)), setup.h_fine)
    for name, val in (("Du", setup.du), ("Dv", setup.dv), ("kr", setup.k)):
        solver.set_param(name, val)
    u0, v0 = initial(setup, seeds, base_seed)
    solver.set_field("u", upsample(u0, setup.factor).reshape(-1))
    solver.set_field("v", upsample(v0, setup.factor).reshape(-1))
    solver.set_lazy_readback(True)
    solver.check_contracts_every(1 << 30)
    data = Dataset(setup, len(seeds))

    def take(t):
        for f in ("u", "v"):
            field_hip.coarsen(solver.field_device(f), n, n, len(seeds), setup.factor, data.at(f, t))

    solver.refresh_derived()  # puts the start on the card
    take(0)
    field_hip.sync()
    t0 = time.perf_counter()
    for t in range(1, setup.snapshots + 1):
        for _ in range(setup.every):
            solver.step(setup.dt_fine)
        take(t)
    field_hip.sync()
    if stats is not None:
        stats["fine_steps"] = setup.snapshots * setup.every
        stats["fine_seconds"] = time.perf_counter() - t0
        stats["fine_cell_updates_per_s"] = stats["fine_steps"] * 2 * n * n * len(seeds) / stats["fine_seconds"]
    return data


def target_scale(data: Dataset, seeds: int = 4) -> float:
    """A flux scale for the net's outputs: h times the spread of what the coarse law misses over one snapshot."""
    from_host = data.snapshot(0), data.snapshot(1)
    s = data.setup
    c = s.coarse
    (u0, v0), (u1, v1) = from_host
    m = seeds * c * c
    u0, v0, u1, v1 = (a[:m].astype(np.float64).reshape(seeds, c, c) for a in (u0, v0, u1, v1))

    def lap(F):
        P = np.pad(F, ((0, 0), (1, 1), (1, 1)), mode="edge")
        return (P[:, 1:-1, :-2] + P[:, 1:-1, 2:] + P[:, :-2, 1:-1] + P[:, 2:, 1:-1] - 4 * F) / s.h_coarse ** 2

    tu = (u1 - u0) / s.dt_coarse - (s.du * lap(u0) + u0 - u0 ** 3 - s.k - v0)
    tv = (v1 - v0) / s.dt_coarse - (s.dv * lap(v0) + u0 - v0)
    return float(s.h_coarse * np.sqrt(0.5 * (np.mean(tu ** 2) + np.mean(tv ** 2))))


def initial_params(seed: int, last_scale: float = 0.0) -> np.ndarray:
    """The net's starting weights in the layout of include/fe_learned.h: per layer W (fan_in x fan_out, row-major) then
    b; Glorot-uniform W, zero b, the last layer times last_scale (0: the net starts at no correction)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out, scale in ((18, 64, 1.0), (64, 64, 1.0), (64, 2, last_scale)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        parts.append((rng.uniform(-lim, lim, (fan_in, fan_out)) * scale).reshape(-1))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def plain_losses(data: Dataset, seeds: slice, scale: float) -> np.ndarray:
    """The loss of the plain coarse law (the net at no correction) on each snapshot pair t -> t + 1 of `seeds`."""
    probe = field_hip.LearnedNet(initial_params(1, 0.0), scale)
    s = data.setup
    c = s.coarse
    count = seeds.stop - seeds.start
    out = np.empty(s.snapshots)
    for t in range(s.snapshots):
        out[t], _ = probe.gradient(data.at("u", t, seeds.start), data.at("v", t, seeds.start), data.at("u", t + 1, seeds.start),
                                   data.at("v", t + 1, seeds.start), c, c, count, s.h_coarse, s.du, s.dv, s.k, s.dt_coarse)
    return out


def train(net: field_hip.LearnedNet, data: Dataset, steps: int, batch: int, lr: float, train_seeds: int, seed: int = 0,
          log_every: int = 0, clip: float = 1e-2, first: int = 0, weights=None):
    """Adam on random (snapshot t >= first, block of `batch` seeds among the first `train_seeds`) pairs, pair t's loss
    times weights[t] when given; lr decays by 10 over the run. Returns the losses logged."""
    rng = np.random.default_rng(seed)
    s = data.setup
    c = s.coarse
    losses = []
    for i in range(steps):
        t = int(rng.integers(first, s.snapshots))
        s0 = int(rng.integers(0, train_seeds - batch + 1))
        want = bool(log_every) and (i % log_every == 0 or i == steps - 1)
        out = net.gradient(data.at("u", t, s0), data.at("v", t, s0), data.at("u", t + 1, s0), data.at("v", t + 1, s0), c, c, batch,
                           s.h_coarse, s.du, s.dv, s.k, s.dt_coarse, want=want,
                           weight=1.0 if weights is None else float(weights[t]))
        if want:
            losses.append((i, out[0]))
        net.adam(lr * 10 ** (-i / max(1, steps)), clip)
    field_hip.sync()
    return losses


def eval_loss(net: field_hip.LearnedNet, data: Dataset, seeds: slice, every: int = 5) -> float:
    """The training loss averaged over snapshot pairs t = 0, every, 2 every, ... of `seeds` (a fixed set: losses
    before and after training compare)."""
    s = data.setup
    c = s.coarse
    count = seeds.stop - seeds.start
    total, n = 0.0, 0
    for t in range(0, s.snapshots, every):
        loss, _ = net.gradient(data.at("u", t, seeds.start), data.at("v", t, seeds.start), data.at("u", t + 1, seeds.start),
                               data.at("v", t + 1, seeds.start), c, c, count, s.h_coarse, s.du, s.dv, s.k, s.dt_coarse)
        total += loss
        n += 1
    return total / n


def rollout(net, data: Dataset, seeds: slice, steps: int, learned: bool, times=None):
    """The coarse law (with the net when `learned`) from snapshot 0 of `seeds`, `steps` snapshots; returns the final
    u, v (host) and, when `times` is a list, appends the card time of the run (ms)."""
    s = data.setup
    c = s.coarse
    count = seeds.stop - seeds.start
    n = count * c * c
    bufs = [field_hip.DeviceArray(n) for _ in range(4)]
    for b, f in ((bufs[0], "u"), (bufs[1], "v")):
        field_hip.check(field_hip.lib().fe_hip_memcpy_dtod_async(b.ptr, data.at(f, 0, seeds.start), 4 * n), "copy")
    ev = field_hip.Events()
    ev.begin()
    cur, nxt = (bufs[0], bufs[1]), (bufs[2], bufs[3])
    for _ in range(steps):
        if learned:
            net.step(cur[0], cur[1], c, c, count, s.h_coarse, s.du, s.dv, s.k, s.dt_coarse, nxt[0], nxt[1], learned=True, bound=50.0)
        else:
            plain_step(cur[0], cur[1], c, c, count, s.h_coarse, s, s.dt_coarse, nxt[0], nxt[1])
        cur, nxt = nxt, cur
    ms = ev.end()
    if times is not None:
        times.append(ms)
    return cur[0].download(), cur[1].download()


def plain_step(u, v, nx, ny, count, h, s: Setup, dt, u2, v2):
    """The plain law's fp32 step (the same kernel the learned step falls back to)."""
    p = lambda a: a.ptr if isinstance(a, field_hip.DeviceArray) else a  # noqa: E731
    if field_hip.lib().fe_lnet_step(None, p(u), p(v), nx, ny, count, h, s.du, s.dv, s.k, dt, 0, 1e30, p(u2), p(v2)) != 0:
        raise field_hip.FieldHipError("plain step failed")


def plain_run(u0, v0, n: int, count: int, s: Setup, t_end: float, substeps: int, times=None):
    """The plain law on an n x n lattice (fp32 on the card) from host states, `substeps` steps to t_end."""
    h = s.width / n
    dt = t_end / substeps
    bufs = [field_hip.DeviceArray.from_host(np.ascontiguousarray(u0, np.float32)),
            field_hip.DeviceArray.from_host(np.ascontiguousarray(v0, np.float32)),
            field_hip.DeviceArray(u0.size), field_hip.DeviceArray(u0.size)]
    ev = field_hip.Events()
    ev.begin()
    cur, nxt = (bufs[0], bufs[1]), (bufs[2], bufs[3])
    for _ in range(substeps):
        plain_step(cur[0], cur[1], n, n, count, h, s, dt, nxt[0], nxt[1])
        cur, nxt = nxt, cur
    ms = ev.end()
    if times is not None:
        times.append(ms)
    return cur[0].download(), cur[1].download()
