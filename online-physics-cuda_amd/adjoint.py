"""Reverse mode over a law's run, with checkpoints and recomputation, all on the card.

The reference's `lawc --vjp` (tools/lawc/main.cpp:274-361) runs the primal forward keeping every state on the host,
then the adjoint law (ir::adjoint_law) backward over them, uploading one state a step. Here the primal keeps only a
checkpoint every `every` steps on the card; the backward pass takes the segments last to first: it restores the
segment's checkpoint, recomputes the segment's states into a ring on the card, then steps the adjoint law over them
in reverse, the adjoint law's static primal fields bound to each state where it lies in the ring (no copy). Memory:
(steps / every + every) states instead of `steps`; work: one extra forward pass.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

import field_hip


@dataclass
class Result:
    loss: float
    field_grads: dict            # dL/d(initial field), host arrays
    param_grads: dict            # dL/dp
    forward_ms: float = 0.0      # card time of the forward pass
    backward_ms: float = 0.0     # card time of the backward pass (recomputation and adjoint steps)
    stats: dict = field(default_factory=dict)


def vjp(law_text: str, shape, h: float, params: dict, init: dict, dt: float, steps: int, loss_field: str, every: int = 8,
        term: str = "") -> Result:
    """dL/d(initial fields) and dL/d(params) for L = sum over the lattice of loss_field^2 after `steps` steps of dt."""
    law = field_hip.Law(law_text)
    adj_text = law.adjoint(term)
    adj_law = field_hip.Law(adj_text)
    primal = field_hip.Solver(law, shape, h)
    adj = field_hip.Solver(adj_law, shape, h)
    for name, val in params.items():
        primal.set_param(name, val)
        adj.set_param(name, val)
    for name, val in init.items():
        primal.set_field(name, np.asarray(val, np.float32).reshape(-1))
    for s in (primal, adj):
        s.set_lazy_readback(True)
        s.check_contracts_every(1 << 30)
    primal.refresh_derived()
    adj.refresh_derived()
    names = law.field_names()
    inputs = [n for n in names if n in adj_law.field_names()]  # primal fields the adjoint reads (its statics)
    cells = int(np.prod(shape))
    n_ckpt = (steps + every - 1) // every
    ckpt = [[field_hip.DeviceArray(cells) for _ in inputs] for _ in range(n_ckpt)]
    ring = [[field_hip.DeviceArray(cells) for _ in inputs] for _ in range(every)]

    def save(dst):
        for arr, name in zip(dst, inputs):
            field_hip.check(field_hip.lib().fe_hip_memcpy_dtod_async(arr.ptr, primal.field_device(name), 4 * cells), "save")

    ev = field_hip.Events()
    ev.begin()
    for s in range(steps):
        if s % every == 0:
            save(ckpt[s // every])
        primal.step(dt)
    forward_ms = ev.end()
    if primal.last_substeps() and max(primal.last_substeps()) != 1:
        raise field_hip.FieldHipError("the differentiated term took substeps: give a dt within its max_dt")
    final = primal.get_field(loss_field).astype(np.float64)
    loss = float(np.sum(final * final))
    adj.set_field(loss_field + "__adj", (2.0 * final).astype(np.float32))
    adj.refresh_derived()  # uploads the seed now: from here on only device copies write the adjoint solver's fields
    # A param's adjoint field takes one step's contribution at a time: its total (fixed-order, compensated, in double
    # on the card) is added to a double on the card, step after step, and the field cleared. Left to accumulate in
    # fp32 over the steps, the cells' sums would round at every step, and the lattice's total of them cancels heavily.
    pnames = [p for p in params if p + "__adj" in adj_law.field_names()]
    pacc = {p: field_hip.DeviceArray.from_host(np.zeros(4, np.float32)) for p in pnames}  # two doubles, added on the card
    ev.begin()
    for seg in range(n_ckpt - 1, -1, -1):
        first = seg * every
        count = min(every, steps - first)
        for arr, name in zip(ckpt[seg], inputs):
            primal.set_field_device(name, arr)
        for j in range(count):
            save(ring[j])
            if j + 1 < count:  # the segment's last state is never needed by its own adjoint steps
                primal.step(dt)
        for j in range(count - 1, -1, -1):
            for arr, name in zip(ring[j], inputs):
                adj.bind_field_device(name, arr)  # the adjoint law reads the state where it lies: no copy
            adj.step(dt)
            for p in pnames:
                adj.total_accumulate(p + "__adj", pacc[p])
                adj.fill_field(p + "__adj", 0.0)
    backward_ms = ev.end()
    for name in inputs:
        adj.bind_field_device(name, None)  # the ring goes with this call
    grads = {n: adj.get_field(n + "__adj").astype(np.float64) for n in inputs}
    pgrads = {p: float(pacc[p].download().view(np.float64)[0]) for p in pnames}
    return Result(loss, grads, pgrads, forward_ms, backward_ms,
                  {"checkpoints": n_ckpt, "every": every, "states_held": n_ckpt + every, "cells": cells, "steps": steps})
