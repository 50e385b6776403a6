"""Reference-order flop counts of the oracle's contractions (test/baseline infrastructure).

The oracle's contraction functions are executed on shape-only stand-ins whose
``np.tensordot`` records 8 * M * N * K real flops per call (a complex
multiply-add counted as 8 flops), so the reference order's cost is counted
exactly without allocating its intermediates (67 GB at the chi=256 ternary root).
"""
from __future__ import annotations

import math
from contextlib import contextmanager
from typing import Dict, Tuple

import numpy as np

from . import contract as C


class ShapeOnly:
    """Minimal array stand-in carrying only a shape."""

    def __init__(self, shape):
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)

    def conj(self):
        return self

    @property
    def size(self):
        return math.prod(self.shape)


class _Counter:
    def __init__(self):
        self.flops = 0.0
        self.peak = 0


@contextmanager
def counting():
    """Patch the oracle's tensordot with a shape-only flop counter."""
    ctr = _Counter()
    real = C.np.tensordot

    def fake(a, b, axes):
        ax_a, ax_b = axes
        ax_a = [ax_a] if isinstance(ax_a, int) else list(ax_a)
        ax_b = [ax_b] if isinstance(ax_b, int) else list(ax_b)
        k = math.prod(a.shape[i] for i in ax_a)
        m = math.prod(s for i, s in enumerate(a.shape) if i not in ax_a)
        n = math.prod(s for i, s in enumerate(b.shape) if i not in ax_b)
        ctr.flops += 8.0 * m * n * k
        ctr.peak = max(ctr.peak, m * n)
        return ShapeOnly([s for i, s in enumerate(a.shape) if i not in ax_a] +
                         [s for i, s in enumerate(b.shape) if i not in ax_b])

    C.np.tensordot = fake
    try:
        yield ctr
    finally:
        C.np.tensordot = real


def heff_flops(topo, nid: str, shapes: Dict[str, tuple], wshapes: Dict[str, tuple]) -> float:
    """Reference-order flops of one H_eff apply at ``nid``."""
    envs = {}
    for nb in topo.neighbours(nid):
        d = shapes[nid][topo.neighbour_index(nid, nb)]
        D = wshapes[nid][topo.neighbour_index(nid, nb)]
        envs[(nb, nid)] = ShapeOnly((d, D, d))
    with counting() as c:
        C.heff_apply(topo, nid, ShapeOnly(shapes[nid]), ShapeOnly(wshapes[nid]), envs)
    return c.flops


def env_flops(topo, nid: str, to: str, shapes, wshapes) -> float:
    """Reference-order flops of one environment update ``nid -> to``."""
    envs = {}
    for nb in topo.neighbours(nid):
        if nb == to:
            continue
        d = shapes[nid][topo.neighbour_index(nid, nb)]
        D = wshapes[nid][topo.neighbour_index(nid, nb)]
        envs[(nb, nid)] = ShapeOnly((d, D, d))
    with counting() as c:
        C.env_update(topo, nid, to, ShapeOnly(shapes[nid]), ShapeOnly(wshapes[nid]), envs)
    return c.flops


def bond_flops(chi: int, D: int) -> float:
    """Reference-order flops of one bond action on a chi x chi link."""
    return 8.0 * chi * (D * chi) * chi + 8.0 * chi * (D * chi) * (D * chi)
