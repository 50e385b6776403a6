"""Synthetic trees and seeded random states for the benchmark workloads.

* complete b-ary trees of a given depth, node ids ``site{i}`` in breadth-first
  (parents-first) order, every node carrying one physical leg;
* bond dims min(chi, d**|subtree|, d**(N - |subtree|)) -- larger values would
  only be zero-padded by a KEEP canonicalisation;
* tensors drawn from one ``numpy.random.default_rng(seed)`` in parents-first
  order with real and imaginary parts ~ N(0, 1) (the reference ``crandn``,
  ``util/crandn.py:36-67``, additionally divides by sqrt(2); a global scale is
  removed by the normalisation anyway).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

from .ttn import Node, TreeTensorNetworkState


def complete_tree(branching: int, depth: int) -> Tuple[List[str], Dict[str, Optional[str]],
                                                         Dict[str, List[str]]]:
    """(ids in BFS order, parent map, children map) of a complete tree; depth 0 = root only."""
    ids = ["site0"]
    parent: Dict[str, Optional[str]] = {"site0": None}
    children: Dict[str, List[str]] = {"site0": []}
    frontier = ["site0"]
    for _ in range(depth):
        nxt = []
        for p in frontier:
            for _ in range(branching):
                nid = f"site{len(ids)}"
                ids.append(nid)
                parent[nid] = p
                children[nid] = []
                children[p].append(nid)
                nxt.append(nid)
        frontier = nxt
    return ids, parent, children


def ternary_tree_40():
    """Root with three children, each with three children, each with three leaves (N = 40)."""
    return complete_tree(3, 3)


def _subtree_sizes(ids, children) -> Dict[str, int]:
    size: Dict[str, int] = {}
    for nid in reversed(ids):
        size[nid] = 1 + sum(size[c] for c in children[nid])
    return size


def bond_dims(ids, parent, children, chi: int, phys_dim: int = 2) -> Dict[str, int]:
    """Dimension of the bond above each non-root node."""
    n = len(ids)
    size = _subtree_sizes(ids, children)
    out = {}
    for nid in ids:
        if parent[nid] is None:
            continue
        s = size[nid]
        cap = min(float(phys_dim) ** s, float(phys_dim) ** (n - s))
        out[nid] = int(min(chi, cap))
    return out


def random_state_tensors(ids, parent, children, chi: int, seed: int, phys_dim: int = 2,
                         bonds: Optional[Dict[str, int]] = None) -> Dict[str, np.ndarray]:
    bonds = bonds or bond_dims(ids, parent, children, chi, phys_dim)
    rng = np.random.default_rng(seed)
    out = {}
    for nid in ids:
        shape = ([bonds[nid]] if parent[nid] is not None else []) + \
                [bonds[c] for c in children[nid]] + [phys_dim]
        out[nid] = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return out


def build_state(ids, parent, children, tensors: Dict[str, np.ndarray],
                device=None) -> TreeTensorNetworkState:
    """TTNS from canonical-layout tensors (parent, children..., phys)."""
    st = TreeTensorNetworkState()
    for nid in ids:
        t = tensors[nid]
        if device is not None:
            import torch
            t = torch.as_tensor(t, dtype=torch.complex128).to(device)
        if parent[nid] is None:
            st.add_root(Node(identifier=nid), t)
        else:
            p = parent[nid]
            pleg = st.nodes[p].nvirt_legs()
            st.add_child_to_parent(Node(identifier=nid), t, 0, p, pleg)
    return st


def random_tree_state(branching: int, depth: int, chi: int, seed: int, phys_dim: int = 2,
                      device=None) -> TreeTensorNetworkState:
    ids, parent, children = complete_tree(branching, depth)
    tensors = random_state_tensors(ids, parent, children, chi, seed, phys_dim)
    return build_state(ids, parent, children, tensors, device=device)
