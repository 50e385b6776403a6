"""Brute-force dense vectors and operators for pinning (oracle, test infrastructure only).

Site order of every dense object is ``topo.node_ids`` (insertion order); the
first site is the most significant index.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np

from .tree import Topology


def _sites_below(topo: Topology, nid: str) -> List[str]:
    out = [nid]
    for c in topo.children[nid]:
        out += _sites_below(topo, c)
    return out


def state_to_vector(topo: Topology, A: Dict[str, np.ndarray]) -> np.ndarray:
    """Contract a TTNS (legs parent, children..., phys) into a dense vector."""
    def rec(nid):
        # returns tensor with legs (parent?, phys of all sites below in _sites_below order)
        t = A[nid]
        has_p = topo.parent[nid] is not None
        nchild = len(topo.children[nid])
        sites = [nid]
        # t legs: (p?, c_1..c_k, phys)
        for i, c in enumerate(topo.children[nid]):
            sub, subsites = rec(c)
            leg = (1 if has_p else 0)  # after each contraction the next child leg moves here
            t = np.tensordot(t, sub, axes=([leg], [0]))
            sites += subsites
        # legs now: (p?, phys_own, phys of child subtrees...)
        return t, sites

    t, sites = rec(topo.root)
    order = [sites.index(s) for s in topo.node_ids]
    return np.transpose(t, order).reshape(-1)


def kron_op(n_sites: int, ops: Dict[int, np.ndarray], d: int = 2) -> np.ndarray:
    mats = [ops.get(i, np.eye(d)) for i in range(n_sites)]
    out = mats[0]
    for m in mats[1:]:
        out = np.kron(out, m)
    return out


def edge_hamiltonian(topo: Topology, site_terms: Dict[str, np.ndarray],
                     edge_terms: List[tuple]) -> np.ndarray:
    """Dense sum_i h_i + sum coef A_{n1} B_{n2}."""
    ids = topo.node_ids
    pos = {n: i for i, n in enumerate(ids)}
    n = len(ids)
    H = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for nid, h in site_terms.items():
        H += kron_op(n, {pos[nid]: h})
    for n1, n2, coef, a, b in edge_terms:
        H += coef * kron_op(n, {pos[n1]: a, pos[n2]: b})
    return H


def ttno_to_matrix(topo: Topology, W: Dict[str, np.ndarray]) -> np.ndarray:
    """Contract a TTNO (legs parent, children..., out, in) into a dense matrix."""
    def rec(nid):
        t = W[nid]
        has_p = topo.parent[nid] is not None
        sites = [nid]
        for c in topo.children[nid]:
            sub, subsites = rec(c)
            leg = 1 if has_p else 0
            t = np.tensordot(t, sub, axes=([leg], [0]))
            sites += subsites
        # legs: (p?, out_own, in_own, [out_s, in_s for each site below, grouped per child])
        return t, sites

    t, sites = rec(topo.root)
    # rec leaves, for each contracted child subtree, its legs in the order the child
    # produced them; normalise by tracking leg roles explicitly.
    roles = _roles(topo, topo.root)
    n = len(sites)
    out_axes = [roles.index(("out", s)) for s in topo.node_ids]
    in_axes = [roles.index(("in", s)) for s in topo.node_ids]
    t = np.transpose(t, out_axes + in_axes)
    dim = int(np.prod(t.shape[:n]))
    return t.reshape(dim, dim)


def _roles(topo: Topology, nid: str):
    roles = [("out", nid), ("in", nid)]
    for c in topo.children[nid]:
        roles += _roles(topo, c)
    return roles
