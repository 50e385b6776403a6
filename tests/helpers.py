"""Shared test helpers: product <-> oracle conversions and small tree builders."""
from __future__ import annotations

import numpy as np

from oracle.tree import Topology
from oracle import dense as OD
from oracle import bug as OB
from oracle.tdvp import OracleState
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import build_edge_ttno_tensors


def topology_from_spec(ids, parent, children) -> Topology:
    t = Topology()
    for nid in ids:
        if parent[nid] is None:
            t.add_root(nid)
        else:
            t.add_child(nid, parent[nid])
    return t


def topology_from_ttn(ttn) -> Topology:
    """Oracle topology with the product's node insertion order and children order."""
    t = Topology()
    for nid, node in ttn.nodes.items():
        if node.parent is None:
            t.add_root(nid)
        else:
            t.add_child(nid, node.parent)
    for nid, node in ttn.nodes.items():
        t.children[nid] = list(node.children)
    return t


def host_tensors(ttn) -> dict:
    return {k: v.detach().cpu().numpy() for k, v in ttn.tensors.items()}


def small_tree(branching=2, depth=2, chi=4, seed=7):
    ids, parent, children = complete_tree(branching, depth)
    tensors = random_state_tensors(ids, parent, children, chi, seed)
    return ids, parent, children, tensors


X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


def _xyz_terms(topo, seed):
    rng = np.random.default_rng(seed)
    site = {n: rng.uniform(-1, 1) * Z for n in topo.node_ids}
    edges = []
    for n in topo.node_ids:
        p = topo.parent[n]
        if p is not None:
            j = rng.uniform(-1, 1, 3)
            edges += [(p, n, -j[0], X, X), (p, n, -j[1], Y, Y), (p, n, -j[2], Z, Z)]
    return site, edges


def oracle_problem(branching, depth, chi, seed):
    """(topology, normalised oracle state with the root as centre, product-built TTNO
    tensors, dense Hamiltonian) of a random XYZ model with random fields on a
    complete tree."""
    ids, parent, children = complete_tree(branching, depth)
    tensors = random_state_tensors(ids, parent, children, chi, seed)
    topo = topology_from_spec(ids, parent, children)
    site, edges = _xyz_terms(topo, seed)
    st = build_state(ids, parent, children, tensors, device="cpu")
    W = build_edge_ttno_tensors(st, site, edges)
    H = OD.edge_hamiltonian(topo, site, edges)
    ost = OracleState(topo, tensors)
    ost.canonical_form(topo.root)
    ost.normalise()
    return topo, ost, W, H


def reversed_children_operator(topo: Topology, W: dict):
    """(operator tree, tensors) of the same TTNO with every node's children in reverse
    order: the operator's own tree then differs from the state's wherever a node
    has two or more children, and each tensor's child legs are permuted to match."""
    op = Topology()
    for nid in topo.node_ids:
        if topo.parent[nid] is None:
            op.add_root(nid)
        else:
            op.add_child(nid, topo.parent[nid])
    Wp = {}
    for nid in topo.node_ids:
        kids = list(topo.children[nid])
        op.children[nid] = list(reversed(kids))
        npar = 0 if topo.parent[nid] is None else 1
        t = W[nid]
        perm = list(range(npar)) + [npar + kids.index(c) for c in op.children[nid]] + \
            [t.ndim - 2, t.ndim - 1]
        Wp[nid] = np.ascontiguousarray(np.transpose(t, perm))
    return op, Wp
