"""Shared test helpers: product <-> oracle conversions and small tree builders."""
from __future__ import annotations

import numpy as np

from oracle.tree import Topology
from pytreenet_amd.random_states import complete_tree, random_state_tensors


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
