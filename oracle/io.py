"""Reader of the reference's TTN interchange files (oracle, test infrastructure only).

Restates ``TreeTensorNetwork.load`` (``core/ttn.py:1786-1835``): ``{path}.json``
holds the root, the orthogonality centre and per node its identifier, parent,
children (in order) and shape; ``{path}.npz`` holds every tensor under its node
id with legs (parent, children..., open) in that order.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional, Tuple

import numpy as np

from .tree import Topology


def load_ttn(path: str) -> Tuple[Topology, Dict[str, np.ndarray], Optional[str]]:
    """(topology with the files' children order, tensors, orthogonality centre)."""
    if not (os.path.exists(f"{path}.npz") and os.path.exists(f"{path}.json")):
        raise FileNotFoundError(f"Could not find required files at {path}!")
    with open(f"{path}.json") as f:
        data = json.load(f)
    topo = Topology()
    for nid, nd in data["nodes"].items():
        topo.parent[nid] = nd["parent"]
        topo.children[nid] = list(nd["children"])
    topo.root = data["root_id"]
    with np.load(f"{path}.npz") as arrays:
        tensors = {nid: np.array(arrays[nid]) for nid in data["nodes"]}
    for nid, t in tensors.items():
        if tuple(t.shape) != tuple(data["nodes"][nid]["shape"]):
            raise ValueError(f"tensor {nid} has shape {t.shape}, the structure says "
                             f"{data['nodes'][nid]['shape']}")
    return topo, tensors, data["orthogonality_center_id"]


def save_ttn(path: str, topo: Topology, tensors: Dict[str, np.ndarray],
             center: Optional[str] = None):
    """Write the two files in the reference's format (``core/ttn.py:1748-1784``)."""
    np.savez(f"{path}.npz", **tensors)
    data = {"root_id": topo.root, "orthogonality_center_id": center,
            "nodes": {nid: {"identifier": nid, "parent": topo.parent[nid],
                            "children": list(topo.children[nid]),
                            "shape": [int(d) for d in tensors[nid].shape]}
                      for nid in topo.node_ids}}
    with open(f"{path}.json", "w") as f:
        json.dump(data, f, indent=2)
