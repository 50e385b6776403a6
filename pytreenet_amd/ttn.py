"""Tree tensor network containers: Node, TreeTensorNetwork, TTNS.

Mirrors the reference data model (``core/node.py``, ``core/graph_node.py``,
``core/tree_structure.py``, ``core/ttn.py``, ``ttns/ttns.py``):

* leg order of every tensor is (parent, children..., open legs) -- the
  reference convention of ``core/node.py:35``;
* neighbour order is parent first, then children (``graph_node.py:340-355``);
* ``add_child_to_parent(child, tensor, child_leg, parent_id, parent_leg)``
  makes ``child_leg`` the child's parent leg and appends ``parent_leg`` to the
  parent's children (``core/ttn.py`` / ``tree_structure.py:154-170``).

The reference stores a lazy leg permutation per node and materialises it on
access (``core/ttn.py:72-115``); here every tensor is kept materialised in the
canonical order on the ROCm device, and the tree's children order never
changes under the gauge moves of the time evolution (the reference may
reorder children after split/merge, which changes only leg order, not any
gauge-invariant quantity).  Node metadata lives on the host.
"""
from __future__ import annotations

import json
import os
import uuid
from typing import Dict, Iterator, List, Optional, Tuple, Union

import numpy as np
import torch

from . import engine as E
from ._lib import PtnBackendError
from .splitting import SplitMode


def default_device() -> torch.device:
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def as_device_tensor(t, device=None, dtype=torch.complex128) -> torch.Tensor:
    if isinstance(t, torch.Tensor):
        out = t.to(device=device or t.device, dtype=dtype)
    else:
        out = torch.as_tensor(np.asarray(t), dtype=dtype)
        out = out.to(device or default_device())
    return out.contiguous()


class Node:
    """Host-side metadata of one tensor node (identifier, parent, children, shape)."""

    def __init__(self, tensor=None, identifier: Optional[str] = None):
        self.identifier = identifier if identifier is not None else str(uuid.uuid1())
        self.parent: Optional[str] = None
        self.children: List[str] = []
        self.shape: Tuple[int, ...] = tuple(tensor.shape) if tensor is not None else ()

    # -- connectivity --------------------------------------------------------------
    def is_root(self) -> bool:
        return self.parent is None

    def is_leaf(self) -> bool:
        return len(self.children) == 0

    def nparents(self) -> int:
        return 0 if self.parent is None else 1

    def nchildren(self) -> int:
        return len(self.children)

    def nneighbours(self) -> int:
        return self.nparents() + self.nchildren()

    def neighbouring_nodes(self) -> List[str]:
        return ([] if self.parent is None else [self.parent]) + list(self.children)

    def neighbour_index(self, node_id: str) -> int:
        if node_id == self.parent:
            return 0
        if node_id in self.children:
            return self.children.index(node_id) + self.nparents()
        from .util import NoConnectionException
        raise NoConnectionException(f"{node_id} is not a neighbour of {self.identifier}!")

    def is_parent_of(self, other: str) -> bool:
        return other in self.children

    def is_child_of(self, other: str) -> bool:
        return self.parent == other

    def has_x_children(self, x: int) -> bool:
        return len(self.children) == x

    # -- legs ----------------------------------------------------------------------
    def nvirt_legs(self) -> int:
        return self.nneighbours()

    def nlegs(self) -> int:
        return len(self.shape)

    @property
    def open_legs(self) -> List[int]:
        return list(range(self.nvirt_legs(), self.nlegs()))

    def nopen_legs(self) -> int:
        return self.nlegs() - self.nvirt_legs()

    def open_dimension(self) -> int:
        return int(np.prod([self.shape[i] for i in self.open_legs])) if self.open_legs else 1

    def virtual_dimension(self) -> int:
        return int(np.prod(self.shape[: self.nvirt_legs()])) if self.nvirt_legs() else 1

    def neighbour_dim(self, node_id: str) -> int:
        return self.shape[self.neighbour_index(node_id)]

    def __repr__(self):
        return (f"Node({self.identifier!r}, parent={self.parent!r}, "
                f"children={self.children!r}, shape={self.shape})")


class TensorDict(dict):
    """id -> device tensor; assignment refreshes the owning node's shape."""

    def __init__(self, owner: "TreeTensorNetwork"):
        super().__init__()
        self._owner = owner

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        node = self._owner._nodes.get(key)
        if node is not None:
            node.shape = tuple(value.shape)


class TreeTensorNetwork:
    """A tree of tensors with host metadata and device data."""

    def __init__(self):
        self._nodes: Dict[str, Node] = {}
        self._tensors = TensorDict(self)
        self._root_id: Optional[str] = None
        self.orthogonality_center_id: Optional[str] = None

    # -- access ---------------------------------------------------------------------
    @property
    def nodes(self) -> Dict[str, Node]:
        return self._nodes

    @property
    def tensors(self) -> TensorDict:
        return self._tensors

    @property
    def root_id(self) -> Optional[str]:
        return self._root_id

    @property
    def root(self) -> Tuple[Node, torch.Tensor]:
        return self[self._root_id]

    def __getitem__(self, key: str) -> Tuple[Node, torch.Tensor]:
        return self._nodes[key], self._tensors[key]

    def __contains__(self, key: str) -> bool:
        return key in self._nodes

    def __len__(self) -> int:
        return len(self._nodes)

    def num_nodes(self) -> int:
        return len(self._nodes)

    @property
    def device(self) -> torch.device:
        return next(iter(self._tensors.values())).device

    # -- construction -------------------------------------------------------------
    def add_root(self, node: Node, tensor) -> "TreeTensorNetwork":
        if self._root_id is not None:
            raise ValueError("root already exists")
        t = as_device_tensor(tensor)
        node.shape = tuple(t.shape)
        self._nodes[node.identifier] = node
        self._tensors[node.identifier] = t
        self._root_id = node.identifier
        return self

    def add_child_to_parent(self, child: Node, tensor, child_leg: int, parent_id: str,
                            parent_leg: int) -> "TreeTensorNetwork":
        """Attach ``child``: its leg ``child_leg`` meets leg ``parent_leg`` of the parent.

        ``parent_leg`` indexes the parent's current tensor, which must be an open leg.
        """
        if child.identifier in self._nodes:
            raise ValueError(f"node {child.identifier} already exists")
        parent = self._nodes[parent_id]
        if parent_leg not in parent.open_legs:
            raise ValueError(f"leg {parent_leg} of {parent_id} is not open")
        t = as_device_tensor(tensor, device=self._tensors[parent_id].device)
        perm = [child_leg] + [i for i in range(t.dim()) if i != child_leg]
        t = t.permute(perm).contiguous()
        pt = self._tensors[parent_id]
        if pt.shape[parent_leg] != t.shape[0]:
            raise ValueError("bond dimensions do not match")
        nv = parent.nvirt_legs()
        pperm = list(range(nv)) + [parent_leg] + [i for i in parent.open_legs if i != parent_leg]
        parent.children.append(child.identifier)
        self._tensors[parent_id] = pt.permute(pperm).contiguous()
        child.parent = parent_id
        child.children = []
        child.shape = tuple(t.shape)
        self._nodes[child.identifier] = child
        self._tensors[child.identifier] = t
        return self

    def to(self, device) -> "TreeTensorNetwork":
        for k in list(self._tensors):
            self._tensors[k] = self._tensors[k].to(device)
        return self

    def copy(self) -> "TreeTensorNetwork":
        """Deep copy: metadata copied, device tensors cloned."""
        new = self.__class__.__new__(self.__class__)
        TreeTensorNetwork.__init__(new)
        for k, n in self._nodes.items():
            m = Node(identifier=n.identifier)
            m.parent, m.children, m.shape = n.parent, list(n.children), n.shape
            new._nodes[k] = m
        for k, t in self._tensors.items():
            dict.__setitem__(new._tensors, k, t.clone())
        new._root_id = self._root_id
        new.orthogonality_center_id = self.orthogonality_center_id
        for attr, val in self.__dict__.items():
            if attr not in new.__dict__:
                new.__dict__[attr] = val
        return new

    __deepcopy__ = lambda self, memo: self.copy()

    # -- interchange files (core/ttn.py:1748-1837) ----------------------------------
    def save(self, filepath: str):
        """Write ``{filepath}.npz`` (every tensor under its node id, in the canonical
        leg order (parent, children..., open)) and ``{filepath}.json`` (root id,
        orthogonality centre, and per node its identifier, parent, children and
        shape).  These are the reference's two files: its ``TreeTensorNetwork.load``
        reads them, and ``load`` below reads the reference's."""
        arrays = {k: t.detach().cpu().numpy() for k, t in self._tensors.items()}
        np.savez(f"{filepath}.npz", **arrays)
        data = {"root_id": self._root_id,
                "orthogonality_center_id": self.orthogonality_center_id,
                "nodes": {nid: {"identifier": n.identifier, "parent": n.parent,
                                "children": list(n.children),
                                "shape": [int(d) for d in self._tensors[nid].shape]}
                          for nid, n in self._nodes.items()}}
        with open(f"{filepath}.json", "w") as f:
            json.dump(data, f, indent=2)

    @classmethod
    def load(cls, filepath: str, device=None) -> "TreeTensorNetwork":
        """Read the two files ``save`` writes (or the reference's).  Tensors go to
        ``device`` (default: the ROCm device when present) as complex128; every
        bond is checked to have the same dimension on both sides."""
        if not (os.path.exists(f"{filepath}.npz") and os.path.exists(f"{filepath}.json")):
            raise FileNotFoundError(f"Could not find required files at {filepath}!")
        with open(f"{filepath}.json") as f:
            data = json.load(f)
        ttn = cls()
        with np.load(f"{filepath}.npz") as arrays:
            for nid, nd in data["nodes"].items():
                node = Node(identifier=nd["identifier"])
                node.parent = nd["parent"]
                node.children = list(nd["children"])
                t = as_device_tensor(arrays[nid], device=device)
                node.shape = tuple(t.shape)
                ttn._nodes[nid] = node
                dict.__setitem__(ttn._tensors, nid, t)
        ttn._root_id = data["root_id"]
        ttn.orthogonality_center_id = data["orthogonality_center_id"]
        for nid, node in ttn._nodes.items():
            for c in node.children:
                child = ttn._nodes[c]
                if child.parent != nid:
                    raise ValueError(f"{c} is listed as a child of {nid} but names "
                                     f"{child.parent} as its parent")
                if node.shape[node.neighbour_index(c)] != child.shape[0]:
                    raise ValueError(f"bond {nid}-{c} has dimension "
                                     f"{node.shape[node.neighbour_index(c)]} at {nid} and "
                                     f"{child.shape[0]} at {c}")
        return ttn

    # -- tree queries (core/tree_structure.py) ----------------------------------------
    def neighbours(self, nid: str) -> List[str]:
        return self._nodes[nid].neighbouring_nodes()

    def get_leaves(self, include_root: bool = True) -> List[str]:
        leaves = [k for k, n in self._nodes.items() if n.is_leaf()]
        if include_root and self._nodes[self._root_id].nneighbours() == 1:
            leaves.append(self._root_id)
        return leaves

    def distance_to_node(self, center: str) -> Dict[str, int]:
        out = {center: 0}
        stack = [(nb, center, 1) for nb in reversed(self.neighbours(center))]
        while stack:
            nid, last, d = stack.pop()
            out[nid] = d
            for nb in reversed(self.neighbours(nid)):
                if nb != last:
                    stack.append((nb, nid, d + 1))
        return out

    def linearise(self) -> List[str]:
        order: List[str] = []

        def rec(nid):
            for c in self._nodes[nid].children:
                rec(c)
            order.append(nid)

        rec(self._root_id)
        return order

    def find_path_to_root(self, nid: str) -> List[str]:
        path = [nid]
        while self._nodes[nid].parent is not None:
            nid = self._nodes[nid].parent
            path.append(nid)
        return path

    def path_from_to(self, start: str, end: str) -> List[str]:
        if start == end:
            return [start]
        up = self.find_path_to_root(start)
        down = self.find_path_to_root(end)
        down_set = set(down)
        lca = next(n for n in up if n in down_set)
        return up[: up.index(lca) + 1] + list(reversed(down[: down.index(lca)]))

    def is_child_of(self, a: str, b: str) -> bool:
        return self._nodes[a].parent == b

    def is_parent_of(self, a: str, b: str) -> bool:
        return self._nodes[b].parent == a

    # -- gauge moves (core/canonical_form.py, core/ttn.py:1530-1582) -----------------
    def _qr_towards(self, nid: str, nb: str, mode=None):
        from .splitting import qr_split_to_neighbour, absorb_into_neighbour
        q, r = qr_split_to_neighbour(self, nid, nb, mode=mode)
        self._tensors[nid] = q
        absorb_into_neighbour(self, nb, nid, r)

    def canonical_form(self, orthogonality_center_id: str, mode=SplitMode.REDUCED):
        """QR-canonicalise towards ``orthogonality_center_id``
        (``core/canonical_form.py``; ``core/ttn.py:1673-1697``: a state that already
        has a centre only moves it there).  ``mode`` is the QR mode of every
        split: REDUCED (the default, as in the reference) lets a bond shrink to
        the product of the node's other dimensions, KEEP keeps every bond."""
        if self.orthogonality_center_id is not None:
            self.move_orthogonalization_center(orthogonality_center_id, mode=mode)
            return
        dist = self.distance_to_node(orthogonality_center_id)
        for d in range(max(dist.values()), 0, -1):
            for nid in [n for n in dist if dist[n] == d]:
                nb = min(self.neighbours(nid), key=lambda x: dist[x])
                self._qr_towards(nid, nb, mode)
        self.orthogonality_center_id = orthogonality_center_id

    def move_orthogonalization_center(self, new_center_id: str, mode=SplitMode.REDUCED):
        """QR splits along the path from the current centre (``core/ttn.py:1530-1582``);
        ``mode`` as in ``canonical_form``."""
        if self.orthogonality_center_id is None:
            raise AssertionError("The TTN is not in canonical form!")
        path = self.path_from_to(self.orthogonality_center_id, new_center_id)
        for a, b in zip(path[:-1], path[1:]):
            self._qr_towards(a, b, mode)
        self.orthogonality_center_id = new_center_id

    def ensure_root_orth_center(self, mode=SplitMode.REDUCED) -> bool:
        """Make the root the orthogonality centre (``core/ttn.py:1631-1645``);
        True when it already was."""
        if self.orthogonality_center_id == self._root_id:
            return True
        if self.orthogonality_center_id is None:
            self.canonical_form(self._root_id, mode=mode)
        else:
            self.move_orthogonalization_center(self._root_id, mode=mode)
        return False

    def bond_dims(self) -> Dict[Tuple[str, str], int]:
        out = {}
        for nid, n in self._nodes.items():
            if n.parent is not None:
                out[(n.parent, nid)] = n.shape[0]
        return out

    def max_bond_dim(self) -> int:
        return max(self.bond_dims().values(), default=1)


class TreeTensorNetworkState(TreeTensorNetwork):
    """TTNS: state tensors with one physical leg per node (``ttns/ttns.py``)."""

    def scalar_product(self, other: Optional["TreeTensorNetworkState"] = None,
                       use_orthogonal_center: bool = True) -> complex:
        """sum self * conj(other) as the reference computes it (``ttns.py:65-93``);
        with ``other`` None the squared norm."""
        from .contractions import scalar_product
        if other is None:
            c = self.orthogonality_center_id
            if use_orthogonal_center and c is not None:
                t = self._tensors[c]
                return complex(_dotc(t, t))
            other = self
        return scalar_product(self, other)

    def norm(self) -> float:
        return float(abs(self.scalar_product()) ** 0.5)

    def normalise(self) -> "TreeTensorNetworkState":
        c = self.orthogonality_center_id
        if c is None:
            raise AssertionError("normalise needs an orthogonality centre")
        nrm = self.norm()
        self._tensors[c] = self._tensors[c] / nrm
        return self

    def operator_expectation_value(self, operator) -> complex:
        """<psi|O|psi> for a TTNO or a TensorProduct (unnormalised, ``ttns.py:278-297``)."""
        from .ttno import TreeTensorNetworkOperator, TensorProduct
        from .contractions import expectation_value, tensor_product_expectation_value
        if isinstance(operator, TreeTensorNetworkOperator):
            return expectation_value(self, operator)
        if isinstance(operator, TensorProduct):
            return tensor_product_expectation_value(self, operator)
        raise TypeError(f"unsupported operator type {type(operator)}")

    def single_site_operator_expectation_value(self, node_id: str, operator) -> complex:
        from .ttno import TensorProduct
        return self.operator_expectation_value(TensorProduct({node_id: operator}))


def _dotc(a: torch.Tensor, b: torch.Tensor) -> complex:
    from .krylov import dotc
    return dotc(a, b)


TTNS = TreeTensorNetworkState
