"""Dense effective Hamiltonians (site, bond and two-site) on the device.

Drop-in counterparts of ``contractions/effective_hamiltonians.py:20-470``, the
functions the non-Krylov time-evolution modes and DMRG use to build the
effective operator as a matrix:

* ``find_tensor_leg_permutation``              -- ``:20-52``
* ``contract_all_except_node``                 -- ``:54-104``
* ``get_effective_single_site_hamiltonian(_nodes)`` -- ``:106-149``
* ``get_effective_bond_hamiltonian_tensor`` / ``_nodes`` / ``get_effective_bond_hamiltonian``
                                               -- ``:151-240``
* ``contract_all_except_two_nodes``            -- ``:242-298``
* ``get_effective_two_site_hamiltonian(_nodes)`` -- ``:394-470``

Leg conventions are the reference's: the effective tensor has the output
(bra-side) legs first, in the state node's neighbour order followed by its
open legs, then the input (ket-side) legs in the same order; the matrix is that
tensor with the two halves flattened.  The TTNO node's children may be in a
different order than the state node's; operator legs are matched to state
legs by neighbour identifier.

The reference contracts the cached blocks into the Hamiltonian tensor one by
one and then permutes the result.  Here the whole network (TTNO tensor and
environment blocks) is one labelled network handed to the contraction engine,
whose output labels are the final leg order, so no transposed copy is made.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import torch

from . import engine as E
from .ttn import Node, TreeTensorNetworkState
from .ttno import TreeTensorNetworkOperator
from .util import NotCompatibleException


def find_tensor_leg_permutation(state_node: Node, hamiltonian_node: Node) -> Tuple[int, ...]:
    """Permutation taking the tensor of ``contract_all_except_node`` as the reference
    first forms it -- legs (out, in, then (ket, bra) per block in the Hamiltonian
    node's neighbour order) -- to (bra legs, out, ket legs, in) in the state node's
    neighbour order."""
    ham_idx = [hamiltonian_node.neighbour_index(nb) for nb in state_node.neighbouring_nodes()]
    n_open = state_node.nopen_legs()
    out_legs = [2 * n_open + 2 * h + 1 for h in ham_idx] + list(range(n_open))
    in_legs = [2 * n_open + 2 * h for h in ham_idx] + list(range(n_open, 2 * n_open))
    return tuple(out_legs + in_legs)


def _half_matrix(t: torch.Tensor) -> torch.Tensor:
    """Reference ``tensor_matricisation_half``: first half of the legs as rows."""
    half = t.dim() // 2
    rows = math.prod(t.shape[:half])
    return t.reshape(rows, -1)


def _env_labels(i: int):
    return (("a", i), ("w", i), ("b", i))


def contract_all_except_node(state_node: Node, hamiltonian_node: Node,
                             hamiltonian_tensor: torch.Tensor, tensor_cache
                             ) -> torch.Tensor:
    """Every cached block into ``state_node`` contracted with the node's TTNO tensor:
    legs (bra per state neighbour, out, ket per state neighbour, in)."""
    nid = state_node.identifier
    nbs = state_node.neighbouring_nodes()
    k = len(nbs)
    if sorted(nbs) != sorted(hamiltonian_node.neighbouring_nodes()):
        raise NotCompatibleException(f"the state and Hamiltonian nodes of {nid} have "
                                     "different neighbours")
    w_lab: List = [None] * (k + 2)
    for i, nb in enumerate(nbs):
        w_lab[hamiltonian_node.neighbour_index(nb)] = ("w", i)
    w_lab[k], w_lab[k + 1] = "q", "p"
    tensors = [hamiltonian_tensor]
    labels = [w_lab]
    for i, nb in enumerate(nbs):
        tensors.append(tensor_cache.get_entry(nb, nid))
        labels.append(_env_labels(i))
    out = [("b", i) for i in range(k)] + ["q"] + [("a", i) for i in range(k)] + ["p"]
    return E.contract_network(tensors, labels, out, out_dtype=torch.complex128)


def get_effective_single_site_hamiltonian_nodes(state_node: Node, hamiltonian_node: Node,
                                                hamiltonian_tensor: torch.Tensor,
                                                tensor_cache) -> torch.Tensor:
    """The effective site Hamiltonian as a matrix (rows: bra legs and out)."""
    return _half_matrix(contract_all_except_node(state_node, hamiltonian_node,
                                                 hamiltonian_tensor, tensor_cache))


def get_effective_single_site_hamiltonian(node_id: str, state: TreeTensorNetworkState,
                                          hamiltonian: TreeTensorNetworkOperator,
                                          tensor_cache) -> torch.Tensor:
    ham_node, ham_tensor = hamiltonian[node_id]
    return get_effective_single_site_hamiltonian_nodes(state.nodes[node_id], ham_node,
                                                       ham_tensor, tensor_cache)


def get_effective_bond_hamiltonian_tensor(state_node: Node, tensor_cache) -> torch.Tensor:
    """The effective bond Hamiltonian of a link node (one parent, one child) as a
    tensor with legs (bra parent side, bra child side, ket parent side, ket child
    side): the two blocks across the link with their TTNO legs traced."""
    if state_node.nneighbours() != 2 or state_node.parent is None:
        raise ValueError("The effective bond Hamiltonian can only be "
                         "computed for a bond with two neighbours.")
    p, c = state_node.parent, state_node.children[0]
    tp, tc = tensor_cache.get_entry(p, c), tensor_cache.get_entry(c, p)
    return E.contract_network([tp, tc], [["ap", "w", "bp"], ["ac", "w", "bc"]],
                              ["bp", "bc", "ap", "ac"], out_dtype=torch.complex128)


def get_effective_bond_hamiltonian_nodes(state_node: Node, tensor_cache) -> torch.Tensor:
    return _half_matrix(get_effective_bond_hamiltonian_tensor(state_node, tensor_cache))


def get_effective_bond_hamiltonian(bond_node_id: str, state: TreeTensorNetworkState,
                                   tensor_cache) -> torch.Tensor:
    return get_effective_bond_hamiltonian_nodes(state.nodes[bond_node_id], tensor_cache)


def contract_all_except_two_nodes(state_node: Node, target_node: Node,
                                  target_tensor: torch.Tensor, next_node: Node,
                                  next_tensor: torch.Tensor, tensor_cache) -> torch.Tensor:
    """The two-site effective tensor of the merged node ``state_node`` (whose
    neighbours are the outer neighbours of the two sites, in any order, and whose
    open legs are the target's then the next site's): legs (bra per merged
    neighbour, out target, out next, ket per merged neighbour, in target, in next)."""
    t_id, n_id = target_node.identifier, next_node.identifier
    outer_t = [nb for nb in target_node.neighbouring_nodes() if nb != n_id]
    outer_n = [nb for nb in next_node.neighbouring_nodes() if nb != t_id]
    nbs = state_node.neighbouring_nodes()
    k = len(nbs)
    owner = {}
    for i, nb in enumerate(nbs):
        if nb in outer_t:
            owner[nb] = t_id
        elif nb in outer_n:
            owner[nb] = n_id
        else:
            raise NotCompatibleException("The two-site Hamiltonian has a neighbour that is "
                                         "not a neighbour of the two sites.")
    if len(owner) != len(outer_t) + len(outer_n):
        raise NotCompatibleException("the merged node does not carry every outer "
                                     "neighbour of the two sites")

    def w_labels(node: Node, other: str, q: str, p: str) -> list:
        lab: List = [None] * (node.nneighbours() + 2)
        for nb in node.neighbouring_nodes():
            lab[node.neighbour_index(nb)] = "wl" if nb == other else ("w", nbs.index(nb))
        lab[-2], lab[-1] = q, p
        return lab

    tensors = [target_tensor, next_tensor]
    labels = [w_labels(target_node, n_id, "q1", "p1"), w_labels(next_node, t_id, "q2", "p2")]
    for i, nb in enumerate(nbs):
        tensors.append(tensor_cache.get_entry(nb, owner[nb]))
        labels.append(_env_labels(i))
    out = [("b", i) for i in range(k)] + ["q1", "q2"] + \
        [("a", i) for i in range(k)] + ["p1", "p2"]
    return E.contract_network(tensors, labels, out, out_dtype=torch.complex128)


def get_effective_two_site_hamiltonian_nodes(state_node: Node, target_node: Node,
                                             target_tensor: torch.Tensor, next_node: Node,
                                             next_tensor: torch.Tensor,
                                             tensor_cache) -> torch.Tensor:
    return _half_matrix(contract_all_except_two_nodes(state_node, target_node, target_tensor,
                                                      next_node, next_tensor, tensor_cache))


def two_site_node(state: TreeTensorNetworkState, target_node_id: str,
                  next_node_id: str) -> Node:
    """Metadata of the merged node in this package's two-site layout (the layout
    ``splitting.merge_two_sites`` writes): the target's outer neighbours in its
    neighbour order, then the next site's, then the two physical legs."""
    t, n = state.nodes[target_node_id], state.nodes[next_node_id]
    outer = [nb for nb in t.neighbouring_nodes() if nb != next_node_id] + \
        [nb for nb in n.neighbouring_nodes() if nb != target_node_id]
    node = Node(identifier=create_two_site_id(target_node_id, next_node_id))
    # the merged node only needs its neighbour order: the first outer neighbour
    # plays the parent, the rest the children
    node.parent = outer[0] if outer else None
    node.children = outer[1:]
    dims = [state.nodes[target_node_id].neighbour_dim(nb) if nb in t.neighbouring_nodes()
            else state.nodes[next_node_id].neighbour_dim(nb) for nb in outer]
    node.shape = tuple(dims) + (t.shape[-1], n.shape[-1])
    return node


def create_two_site_id(node_id: str, next_node_id: str) -> str:
    return "TwoSite_" + node_id + "_contr_" + next_node_id


def get_effective_two_site_hamiltonian(target_node_id: str, next_node_id: str,
                                       state: TreeTensorNetworkState,
                                       hamiltonian: TreeTensorNetworkOperator,
                                       tensor_cache,
                                       state_node: Optional[Node] = None) -> torch.Tensor:
    """The effective two-site Hamiltonian of ``target_node_id`` and ``next_node_id``
    as a matrix.  The reference reads the merged node from the state, where its
    two-site step has put it; the tree here is never changed by a merge, so the
    merged node's leg order is ``state_node`` when given, else this package's
    two-site layout (``two_site_node``)."""
    t_node, t_tensor = hamiltonian[target_node_id]
    n_node, n_tensor = hamiltonian[next_node_id]
    if state_node is None:
        state_node = two_site_node(state, target_node_id, next_node_id)
    return get_effective_two_site_hamiltonian_nodes(state_node, t_node, t_tensor, n_node,
                                                    n_tensor, tensor_cache)
