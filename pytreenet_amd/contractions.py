"""Sandwich contractions, the environment cache and observables on the device.

Drop-in counterparts of the reference functions (same names and argument
meaning):

* ``PartialTreeCachDict``        -- ``contractions/tree_cach_dict.py:16-150``
* ``SandwichCache``              -- ``contractions/sandwich_caching.py:14-215``
* ``contract_any``               -- ``contractions/state_operator_contraction.py:407-461``
* ``contract_ket_ham_with_envs`` -- ``:227-273`` (H_eff x)
* ``contract_bond_tensor``       -- ``:984-1040`` (bond action)
* ``expectation_value``          -- ``:26-74`` (<psi|H|psi>, unnormalised)
* ``contract_two_ttns``          -- ``contractions/state_state_contraction.py:29-66``

Every contraction is a labelled network handed to ``engine.contract_network``:
the pairwise order is the flop-optimal one, each pairwise step is one
gather-GEMM launch, and the environment blocks keep the reference layout
(ket, ham, bra).  With ``precision="fp32"`` the blocks are stored as
complex64 and the GEMMs run as 3M FP32 MFMA; inputs in complex128 (Krylov
vectors) are rounded while being staged.
"""
from __future__ import annotations

from typing import Callable, Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import engine as E
from .ttn import Node, TreeTensorNetwork, TreeTensorNetworkState
from .ttno import TensorProduct, TreeTensorNetworkOperator


class PartialTreeCachDict(dict):
    """(node_id, next_node_id) -> environment block with legs (ket, ham, bra)."""

    def add_entry(self, node_id: str, next_node_id: str, tensor: torch.Tensor):
        self[(node_id, next_node_id)] = tensor

    def get_entry(self, node_id: str, next_node_id: str) -> torch.Tensor:
        return self[(node_id, next_node_id)]

    def delete_entry(self, node_id: str, next_node_id: str):
        del self[(node_id, next_node_id)]

    def contains(self, node_id: str, next_node_id: str) -> bool:
        return (node_id, next_node_id) in self

    def change_next_id_for_entry(self, node_id: str, old_next_id: str, new_next_id: str):
        self[(node_id, new_next_id)] = self.pop((node_id, old_next_id))

    def num_legs(self, node_id: str, next_node_id: str) -> int:
        return self[(node_id, next_node_id)].dim()


# ---------------------------------------------------------------------------------------
# network builders
# ---------------------------------------------------------------------------------------
def _env_labels(i: int):
    return (("a", i), ("w", i), ("b", i))


def _ham_legs_map(ket_node: Node, op_node: Node) -> List[int]:
    """Operator leg index for each ket neighbour index (child orders may differ)."""
    return [op_node.neighbour_index(nb) for nb in ket_node.neighbouring_nodes()]


def _op_labels(ket_node: Node, op_node: Node) -> list:
    k = ket_node.nneighbours()
    labels = [None] * (k + 2)
    for i, j in enumerate(_ham_legs_map(ket_node, op_node)):
        labels[j] = ("w", i)
    labels[k] = "q"      # out, meets the bra
    labels[k + 1] = "p"  # in, meets the ket
    return labels


USE_BLOCK_SPARSE = True
# Site tensors below this many elements meet the dense TTNO tensor: their networks
# are bound by launches, and block sparsity splits each operator step into one GEMM
# per block group (an environment update of a 256-element site tensor: about 30
# launches instead of 9).
BLOCK_SPARSE_MIN_SIZE = 1 << 16


def operator_tensor(operator: TreeTensorNetworkOperator, node_id: str, precision: str,
                    block: Optional[bool] = None, size: Optional[int] = None):
    """TTNO node tensor as handed to the networks: block sparse over its bond-state
    legs (structural zeros skipped exactly) unless ``block`` is off; by default
    (``block=None``) when ``USE_BLOCK_SPARSE`` is on and the site tensor it meets has
    ``size`` >= ``BLOCK_SPARSE_MIN_SIZE`` elements (or its size is not given)."""
    dt = E.storage_dtype(precision)
    if block is None:
        block = USE_BLOCK_SPARSE and (size is None or size >= BLOCK_SPARSE_MIN_SIZE)
    if block:
        return operator.block_tensor(node_id, dt)
    return operator.tensor_in(node_id, dt)


# Environment blocks in which the operator on the environment's side is the
# identity (``TTNO.identity_states``) equal the identity matrix when that side of
# the state is an isometry towards the node the block points to.  The time
# evolution keeps that canonical form, so its effective Hamiltonians and
# environment updates skip those blocks (``engine._split_identity``) for site
# tensors of at least IDENTITY_MIN_SIZE elements (the ones that meet the
# block-sparse operator tensor).
IDENTITY_BLOCKS = True
IDENTITY_MIN_SIZE = 1 << 16


def identity_slots(node_id: str, neighbours: Sequence[Optional[str]],
                   operator: TreeTensorNetworkOperator, first_slot: int
                   ) -> Dict[int, Tuple[int, ...]]:
    """{network slot: identity states} of the environment blocks E[(nb, node_id)]
    sitting at slots first_slot, first_slot + 1, ... (None: no block there)."""
    out = {}
    slot = first_slot
    for nb in neighbours:
        if nb is None:
            continue
        out[slot] = operator.identity_states(nb, node_id)
        slot += 1
    return out


def contract_any(node_id: str, next_node_id: str, state: TreeTensorNetworkState,
                 operator: TreeTensorNetworkOperator, dictionary: PartialTreeCachDict,
                 bra_state: Optional[TreeTensorNetworkState] = None,
                 bra_state_conjugated: bool = True, id_trafo_op=None, id_trafo_bra=None,
                 precision: str = E.FP64, canonical: bool = False) -> torch.Tensor:
    """Directed environment of the subtree at ``node_id`` with open legs to ``next_node_id``.

    ``next_node_id`` may be a node that is not a tree neighbour (the link tensor
    of a one-site update): then the block points along the leg that
    ``dictionary`` does not hold for ``node_id``.  ``canonical`` states that the
    subtrees behind the blocks used are isometries towards ``node_id`` (and the
    bra is the ket), so their identity blocks are skipped.
    """
    ket_node, ket = state[node_id]
    op_node = operator.nodes[node_id]
    nbs = ket_node.neighbouring_nodes()
    if next_node_id in nbs:
        to = nbs.index(next_node_id)
    else:
        missing = [i for i, nb in enumerate(nbs) if (nb, node_id) not in dictionary]
        if len(missing) != 1:
            raise KeyError(f"cannot infer the open leg of {node_id} towards {next_node_id}")
        to = missing[0]
    ident = None
    if canonical and bra_state is None and IDENTITY_BLOCKS and ket.numel() >= IDENTITY_MIN_SIZE:
        ident = identity_slots(node_id, [nb if i != to else None for i, nb in enumerate(nbs)],
                               operator, 3)
    return env_block(ket_node, ket, op_node,
                     operator_tensor(operator, node_id, precision, size=ket.numel()),
                     [dictionary.get((nb, node_id)) if i != to else None
                      for i, nb in enumerate(nbs)], to,
                     bra=None if bra_state is None else bra_state.tensors[node_id],
                     bra_conj=bra_state_conjugated, precision=precision, identity=ident)


def env_block(ket_node: Node, ket: torch.Tensor, op_node: Node, op: torch.Tensor,
              envs: List[Optional[torch.Tensor]], to: int, bra: Optional[torch.Tensor] = None,
              bra_conj: bool = True, precision: str = E.FP64,
              identity: Optional[Dict[int, Tuple[int, ...]]] = None) -> torch.Tensor:
    k = ket_node.nneighbours()
    a_lab = [("a", i) for i in range(k)] + ["p"]
    b_lab = [("b", i) for i in range(k)] + ["q"]
    tensors = [ket, op, ket if bra is None else bra]
    labels = [a_lab, _op_labels(ket_node, op_node), b_lab]
    conj = [False, False, True if bra is None else bra_conj]
    for i in range(k):
        if i == to:
            continue
        tensors.append(envs[i])
        labels.append(_env_labels(i))
        conj.append(False)
    return E.contract_network_cached(tensors, labels, _env_labels(to), conj=conj,
                                     precision=precision, out_dtype=E.storage_dtype(precision),
                                     identity=identity)


def heff_network(ket_node: Node, op_node: Node, op: torch.Tensor,
                 envs: List[torch.Tensor]):
    """(tensors, labels, out_labels) of H_eff without the ket (slot 0 left to the caller)."""
    k = ket_node.nneighbours()
    labels = [[("a", i) for i in range(k)] + ["p"], _op_labels(ket_node, op_node)]
    tensors = [None, op]
    for i in range(k):
        tensors.append(envs[i])
        labels.append(_env_labels(i))
    out = [("b", i) for i in range(k)] + ["q"]
    return tensors, labels, out


def contract_ket_ham_with_envs(ket_node: Node, ket_tensor: torch.Tensor, ham_node: Node,
                               ham_tensor: torch.Tensor, dictionary: PartialTreeCachDict,
                               precision: str = E.FP64,
                               out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """y = H_eff x; output legs (bra neighbour legs..., out) in the ket's node order."""
    envs = [dictionary.get_entry(nb, ket_node.identifier)
            for nb in ket_node.neighbouring_nodes()]
    tensors, labels, out_l = heff_network(ket_node, ham_node, ham_tensor, envs)
    tensors[0] = ket_tensor
    return E.contract_network(tensors, labels, out_l, precision=precision, out=out)


def contract_bond_tensor(ket_tensor: torch.Tensor, ket_node: Node,
                         tensor_cache: PartialTreeCachDict, precision: str = E.FP64,
                         out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Bond action on a link tensor with legs (parent side, child side).

    The link node's parent ``p`` and only child ``c`` give the blocks
    E[(p, c)] and E[(c, p)]; the TTNO legs of the two blocks are traced together.
    """
    if ket_node.nneighbours() != 2:
        raise ValueError("The node must have exactly two neighbours to perform effective "
                         f"bond time evolution! Node {ket_node.identifier} has "
                         f"{ket_node.nneighbours()} neighbours.")
    p, c = ket_node.parent, ket_node.children[0]
    return bond_apply(ket_tensor, tensor_cache.get_entry(p, c), tensor_cache.get_entry(c, p),
                      precision=precision, out=out)


def bond_apply(x: torch.Tensor, env_0: torch.Tensor, env_1: torch.Tensor,
               precision: str = E.FP64, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """y[b0, b1] = sum x[a0, a1] env_0[a0, w, b0] env_1[a1, w, b1]."""
    return E.contract_network([x, env_0, env_1],
                              [["a0", "a1"], ["a0", "w", "b0"], ["a1", "w", "b1"]],
                              ["b0", "b1"], precision=precision, out=out)


# ---------------------------------------------------------------------------------------
# cache
# ---------------------------------------------------------------------------------------
class SandwichCache(PartialTreeCachDict):
    """Environment cache bound to a state and a Hamiltonian (``sandwich_caching.py``)."""

    def __init__(self, state: TreeTensorNetworkState, hamiltonian: TreeTensorNetworkOperator,
                 bra_state: Optional[TreeTensorNetworkState] = None,
                 precision: str = E.FP64):
        super().__init__()
        self.state = state
        self.hamiltonian = hamiltonian
        self.bra_state = bra_state
        self.precision = precision

    @classmethod
    def init_cache_but_one(cls, state: TreeTensorNetworkState,
                           hamiltonian: TreeTensorNetworkOperator, node_id: str,
                           bra_state=None, precision: str = E.FP64,
                           canonical: bool = False) -> "SandwichCache":
        """All directed blocks pointing towards ``node_id``, leaves first."""
        cache = cls(state, hamiltonian, bra_state=bra_state, precision=precision)
        cache.canonical = canonical
        for nid, to in caching_path(state, node_id):
            cache.update_tree_cache(nid, to)
        return cache

    # set by the time evolution, which keeps the state canonical around the
    # blocks it builds (``contract_any(canonical=...)``)
    canonical = False

    def update_tree_cache(self, node_id: str, next_node_id: str):
        self.add_entry(node_id, next_node_id,
                       contract_any(node_id, next_node_id, self.state, self.hamiltonian, self,
                                    bra_state=self.bra_state, precision=self.precision,
                                    canonical=self.canonical))


def caching_path(state: TreeTensorNetwork, target: str) -> List[Tuple[str, str]]:
    """(node, next) pairs so that every block's inputs exist before it is built."""
    dist = state.distance_to_node(target)
    out = []
    for d in range(max(dist.values()), 0, -1):
        for nid in [n for n in dist if dist[n] == d]:
            out.append((nid, min(state.neighbours(nid), key=lambda x: dist[x])))
    return out


# ---------------------------------------------------------------------------------------
# observables
# ---------------------------------------------------------------------------------------
def expectation_value(state: TreeTensorNetworkState, operator: TreeTensorNetworkOperator,
                      precision: str = E.FP64) -> complex:
    """<psi|H|psi> (unnormalised): blocks children-first, root contracted last."""
    cache = PartialTreeCachDict()
    for nid in state.linearise()[:-1]:
        node = state.nodes[nid]
        cache.add_entry(nid, node.parent,
                        contract_any(nid, node.parent, state, operator, cache,
                                     precision=precision))
        for c in node.children:
            cache.delete_entry(c, nid)
    r = state.root_id
    node, ket = state[r]
    op_node = operator.nodes[r]
    k = node.nneighbours()
    tensors = [ket, operator_tensor(operator, r, precision, size=ket.numel()), ket]
    labels = [[("a", i) for i in range(k)] + ["p"], _op_labels(node, op_node),
              [("b", i) for i in range(k)] + ["q"]]
    for i, nb in enumerate(node.neighbouring_nodes()):
        tensors.append(cache.get_entry(nb, r))
        labels.append(_env_labels(i))
    res = E.contract_network(tensors, labels, [], conj=[False, False, True] + [False] * k,
                             precision=precision)
    return complex(res.cpu().item())


def _overlap_block(node: Node, ket: torch.Tensor, bra: torch.Tensor,
                   envs: List[Optional[torch.Tensor]], to: Optional[int],
                   bra_conj: bool) -> torch.Tensor:
    k = node.nneighbours()
    tensors = [ket, bra]
    labels = [[("a", i) for i in range(k)] + ["p"], [("b", i) for i in range(k)] + ["p"]]
    conj = [False, bra_conj]
    for i in range(k):
        if i == to:
            continue
        tensors.append(envs[i])
        labels.append([("a", i), ("b", i)])
        conj.append(False)
    out = [] if to is None else [("a", to), ("b", to)]
    return E.contract_network(tensors, labels, out, conj=conj)


def contract_two_ttns(ttn1: TreeTensorNetworkState, ttn2: TreeTensorNetworkState,
                      conj_second: bool = False) -> complex:
    """Full contraction sum ttn1 * ttn2 (optionally with ttn2 conjugated on load)."""
    cache: Dict[Tuple[str, str], torch.Tensor] = {}
    for nid in ttn1.linearise()[:-1]:
        node = ttn1.nodes[nid]
        nbs = node.neighbouring_nodes()
        envs = [cache.get((nb, nid)) for nb in nbs]
        cache[(nid, node.parent)] = _overlap_block(node, ttn1.tensors[nid], ttn2.tensors[nid],
                                                   envs, 0, conj_second)
        for c in node.children:
            cache.pop((c, nid))
    r = ttn1.root_id
    node = ttn1.nodes[r]
    envs = [cache[(nb, r)] for nb in node.neighbouring_nodes()]
    res = _overlap_block(node, ttn1.tensors[r], ttn2.tensors[r], envs, None, conj_second)
    return complex(res.cpu().item())


def scalar_product(ttn1: TreeTensorNetworkState, ttn2: TreeTensorNetworkState) -> complex:
    """Reference ``TTNS.scalar_product(other)``: sum ttn1 * conj(ttn2)."""
    return contract_two_ttns(ttn1, ttn2, conj_second=True)


def tensor_product_expectation_value(state: TreeTensorNetworkState,
                                     operator: TensorProduct) -> complex:
    """<psi| prod_i O_i |psi> (unnormalised) by one ket/bra sweep with O applied on load."""
    if len(operator) == 0:
        return state.scalar_product()
    cache: Dict[Tuple[str, str], torch.Tensor] = {}
    dev = state.device
    ops = {nid: torch.as_tensor(np.asarray(o), dtype=torch.complex128).to(dev)
           for nid, o in operator.items()}

    def block(nid, to_idx):
        node, ket = state[nid]
        k = node.nneighbours()
        tensors = [ket, ket]
        labels = [[("a", i) for i in range(k)] + ["p"], [("b", i) for i in range(k)] + ["q"]]
        conj = [False, True]
        if nid in ops:
            tensors.append(ops[nid])
            labels.append(["q", "p"])
            conj.append(False)
        else:
            labels[1][-1] = "p"
        nbs = node.neighbouring_nodes()
        for i in range(k):
            if i == to_idx:
                continue
            tensors.append(cache[(nbs[i], nid)])
            labels.append([("a", i), ("b", i)])
            conj.append(False)
        out = [] if to_idx is None else [("a", to_idx), ("b", to_idx)]
        return E.contract_network(tensors, labels, out, conj=conj)

    for nid in state.linearise()[:-1]:
        node = state.nodes[nid]
        cache[(nid, node.parent)] = block(nid, 0)
        for c in node.children:
            cache.pop((c, nid))
    return complex(block(state.root_id, None).cpu().item())


def _as_matrix(op, nid: str):
    if isinstance(op, TensorProduct):
        if list(op.keys()) != [nid]:
            raise ValueError(f"operator {op} is not a single-site operator on {nid}")
        op = op[nid]
    return np.asarray(op)


def multi_single_site_expectation_value(ttns: TreeTensorNetworkState, operators,
                                        move_orth: bool = False) -> Dict[str, List[complex]]:
    """<psi| O |psi> (unnormalised) for many single-site operators in two passes
    (``ttns/ttns_util.py:95-158``).

    ``operators`` maps node ids to a matrix, a single-site ``TensorProduct`` or a
    list of them (or is one ``TensorProduct`` of single-site factors); the result
    maps each node id to the list of its values.  One upward and one downward pass
    build every directed ket/bra block of the tree (2(N-1) contractions), and each
    value is one contraction of the node with its operator and its blocks, so the
    cost is O(N) contractions instead of one full sweep per operator.  The
    reference moves the orthogonality centre to the root and uses identity blocks
    below it; these blocks are contracted, so the values are exact in any gauge
    and the state is not modified (``move_orth`` moves the centre to the root
    afterwards, with KEEP QR, as the reference's option does).
    """
    from .splitting import SplitMode
    if isinstance(operators, TensorProduct):
        prepared = {nid: [op] for nid, op in operators.items()}
    else:
        prepared = {nid: (list(ops) if isinstance(ops, (list, tuple)) else [ops])
                    for nid, ops in operators.items()}
    dev = ttns.device
    blocks: Dict[Tuple[str, str], torch.Tensor] = {}
    for nid in ttns.linearise()[:-1]:          # children first: blocks towards the root
        node = ttns.nodes[nid]
        nbs = node.neighbouring_nodes()
        blocks[(nid, node.parent)] = _overlap_block(
            node, ttns.tensors[nid], ttns.tensors[nid], [blocks.get((nb, nid)) for nb in nbs],
            0, True)
    for nid in reversed(ttns.linearise()):     # parents first: blocks away from the root
        node = ttns.nodes[nid]
        nbs = node.neighbouring_nodes()
        for c in node.children:
            blocks[(nid, c)] = _overlap_block(
                node, ttns.tensors[nid], ttns.tensors[nid],
                [blocks.get((nb, nid)) for nb in nbs], nbs.index(c), True)
    out: Dict[str, List[complex]] = {}
    for nid, ops in prepared.items():
        node, ket = ttns[nid]
        k = node.nneighbours()
        base_t = [ket, ket]
        base_l = [[("a", i) for i in range(k)] + ["p"], [("b", i) for i in range(k)] + ["q"]]
        for i, nb in enumerate(node.neighbouring_nodes()):
            base_t.append(blocks[(nb, nid)])
            base_l.append([("a", i), ("b", i)])
        vals = []
        for op in ops:
            m = torch.as_tensor(_as_matrix(op, nid), dtype=torch.complex128).to(dev)
            res = E.contract_network(base_t + [m], base_l + [["q", "p"]], [],
                                     conj=[False, True] + [False] * (k + 1))
            vals.append(res)
        out[nid] = [complex(v) for v in torch.stack(vals).cpu().tolist()] if vals else []
    if move_orth:
        ttns.ensure_root_orth_center(mode=SplitMode.KEEP)
    return out
