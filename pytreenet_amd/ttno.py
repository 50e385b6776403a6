"""Tree tensor network operators and the Hamiltonians of the benchmark models.

``TreeTensorNetworkOperator`` mirrors ``ttno/ttno_class.py:35-138``: one tensor
per node with legs (parent, children..., out, in); ``out`` meets the bra and
``in`` the ket (``contractions/state_operator_contraction.py:1042-1069``).

The reference builds TTNOs from symbolic Hamiltonians through a state diagram
and symbolic Gaussian elimination (``ttno_class.py:59-138``), a host-side
one-off construction outside the hot path.  Here ``from_edge_terms`` builds the
operator directly for the Hamiltonians the time evolution uses: one-body terms
on nodes and two-body terms on tree edges, with the finite-state construction

    bond state 0   nothing placed in the subtree below the edge (identity),
    bond state 1   one complete term placed in the subtree,
    bond state 2+j operator of edge term j placed at the child, waiting for
                   its partner at the parent (coefficient applied there),

so an edge carrying t two-body terms has TTNO bond dimension 2 + t (3 for the
transverse-field Ising model, 5 for XXZ/XYZ/Heisenberg), the dimension the
reference's state diagram gives for these models.  Sign convention of the
reference models (``operators/models/two_site_model.py:126-288``):
H = -sum_e (Jx XX + Jy YY + Jz ZZ) - sum_i g_i Z_i.
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .ttn import Node, TreeTensorNetwork, as_device_tensor

PAULI_X = np.array([[0, 1], [1, 0]], dtype=complex)
PAULI_Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
PAULI_Z = np.array([[1, 0], [0, -1]], dtype=complex)
IDENTITY2 = np.eye(2, dtype=complex)


def pauli_matrices() -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    return PAULI_X, PAULI_Y, PAULI_Z


class NumericOperator:
    """An operator on one or more nodes (``operators/operator.py:20-108``).

    ``operator`` is a host array: a matrix (out, in) or a tensor with legs
    (out_1..out_n, in_1..in_n), the legs in ``node_identifiers`` order."""

    def __init__(self, operator: np.ndarray, node_identifiers):
        operator = np.asarray(operator)
        if operator.ndim % 2 != 0:
            raise ValueError("an operator needs as many output legs as input legs")
        self.operator = operator
        self.node_identifiers = ([node_identifiers] if isinstance(node_identifiers, str)
                                 else list(node_identifiers))

    def to_matrix(self) -> "NumericOperator":
        half = self.operator.ndim // 2
        rows = int(np.prod(self.operator.shape[:half]))
        return NumericOperator(self.operator.reshape(rows, -1), list(self.node_identifiers))

    def to_tensor(self, dim: Optional[int] = None, ttn=None) -> "NumericOperator":
        """The operator with one output and one input leg per node: every node of
        dimension ``dim``, or each node's open dimension in ``ttn``."""
        if dim is None and ttn is None:
            raise ValueError("`dim` and `ttn` cannot both be `None`!")
        if dim is not None:
            if dim <= 0:
                raise ValueError("dimension has to be positive")
            shape = [dim] * len(self.node_identifiers)
        else:
            shape = [ttn.nodes[n].open_dimension() for n in self.node_identifiers]
        return NumericOperator(self.operator.reshape(shape * 2), list(self.node_identifiers))

    def is_unitary(self) -> bool:
        m = self.to_matrix().operator
        return bool(np.allclose(m @ m.conj().T, np.eye(m.shape[0])))


class TensorProduct(dict):
    """{node_id: operator matrix} -- a product of single-site operators
    (``operators/tensorproduct.py``)."""

    def into_operator(self, conversion_dict: Optional[Dict[str, np.ndarray]] = None,
                      order: Optional[List[str]] = None) -> NumericOperator:
        """The Kronecker product of the factors, the first identifier most significant
        (``tensorproduct.py:170-205``).  Symbolic (string) factors are looked up in
        ``conversion_dict``."""
        order = list(self.keys()) if order is None else list(order)
        total = np.ones((1, 1), dtype=complex)
        for nid in order:
            op = self[nid]
            if isinstance(op, str):
                if conversion_dict is None:
                    raise TypeError("If the tensor_product contains symbolic operators, "
                                    "there must be a dictionary for conversion!")
                op = conversion_dict[op]
            total = np.kron(total, np.asarray(op))
        return NumericOperator(total, list(self.keys()))

    def exp(self, factor=1) -> NumericOperator:
        """exp(factor * product) as a matrix on the nodes in key order
        (``tensorproduct.py:207-224``)."""
        from scipy.linalg import expm
        total = self.into_operator()
        return NumericOperator(expm(factor * total.operator), total.node_identifiers)


class TreeTensorNetworkOperator(TreeTensorNetwork):
    """TTNO with legs (parent, children..., out, in) per node."""

    def __init__(self):
        super().__init__()
        self._lowp: Dict[str, torch.Tensor] = {}
        self._blocks: Dict[tuple, object] = {}

    def block_tensor(self, node_id: str, dtype: torch.dtype):
        """The node tensor in block-sparse form over its virtual (bond-state) legs."""
        from .engine import BlockSparse
        t = self.tensor_in(node_id, dtype)
        hit = self._blocks.get((node_id, dtype))
        if hit is None or hit.shape != tuple(t.shape) or hit.source is not t:
            hit = BlockSparse.from_dense(t, range(self._nodes[node_id].nneighbours()))
            hit.source = t
            self._blocks[(node_id, dtype)] = hit
        return hit

    def identity_states(self, from_id: str, to_id: str) -> Tuple[int, ...]:
        """Bond states of the leg of ``from_id`` towards its neighbour ``to_id`` in
        which the operator on ``from_id``'s side of that bond is exactly the
        identity.

        A state s qualifies when the node tensor, with that leg at s, has exactly
        one non-zero block over its other virtual legs, that block equals the
        identity on (out, in), and each of those other legs sits at a qualifying
        state of its own neighbour's side (recursively, away from ``to_id``).
        The environment of that side is then, in state s, the Gram matrix of the
        side's state tensors: the identity whenever the side is an isometry
        towards ``to_id`` (the canonical form a time evolution keeps).  In the
        edge-term construction these are state 0 towards a parent and state 1
        towards a child."""
        key = tuple((nid, id(t)) for nid, t in self._tensors.items())
        if getattr(self, "_ident_key", None) != key:
            self._ident_key = key
            self._ident: Dict[Tuple[str, str], Tuple[int, ...]] = {}
            self._ident_host: Dict[str, np.ndarray] = {}
        return self._identity_rec(from_id, to_id)

    def _identity_rec(self, c: str, n: str) -> Tuple[int, ...]:
        hit = self._ident.get((c, n))
        if hit is not None:
            return hit
        node = self._nodes[c]
        W = self._ident_host.get(c)
        if W is None:
            W = self._ident_host[c] = self._tensors[c].detach().cpu().numpy()
        leg_n = node.neighbour_index(n)
        need = {}
        for m in node.neighbouring_nodes():
            if m == n:
                continue
            sm = self._identity_rec(m, c)
            if len(sm) != 1:
                self._ident[(c, n)] = ()
                return ()
            need[node.neighbour_index(m)] = sm[0]
        nv = node.nneighbours()
        d_out, d_in = W.shape[nv], W.shape[nv + 1]
        eye = np.eye(d_out, d_in, dtype=W.dtype)
        found = []
        for s in range(W.shape[leg_n]):
            sl = np.take(W, s, axis=leg_n)          # other virtual legs, out, in
            other = [a for a in range(nv) if a != leg_n]
            flat = sl.reshape(-1, d_out * d_in)
            nz = np.nonzero(np.any(flat != 0, axis=1))[0]
            if len(nz) != 1:
                continue
            idx = np.unravel_index(nz[0], sl.shape[:nv - 1]) if nv > 1 else ()
            if any(int(i) != need[a] for i, a in zip(idx, other)):
                continue
            if np.array_equal(flat[nz[0]].reshape(d_out, d_in), eye):
                found.append(s)
        self._ident[(c, n)] = tuple(found)
        return self._ident[(c, n)]

    def two_site_block_tensor(self, n1: str, n2: str, nbs1: Sequence[str],
                              nbs2: Sequence[str], dtype: torch.dtype):
        """The operator of the neighbouring nodes n1, n2 contracted over their bond,
        block sparse over its outer bond legs: legs (legs of n1 towards nbs1...,
        legs of n2 towards nbs2..., out1, out2, in1, in2) in the order given.
        Built on the host once per pair and dtype (cached)."""
        from .engine import BlockSparse
        key = ("two", n1, n2, tuple(nbs1), tuple(nbs2), dtype)
        t1, t2 = self._tensors[n1], self._tensors[n2]
        hit = self._blocks.get(key)
        if hit is not None and hit.source[0] is t1 and hit.source[1] is t2:
            return hit
        a1, a2 = self._nodes[n1], self._nodes[n2]
        W1 = t1.detach().cpu().numpy()
        W2 = t2.detach().cpu().numpy()
        k1, k2 = len(nbs1), len(nbs2)
        # einsum letters: outer legs, the shared bond, physical legs
        o1 = [chr(ord("a") + i) for i in range(k1)]
        o2 = [chr(ord("a") + k1 + i) for i in range(k2)]
        s1 = [None] * a1.nneighbours()
        for i, nb in enumerate(nbs1):
            s1[a1.neighbour_index(nb)] = o1[i]
        s1[a1.neighbour_index(n2)] = "z"
        s2 = [None] * a2.nneighbours()
        for i, nb in enumerate(nbs2):
            s2[a2.neighbour_index(nb)] = o2[i]
        s2[a2.neighbour_index(n1)] = "z"
        spec = "".join(s1) + "QP," + "".join(s2) + "RS->" + "".join(o1 + o2) + "QRPS"
        W12 = np.einsum(spec, W1, W2)
        dense = torch.as_tensor(W12).to(device=t1.device, dtype=dtype)
        hit = BlockSparse.from_dense(dense, range(k1 + k2))
        hit.source = (t1, t2)
        self._blocks[key] = hit
        return hit

    def tensor_in(self, node_id: str, dtype: torch.dtype) -> torch.Tensor:
        """The node tensor in ``dtype`` (complex64 copies are cached)."""
        t = self._tensors[node_id]
        if dtype == t.dtype:
            return t
        hit = self._lowp.get(node_id)
        if hit is None or hit.shape != t.shape:
            hit = t.to(dtype)
            self._lowp[node_id] = hit
        return hit

    @classmethod
    def from_tensors(cls, structure: TreeTensorNetwork, tensors: Dict[str, np.ndarray],
                     children: Optional[Dict[str, Sequence[str]]] = None
                     ) -> "TreeTensorNetworkOperator":
        """TTNO on ``structure``'s tree (same nodes and parents).

        Each tensor has legs (parent, children..., out, in) with the children in
        the operator's own order: ``children[nid]`` when given, else the order of
        ``structure``.  The reference allows a TTNO whose children order differs
        from the state's and matches legs by neighbour identifier
        (``contractions/contraction_util.py:40-80``); every contraction here does
        the same.  Raises ValueError when a node's children are not the
        structure's, a tensor has the wrong number of legs, a bond has different
        dimensions at its two ends, or the out/in legs do not match the state's
        physical dimension."""
        op = cls()
        dev = structure.device
        for nid in _parents_first(structure):
            snode = structure.nodes[nid]
            n = Node(identifier=nid)
            kids = list(snode.children) if children is None or nid not in children \
                else list(children[nid])
            if sorted(kids) != sorted(snode.children):
                raise ValueError(f"operator children {kids} of {nid} are not the "
                                 f"state's children {list(snode.children)}")
            n.parent, n.children = snode.parent, kids
            t = as_device_tensor(tensors[nid], device=dev)
            n.shape = tuple(t.shape)
            op._nodes[nid] = n
            op._tensors[nid] = t
        op._root_id = structure.root_id
        op.check_structure(structure)
        return op

    def check_structure(self, state: Optional[TreeTensorNetwork] = None):
        """Check every node tensor's leg count and every bond's two dimensions, and,
        with ``state``, the out/in dimensions against the state's physical legs."""
        for nid, n in self._nodes.items():
            shape = tuple(self._tensors[nid].shape)
            if len(shape) != n.nneighbours() + 2:
                raise ValueError(f"operator tensor {nid} has {len(shape)} legs; expected "
                                 f"{n.nneighbours()} virtual legs plus out and in")
            for c in n.children:
                dc = self._tensors[c].shape[0]
                if shape[n.neighbour_index(c)] != dc:
                    raise ValueError(f"bond {nid}-{c} of the operator has dimension "
                                     f"{shape[n.neighbour_index(c)]} at {nid} and {dc} at {c}")
            if state is not None:
                phys = state.nodes[nid].shape[state.nodes[nid].nneighbours():]
                if tuple(shape[-2:]) != (int(np.prod(phys)),) * 2:
                    raise ValueError(f"operator tensor {nid} acts on {shape[-2:]}; the state "
                                     f"node has physical legs {tuple(phys)}")

    def with_children_order(self, children: Dict[str, Sequence[str]]
                            ) -> "TreeTensorNetworkOperator":
        """The same operator with the children of the listed nodes reordered (legs
        permuted to match); the other nodes keep their order."""
        op = type(self)()
        for nid, n in self._nodes.items():
            m = Node(identifier=nid)
            kids = list(children.get(nid, n.children))
            if sorted(kids) != sorted(n.children):
                raise ValueError(f"{kids} is not a reordering of {nid}'s children")
            m.parent, m.children = n.parent, kids
            np_ = n.nparents()
            perm = list(range(np_)) + [np_ + n.children.index(c) for c in kids] + \
                [len(n.shape) - 2, len(n.shape) - 1]
            t = self._tensors[nid].permute(perm).contiguous()
            m.shape = tuple(t.shape)
            op._nodes[nid] = m
            op._tensors[nid] = t
        op._root_id = self._root_id
        return op

    @classmethod
    def from_edge_terms(cls, structure: TreeTensorNetwork,
                        site_terms: Optional[Dict[str, np.ndarray]] = None,
                        edge_terms: Optional[Iterable[tuple]] = None,
                        phys_dim: int = 2) -> "TreeTensorNetworkOperator":
        """H = sum_i site_terms[i] + sum (coef * A_{n1} B_{n2}) over edge_terms.

        ``edge_terms`` holds (n1, n2, coef, A, B) with n1, n2 adjacent in the tree.
        """
        return cls.from_tensors(structure, build_edge_ttno_tensors(structure, site_terms,
                                                                   edge_terms, phys_dim))


TTNO = TreeTensorNetworkOperator


def _parents_first(structure: TreeTensorNetwork) -> List[str]:
    order, stack = [], [structure.root_id]
    while stack:
        nid = stack.pop()
        order.append(nid)
        stack.extend(reversed(structure.nodes[nid].children))
    return order


def build_edge_ttno_tensors(structure: TreeTensorNetwork,
                            site_terms: Optional[Dict[str, np.ndarray]],
                            edge_terms: Optional[Iterable[tuple]],
                            phys_dim: int = 2) -> Dict[str, np.ndarray]:
    site_terms = dict(site_terms or {})
    per_edge: Dict[Tuple[str, str], List[Tuple[complex, np.ndarray, np.ndarray]]] = {}
    for n1, n2, coef, a, b in (edge_terms or []):
        a = np.asarray(a, dtype=complex)
        b = np.asarray(b, dtype=complex)
        if structure.nodes[n1].parent == n2:
            child, parent, op_c, op_p = n1, n2, a, b
        elif structure.nodes[n2].parent == n1:
            child, parent, op_c, op_p = n2, n1, b, a
        else:
            raise ValueError(f"{n1} and {n2} are not adjacent")
        per_edge.setdefault((child, parent), []).append((complex(coef), op_c, op_p))
    eye = np.eye(phys_dim, dtype=complex)
    out: Dict[str, np.ndarray] = {}
    for nid, node in structure.nodes.items():
        kids = list(node.children)
        has_parent = node.parent is not None
        dims = ([2 + len(per_edge.get((nid, node.parent), []))] if has_parent else [])
        dims += [2 + len(per_edge.get((c, nid), [])) for c in kids]
        W = np.zeros(dims + [phys_dim, phys_dim], dtype=complex)
        nk = len(kids)

        def put(parent_state: Optional[int], child_states: Sequence[int], op: np.ndarray):
            idx = ([parent_state] if has_parent else []) + list(child_states)
            W[tuple(idx)] += op

        zeros = [0] * nk
        # parent state I: identity everywhere below
        if has_parent:
            put(0, zeros, eye)
        # parent state F (or the root's only slice)
        f_state = 1 if has_parent else None
        for i in range(nk):
            cs = list(zeros)
            cs[i] = 1
            put(f_state, cs, eye)
            for j, (coef, _op_c, op_p) in enumerate(per_edge.get((kids[i], nid), [])):
                cs = list(zeros)
                cs[i] = 2 + j
                put(f_state, cs, coef * op_p)
        if nid in site_terms:
            put(f_state, zeros, np.asarray(site_terms[nid], dtype=complex))
        # parent state 2+j: this node holds the child-side operator of edge term j
        if has_parent:
            for j, (_coef, op_c, _op_p) in enumerate(per_edge.get((nid, node.parent), [])):
                put(2 + j, zeros, op_c)
        out[nid] = W
    return out


# ---- benchmark models (reference sign convention) -------------------------------------
def _edges(structure: TreeTensorNetwork) -> List[Tuple[str, str]]:
    return [(n.parent, nid) for nid, n in structure.nodes.items() if n.parent is not None]


def transverse_field_ising(structure: TreeTensorNetwork, coupling=1.0, field=0.5,
                           edge_couplings: Optional[Dict[Tuple[str, str], float]] = None
                           ) -> TreeTensorNetworkOperator:
    """H = -sum_e J_e X X - g sum_i Z_i."""
    X, _, Z = pauli_matrices()
    terms = []
    for p, c in _edges(structure):
        j = coupling if edge_couplings is None else edge_couplings[(p, c)]
        terms.append((p, c, -j, X, X))
    site = {nid: -field * Z for nid in structure.nodes}
    return TreeTensorNetworkOperator.from_edge_terms(structure, site, terms)


def xyz_model(structure: TreeTensorNetwork, couplings: Dict[Tuple[str, str], Sequence[float]],
              fields: Optional[Dict[str, float]] = None) -> TreeTensorNetworkOperator:
    """H = -sum_e (Jx_e XX + Jy_e YY + Jz_e ZZ) - sum_i h_i Z_i."""
    X, Y, Z = pauli_matrices()
    terms = []
    for p, c in _edges(structure):
        jx, jy, jz = couplings[(p, c)]
        terms += [(p, c, -jx, X, X), (p, c, -jy, Y, Y), (p, c, -jz, Z, Z)]
    site = {nid: -h * Z for nid, h in (fields or {}).items()}
    return TreeTensorNetworkOperator.from_edge_terms(structure, site, terms)


def xxz_model(structure: TreeTensorNetwork, jxy=1.0, jz=0.8, field=0.0):
    return xyz_model(structure, {e: (jxy, jxy, jz) for e in _edges(structure)},
                     {nid: field for nid in structure.nodes} if field else None)


def random_xyz_model(structure: TreeTensorNetwork, seed: int) -> TreeTensorNetworkOperator:
    """J_e^{x,y,z} ~ U[-1, 1] per edge and axis, edges in tree insertion order."""
    rng = np.random.default_rng(seed)
    return xyz_model(structure, {e: tuple(rng.uniform(-1, 1, 3)) for e in _edges(structure)})


def random_field_heisenberg(structure: TreeTensorNetwork, disorder: float, seed: int):
    """H = -sum_e (XX + YY + ZZ) - sum_i h_i Z_i, h_i ~ U[-W, W] in node insertion order."""
    rng = np.random.default_rng(seed)
    fields = {nid: float(rng.uniform(-disorder, disorder)) for nid in structure.nodes}
    return xyz_model(structure, {e: (1.0, 1.0, 1.0) for e in _edges(structure)}, fields)
