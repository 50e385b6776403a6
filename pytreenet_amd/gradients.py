"""Gradients of the Rayleigh quotient E = <psi|H|psi> / <psi|psi> on the device.

The reference has no gradients.  This is the hand-written adjoint of its
sandwich contraction (``contractions/state_operator_contraction.py:26-74``):

* site tensors:   dE/d conj(A_n) = (H_eff,n A_n - E N_eff,n A_n) / <psi|psi>,
  where H_eff,n is the action of ``contract_ket_ham_with_envs`` at n and N_eff,n
  the same network with the identity operator (norm environments).  PyTorch's
  convention for a real loss of complex inputs is grad = 2 dL/d conj(z).
* TTNO tensors:   dE/dW_n = G_n / <psi|psi>, G_n the network {A_n, conj(A_n),
  all environment blocks into n} with W_n's legs left open (the "environment of
  the operator tensor").
* couplings:      the TTNO is linear in each coupling J_e, W_n = W_n^0 +
  sum_e J_e B_{n,e}, so dE/dJ_e = sum_n Re <G_n, B_{n,e}> / <psi|psi>.

Every directed environment is needed once: an upward pass (children first)
and a downward pass (parents first) give all 2(N-1) blocks of both the
Hamiltonian and the norm sandwich.  Environment checkpointing is recompute
instead of store: ``RayleighQuotient.forward`` keeps no block (it evaluates E
with the upward pass only), ``backward`` recomputes both passes.
"""
from __future__ import annotations

import math

from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import engine as E
from .contractions import (PartialTreeCachDict, _env_labels, _op_labels, caching_path,
                           contract_any, contract_ket_ham_with_envs, operator_tensor)
from .ttn import TreeTensorNetworkState
from .ttno import TreeTensorNetworkOperator, build_edge_ttno_tensors


def _norm_block(state, nid: str, to_idx: Optional[int], cache) -> torch.Tensor:
    node, ket = state[nid]
    k = node.nneighbours()
    tensors = [ket, ket]
    labels = [[("a", i) for i in range(k)] + ["p"], [("b", i) for i in range(k)] + ["p"]]
    conj = [False, True]
    for i, nb in enumerate(node.neighbouring_nodes()):
        if i == to_idx:
            continue
        tensors.append(cache[(nb, nid)])
        labels.append([("a", i), ("b", i)])
        conj.append(False)
    out = [] if to_idx is None else [("a", to_idx), ("b", to_idx)]
    return E.contract_network(tensors, labels, out, conj=conj)


def all_environments(state: TreeTensorNetworkState, ttno: TreeTensorNetworkOperator,
                     precision: str = E.FP64):
    """Both directions of every Hamiltonian and norm environment block."""
    hcache = PartialTreeCachDict()
    ncache: Dict[Tuple[str, str], torch.Tensor] = {}
    root = state.root_id
    for nid, to in caching_path(state, root):
        hcache.add_entry(nid, to, contract_any(nid, to, state, ttno, hcache,
                                               precision=precision))
        ncache[(nid, to)] = _norm_block(state, nid, state.nodes[nid].neighbour_index(to),
                                        ncache)
    order, stack = [], [root]
    while stack:
        nid = stack.pop()
        order.append(nid)
        stack.extend(reversed(state.nodes[nid].children))
    for nid in order:
        node = state.nodes[nid]
        for i, c in enumerate(node.children):
            hcache.add_entry(nid, c, contract_any(nid, c, state, ttno, hcache,
                                                  precision=precision))
            ncache[(nid, c)] = _norm_block(state, nid, node.neighbour_index(c), ncache)
    return hcache, ncache


def norm_apply(state, nid: str, x: torch.Tensor, ncache) -> torch.Tensor:
    """N_eff x at node ``nid`` (identity operator sandwich); output in the ket layout."""
    node = state.nodes[nid]
    k = node.nneighbours()
    tensors = [x]
    labels = [[("a", i) for i in range(k)] + ["p"]]
    for i, nb in enumerate(node.neighbouring_nodes()):
        tensors.append(ncache[(nb, nid)])
        labels.append([("a", i), ("b", i)])
    return E.contract_network(tensors, labels, [("b", i) for i in range(k)] + ["p"])


def operator_environment(state, ttno, nid: str, hcache) -> torch.Tensor:
    """G_n = dN/dW_n: {A_n, conj(A_n), all blocks into n} with W_n's legs open."""
    node, ket = state[nid]
    op_node = ttno.nodes[nid]
    k = node.nneighbours()
    tensors = [ket, ket]
    labels = [[("a", i) for i in range(k)] + ["p"], [("b", i) for i in range(k)] + ["q"]]
    conj = [False, True]
    for i, nb in enumerate(node.neighbouring_nodes()):
        tensors.append(hcache.get_entry(nb, nid))
        labels.append(list(_env_labels(i)))
        conj.append(False)
    return E.contract_network(tensors, labels, _op_labels(node, op_node), conj=conj)


def rayleigh_quotient_gradients(state: TreeTensorNetworkState, ttno: TreeTensorNetworkOperator,
                                precision: str = E.FP64, with_operator: bool = True):
    """(E, {node: dE/d conj(A_n)}, {node: dE/dW_n} or None, <psi|psi>)."""
    hcache, ncache = all_environments(state, ttno, precision)
    r = state.root_id
    node, A = state[r]
    hy = contract_ket_ham_with_envs(node, A, ttno.nodes[r], operator_tensor(ttno, r, precision),
                                    hcache, precision=precision)
    from .krylov import dotc
    num = dotc(A, hy).real
    den = dotc(A, norm_apply(state, r, A, ncache)).real
    if not (math.isfinite(num) and math.isfinite(den)) or den <= 0.0:
        # complex64 environments of an unnormalised state overflow (the subtree
        # norms of random tensors grow geometrically with the subtree size)
        raise FloatingPointError(
            f"<psi|H|psi> = {num}, <psi|psi> = {den} in precision {precision!r}: "
            "bring the state to canonical form (environments of order one) or use fp64")
    energy = num / den
    gsite: Dict[str, torch.Tensor] = {}
    gop: Optional[Dict[str, torch.Tensor]] = {} if with_operator else None
    for nid in state.nodes:
        nd, An = state[nid]
        h = contract_ket_ham_with_envs(nd, An, ttno.nodes[nid],
                                       operator_tensor(ttno, nid, precision), hcache,
                                       precision=precision)
        nrm = norm_apply(state, nid, An, ncache)
        from .krylov import lincomb
        gsite[nid] = lincomb([h, nrm], [1.0 / den, -energy / den])
        if with_operator:
            gop[nid] = lincomb([operator_environment(state, ttno, nid, hcache)], [1.0 / den])
    return energy, gsite, gop, den


class CouplingTTNO:
    """A TTNO that is linear in a coupling vector J: W_n = W_n^0 + sum_e J_e B_{n,e}.

    ``builder(J)`` must return (site_terms, edge_terms) for ``from_edge_terms``
    with every coefficient linear in J (the benchmark models are).
    """

    def __init__(self, structure: TreeTensorNetworkState,
                 builder: Callable[[np.ndarray], tuple], n_params: int):
        self.structure = structure
        self.builder = builder
        self.n = n_params
        base = build_edge_ttno_tensors(structure, *builder(np.zeros(n_params)))
        self.base = base
        self.basis: List[Dict[str, np.ndarray]] = []
        for e in range(n_params):
            J = np.zeros(n_params)
            J[e] = 1.0
            We = build_edge_ttno_tensors(structure, *builder(J))
            self.basis.append({k: We[k] - base[k] for k in base})

    def ttno(self, J: np.ndarray) -> TreeTensorNetworkOperator:
        W = build_edge_ttno_tensors(self.structure, *self.builder(np.asarray(J, dtype=float)))
        return TreeTensorNetworkOperator.from_tensors(self.structure, W)

    def coupling_gradient(self, gop: Dict[str, torch.Tensor]) -> np.ndarray:
        host = {k: v.detach().cpu().numpy() for k, v in gop.items()}
        out = np.zeros(self.n)
        for e, B in enumerate(self.basis):
            out[e] = sum(float(np.real(np.vdot(B[k].conj(), host[k]))) for k in B
                         if np.any(B[k]))
        return out


def tfim_coupling_builder(structure, field: float = 0.5):
    """H = -sum_e J_e X X - g sum Z with one coupling per tree edge (insertion order)."""
    from .ttno import pauli_matrices
    X, _, Z = pauli_matrices()
    edges = [(n.parent, nid) for nid, n in structure.nodes.items() if n.parent is not None]

    def build(J):
        return ({nid: -field * Z for nid in structure.nodes},
                [(p, c, -float(J[i]), X, X) for i, (p, c) in enumerate(edges)])

    return build, len(edges)


class RayleighQuotient(torch.autograd.Function):
    """E(A_1..A_N, J) with a hand-written backward (adjoint sandwich contraction).

    apply(model, state_template, J, *site_tensors) -> E (real, 0-dim tensor).
    Gradients flow to every site tensor (complex) and to J (real, float64).
    """

    @staticmethod
    def forward(ctx, model: CouplingTTNO, template: TreeTensorNetworkState,
                J: torch.Tensor, *tensors):
        st = template.copy()
        for nid, t in zip(st.nodes, tensors):
            st.tensors[nid] = t.detach()
        H = model.ttno(J.detach().cpu().numpy())
        from .contractions import expectation_value
        num = expectation_value(st, H).real
        den = st.scalar_product(use_orthogonal_center=False).real
        ctx.model, ctx.st, ctx.H = model, st, H
        return torch.tensor(num / den, dtype=torch.float64, device=tensors[0].device)

    @staticmethod
    def backward(ctx, grad_out):
        energy, gsite, gop, den = rayleigh_quotient_gradients(ctx.st, ctx.H)
        go = float(grad_out.item())
        grads = [2.0 * go * gsite[nid] for nid in ctx.st.nodes]
        gJ = torch.as_tensor(go * ctx.model.coupling_gradient(gop), dtype=torch.float64)
        return (None, None, gJ.to(grad_out.device), *grads)


def ground_state_descent(state: TreeTensorNetworkState, ttno: TreeTensorNetworkOperator,
                         steps: int = 200, lr: float = 0.05, precision: str = E.FP64
                         ) -> List[float]:
    """Plain gradient descent of the Rayleigh quotient over all site tensors.

    Each step computes every gradient with one environment sweep pair and updates
    A_n <- A_n - lr * dE/d conj(A_n) in place (the quotient is scale invariant;
    the state is rescaled at its root so norms stay O(1)).  Returns E per step.
    """
    from .krylov import lincomb
    hist = []
    for _ in range(steps):
        energy, gsite, _, den = rayleigh_quotient_gradients(state, ttno, precision,
                                                            with_operator=False)
        hist.append(energy)
        for nid, g in gsite.items():
            state.tensors[nid] = lincomb([state.tensors[nid], g], [1.0, -lr])
        r = state.root_id
        state.tensors[r] = lincomb([state.tensors[r]], [float(den) ** -0.5])
    return hist
