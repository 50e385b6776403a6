"""Sandwich contractions in the reference contraction order (oracle).

Leg conventions (reference ``core/node.py:35``,
``contractions/state_operator_contraction.py:1042-1093``):

* state tensor  A_n : (parent, child_0, ..., child_{c-1}, phys)  -- the root has
  no parent leg;
* operator      W_n : (parent, children..., out, in); ``out`` meets the bra,
  ``in`` meets the ket;
* environment   E[(n, m)] : (ket, ham, bra), the subtree rooted at ``n`` with
  open legs pointing to ``m``.

Every function below follows the order of ``np.tensordot`` calls the reference
makes, so the CPU timing reflects the reference path:

* ``env_update``       -- ``contract_any`` ket-first branch
  (``state_operator_contraction.py:407-461,533-694``; ``contraction_util.py:176-200``).
  ``compare_contr_orders`` (``:142-162``) picks the operator-first branch when
  (d-1)*prod(D)*prod(chi) < d*(prod(D)-prod(chi)).  For d = 2 that happens only
  when the ket node's virtual dimensions multiply to 1 and prod(D) >= 3 (any
  virtual product p >= 2 gives prod(D)*p >= 2 prod(D) > 2(prod(D) - p)), i.e.
  for product-state tensors with all bonds of dimension 1.  The branch changes
  only the summation order (rounding-level differences), so it is not restated;
  every tensor of the benchmark and test states has bonds of dimension >= 2.
* Operator legs: the TTNO's children order may differ from the state's.  The
  reference finds the operator leg of each ket neighbour by identifier
  (``contraction_util.py:40-80``, ``get_equivalent_legs``); here every function
  takes an optional ``op_topo`` (the operator's own tree, same nodes and
  parents, its own children order; default: the state's) and looks the W leg
  of a ket neighbour up in it (``_wleg``).
* ``heff_apply``       -- ``contract_ket_ham_with_envs`` (``:227-273``) via
  ``contract_all_neighbour_blocks_to_ket`` (``contraction_util.py:202-290``).
* ``bond_apply``       -- ``contract_bond_tensor`` (``:984-1040``).
* ``expectation_value``-- ``expectation_value`` (``:26-74``), unnormalised.
* ``overlap``          -- ``contract_two_ttns`` (``state_state_contraction.py:29-66``).
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np

from .tree import Topology

Env = Dict[Tuple[str, str], np.ndarray]


def _phys_leg(tensor: np.ndarray) -> int:
    return tensor.ndim - 1


def _wleg(topo: Topology, op_topo, nid: str, nb: str) -> int:
    """Leg of the operator tensor at ``nid`` that meets neighbour ``nb``."""
    return (op_topo or topo).neighbour_index(nid, nb)


def env_leaf(A: np.ndarray, W: np.ndarray, B: np.ndarray) -> np.ndarray:
    """contract_leaf: A (p, phys), W (p, out, in), B = conj(A) (p, phys) -> (ket, ham, bra)."""
    bra_ham = np.tensordot(W, B, axes=([1], [1]))          # (w, in, b)
    return np.tensordot(A, bra_ham, axes=([1], [1]))       # (a, w, b)


def env_update(topo: Topology, nid: str, to: str, A: np.ndarray, W: np.ndarray,
               envs: Env, B: np.ndarray = None, op_topo: Topology = None) -> np.ndarray:
    """Directed environment E[(nid, to)] (ket-first reference order)."""
    if B is None:
        B = A.conj()
    nbs = topo.neighbours(nid)
    if len(nbs) == 1:
        return env_leaf(A, W, B)
    ignore = nbs.index(to)
    # contract_all_but_one_neighbour_block_to_ket: the leg to ``to`` stays; each
    # contracted leg disappears and (ham, bra) are appended at the end.
    t = A
    for k, nb in enumerate(nbs):
        if nb == to:
            continue
        leg = 1 if ignore < k else 0
        t = np.tensordot(t, envs[(nb, nid)], axes=([leg], [0]))
    # t legs: (a_to, phys, w_1, b_1, ..., w_j, b_j) in neighbour order without ``to``
    # contract_operator_tensor_ignoring_one_leg: W legs (nbs..., out, in)
    nrest = len(nbs) - 1
    t_legs = [2 + 2 * i for i in range(nrest)] + [1]
    w_legs = [_wleg(topo, op_topo, nid, nb) for nb in nbs if nb != to] + [len(nbs) + 1]
    t = np.tensordot(t, W, axes=(t_legs, w_legs))
    # t legs: (a_to, b_1..b_j, w_to, out)
    # contract_bra_tensor_ignore_one_leg: B legs (nbs..., phys)
    t_legs = list(range(1, 1 + nrest)) + [nrest + 2]
    b_legs = [i for i in range(len(nbs)) if i != ignore] + [len(nbs)]
    return np.tensordot(t, B, axes=(t_legs, b_legs))        # (a_to, w_to, b_to)


def heff_apply(topo: Topology, nid: str, x: np.ndarray, W: np.ndarray,
               envs: Env, op_topo: Topology = None) -> np.ndarray:
    """y = H_eff x at node ``nid`` (reference order: all envs into the ket, then W)."""
    t = x
    nbs = topo.neighbours(nid)
    for nb in nbs:
        t = np.tensordot(t, envs[(nb, nid)], axes=([0], [0]))
    # t legs: (phys, w_1, b_1, ..., w_k, b_k)
    k = len(nbs)
    t_legs = [1 + 2 * i for i in range(k)] + [0]
    w_legs = [_wleg(topo, op_topo, nid, nb) for nb in nbs] + [k + 1]
    return np.tensordot(t, W, axes=(t_legs, w_legs))        # (b_1..b_k, out)


def bond_apply(x: np.ndarray, env_p: np.ndarray, env_c: np.ndarray) -> np.ndarray:
    """Link action: x (a_p, a_c), env_p = E[(p, c)], env_c = E[(c, p)] -> (b_p, b_c)."""
    t = np.tensordot(x, env_p, axes=([0], [0]))             # (a_c, w, b_p)
    t = np.tensordot(t, env_c, axes=([0], [0]))             # (w, b_p, w', b_c)
    return np.einsum("ijil->jl", t)


def two_site_heff_apply(topo: Topology, n1: str, n2: str, x: np.ndarray,
                        W1: np.ndarray, W2: np.ndarray, envs: Env,
                        op_topo: Topology = None) -> np.ndarray:
    """Two-site action on x with legs (outer neighbours of n1, outer of n2, p1, p2).

    The outer neighbours are listed in each node's neighbour order, skipping the
    shared bond (the layout ``contract_nodes`` produces up to a leg permutation,
    ``core/ttn.py:1072-1115``); the TTNO tensors are merged over their shared bond
    (``contractions/node_contraction.py:12-59``).
    """
    nb1 = [n for n in topo.neighbours(n1) if n != n2]
    nb2 = [n for n in topo.neighbours(n2) if n != n1]
    otopo = op_topo or topo
    i12 = otopo.neighbour_index(n1, n2)
    i21 = otopo.neighbour_index(n2, n1)
    # merged operator legs: (outer1..., out1, in1, outer2..., out2, in2) with the
    # outer legs in the operator's own neighbour order
    Wm = np.tensordot(W1, W2, axes=([i12], [i21]))
    k1, k2 = len(nb1), len(nb2)
    onb1 = [n for n in otopo.neighbours(n1) if n != n2]
    onb2 = [n for n in otopo.neighbours(n2) if n != n1]
    # reorder to (outer1, outer2 in the state's neighbour order, out1, out2, in1, in2)
    perm = [onb1.index(n) for n in nb1] + [k1 + 2 + onb2.index(n) for n in nb2] + \
        [k1, k1 + 2 + k2, k1 + 1, k1 + 3 + k2]
    Wm = Wm.transpose(perm)
    t = x
    for nb, owner in [(n, n1) for n in nb1] + [(n, n2) for n in nb2]:
        t = np.tensordot(t, envs[(nb, owner)], axes=([0], [0]))
    k = k1 + k2
    # t legs: (p1, p2, w_1, b_1, ...)
    t_legs = [2 + 2 * i for i in range(k)] + [0, 1]
    w_legs = list(range(k)) + [k + 2, k + 3]
    return np.tensordot(t, Wm, axes=(t_legs, w_legs))       # (b..., out1, out2)


def build_envs_to_root(topo: Topology, A: dict, W: dict, op_topo: Topology = None) -> Env:
    envs: Env = {}
    for nid in topo.linearise()[:-1]:
        p = topo.parent[nid]
        envs[(nid, p)] = env_update(topo, nid, p, A[nid], W[nid], envs, op_topo=op_topo)
    return envs


def expectation_value(topo: Topology, A: dict, W: dict, op_topo: Topology = None) -> complex:
    """<psi|H|psi> (unnormalised), children-first, root contracted last."""
    envs = build_envs_to_root(topo, A, W, op_topo)
    r = topo.root
    y = heff_apply(topo, r, A[r], W[r], envs, op_topo=op_topo)
    return complex(np.vdot(A[r], y))


def _overlap_env(topo: Topology, nid: str, to: str, ket: np.ndarray,
                 bra: np.ndarray, envs: Env) -> np.ndarray:
    """Two-leg environment (ket, bra) of the subtree at ``nid`` pointing to ``to``."""
    nbs = topo.neighbours(nid)
    t = ket
    ignore = nbs.index(to)
    for k, nb in enumerate(nbs):
        if nb == to:
            continue
        leg = 1 if ignore < k else 0
        t = np.tensordot(t, envs[(nb, nid)], axes=([leg], [0]))
    # t legs: (a_to, phys, b_1, ..., b_j)
    nrest = len(nbs) - 1
    t_legs = list(range(2, 2 + nrest)) + [1]
    b_legs = [i for i in range(len(nbs)) if i != ignore] + [len(nbs)]
    return np.tensordot(t, bra, axes=(t_legs, b_legs))


def overlap(topo: Topology, bra: dict, ket: dict) -> complex:
    """<bra|ket> with ``bra`` conjugated."""
    envs: Env = {}
    for nid in topo.linearise()[:-1]:
        p = topo.parent[nid]
        envs[(nid, p)] = _overlap_env(topo, nid, p, ket[nid], bra[nid].conj(), envs)
    r = topo.root
    t = ket[r]
    k = len(topo.neighbours(r))
    for nb in topo.neighbours(r):
        t = np.tensordot(t, envs[(nb, r)], axes=([0], [0]))
    # t legs: (phys, b_1, ..., b_k)
    return complex(np.tensordot(t, bra[r].conj(),
                                axes=(list(range(1, k + 1)) + [0], list(range(k + 1)))))


def single_site_expectation(topo: Topology, A: dict, nid: str, op: np.ndarray) -> complex:
    """<psi| op_nid |psi> (unnormalised) via an overlap with the operator applied to the ket."""
    ket = dict(A)
    ket[nid] = np.tensordot(A[nid], op, axes=([A[nid].ndim - 1], [1]))
    return overlap(topo, A, ket)
