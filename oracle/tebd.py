"""TEBD gate sweep (oracle, test infrastructure only).

Restates ``time_evolution/tebd.py:84-155`` on an ``OracleState`` with fixed leg
layout (parent, children..., phys):

* ``apply_single_site``  -- ``_apply_one_trotter_step_single_site`` (``:84-96``):
  the gate's input leg contracted with the node's physical leg
  (``TTN.absorb_into_open_legs``, ``core/ttn.py:929-964``).
* ``apply_two_site``     -- ``_apply_one_trotter_step_two_site`` (``:98-120``):
  contract the two nodes over their bond, contract the gate's input legs with
  the two physical legs (first identifier first), truncated SVD with S absorbed
  into the second node (``split_node_svd``, ``core/ttn.py:1476-1497``).
* ``run_one_time_step``  -- ``:148-155``: every gate once, in order.

``apply_gates_dense`` is the brute-force counterpart on a dense state vector:
without truncation the two agree exactly (the SVD split is exact).
"""
from __future__ import annotations

from typing import Dict, List, Sequence

import numpy as np

from .splitting import svd_split
from .tdvp import OracleState


def apply_single_site(state: OracleState, nid: str, gate: np.ndarray):
    t = state.A[nid]
    state.A[nid] = np.tensordot(t, gate, axes=([t.ndim - 1], [1]))


def apply_two_site(state: OracleState, n1: str, n2: str, gate: np.ndarray,
                   svd_params: Dict) -> np.ndarray:
    """Returns the kept singular values."""
    topo = state.topo
    A1, A2 = state.A[n1], state.A[n2]
    i12 = topo.neighbour_index(n1, n2)
    i21 = topo.neighbour_index(n2, n1)
    x = np.tensordot(A1, A2, axes=([i12], [i21]))
    k1, k2 = A1.ndim - 2, A2.ndim - 2
    # (outer1..., p1, outer2..., p2) -> (outer1..., outer2..., p1, p2)
    perm = list(range(k1)) + list(range(k1 + 1, k1 + 1 + k2)) + [k1, k1 + 1 + k2]
    x = np.transpose(x, perm)
    x = np.tensordot(x, gate, axes=([k1 + k2, k1 + k2 + 1], [2, 3]))
    u_legs = list(range(k1)) + [k1 + k2]
    v_legs = list(range(k1, k1 + k2)) + [k1 + k2 + 1]
    u, sv, s_kept = svd_split(x, u_legs, v_legs, **svd_params)
    outer1 = list(range(k1))
    state.A[n1] = np.transpose(u, outer1[:i12] + [k1 + 1] + outer1[i12:] + [k1])
    outer2 = list(range(1, 1 + k2))
    state.A[n2] = np.transpose(sv, outer2[:i21] + [0] + outer2[i21:] + [1 + k2])
    return s_kept


def run_one_time_step(state: OracleState, gates: Sequence, svd_params: Dict) -> List[int]:
    """gates: (node_ids, tensor) with legs (out..., in...).  Returns the kept
    bond dimension of every two-site gate."""
    kept = []
    for ids, g in gates:
        if len(ids) == 1:
            apply_single_site(state, ids[0], g)
        elif len(ids) == 2:
            kept.append(len(apply_two_site(state, ids[0], ids[1], g, svd_params)))
        elif len(ids) > 2:
            raise NotImplementedError("More than two-site interactions are not implemented.")
    state.center = None
    return kept


def apply_gates_dense(psi: np.ndarray, site_ids: Sequence[str], dims: Sequence[int],
                      gates: Sequence) -> np.ndarray:
    """Apply the gates to a dense state vector whose sites are ``site_ids`` (first
    most significant) with local dimensions ``dims``."""
    pos = {n: i for i, n in enumerate(site_ids)}
    t = psi.reshape(dims)
    for ids, g in gates:
        k = len(ids)
        axes = [pos[n] for n in ids]
        t = np.tensordot(g, t, axes=(list(range(k, 2 * k)), axes))
        # the gate's output legs come first: move them back to the sites' axes
        rest = [a for a in range(len(dims)) if a not in axes]
        order = [0] * len(dims)
        for j, a in enumerate(axes):
            order[a] = j
        for j, a in enumerate(rest):
            order[a] = k + j
        t = np.transpose(t, order)
    return t.reshape(-1)
