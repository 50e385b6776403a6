"""Basis-update & Galerkin integrators (oracle, test infrastructure only).

Restates, on an ``OracleState`` with fixed leg layout (parent, children..., phys):

* ``bug_step``          -- ``root_update`` / ``update_node`` / ``update_leaf_node`` /
  ``update_non_leaf_node`` (``time_evolution/time_evo_util/common_bug.py:47-378``)
  with ``compute_new_basis_tensor``, ``compute_fixed_size_new_basis_tensor`` and
  ``compute_basis_change_tensor`` (``bug_util.py:78-170``).  The centre moves
  from the parent by QR (REDUCED, KEEP for the fixed rank form); leaves stack
  their old basis tensor on the evolved one, inner nodes the coefficient tensor
  with the children's basis changes contracted in.
* ``leaf_to_leaf_path`` -- ``TDVPUpdatePathFinder_LeafToLeaf``
  (``time_evo_util/update_path.py:256-398``).
* ``svd_truncation``    -- ``core/truncation/svd_truncation.py:22-81`` (one-site
  form), followed in ``BUG.truncation`` (``bug.py:100-108``) by a KEEP move of the
  centre back to the root.

The local evolutions use the exact exponential of the dense effective
Hamiltonian (``expm`` of H_eff built column by column), the dense mode of the
reference's ``TimeEvoMode``.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np
from scipy.linalg import expm

from . import contract as C
from .splitting import qr_split, svd_split
from .tdvp import OracleState
from .tree import Topology


def _move_to(t: np.ndarray, src: int, dst: int) -> np.ndarray:
    return np.moveaxis(t, src, dst)


def _absorb(t: np.ndarray, leg: int, r: np.ndarray) -> np.ndarray:
    """t[.., new, ..] = sum_old r[new, old] t[.., old, ..] on ``leg``."""
    return _move_to(np.tensordot(r, t, axes=([1], [leg])), 0, leg)


def _qr_off(t: np.ndarray, leg: int, keep: bool) -> Tuple[np.ndarray, np.ndarray]:
    """Q (new bond at ``leg``) and R (new, old) of t split off ``leg``."""
    others = [i for i in range(t.ndim) if i != leg]
    q, r = qr_split(t, others, [leg], keep=keep)
    return _move_to(q, q.ndim - 1, leg), r


def heff_matrix(topo: Topology, nid: str, W: np.ndarray, envs: C.Env,
                shape) -> np.ndarray:
    n = int(np.prod(shape))
    cols = []
    for j in range(n):
        e = np.zeros(n, dtype=complex)
        e[j] = 1.0
        cols.append(C.heff_apply(topo, nid, e.reshape(shape), W, envs).reshape(-1))
    return np.stack(cols, axis=1)


def _evolve(topo, nid, x, W, envs, dt):
    H = heff_matrix(topo, nid, W, envs, x.shape)
    return (expm(-1j * dt * H) @ x.reshape(-1)).reshape(x.shape)


def bug_step(state: OracleState, W: Dict[str, np.ndarray], dt: float,
             fixed_rank: bool = False) -> OracleState:
    topo = state.topo
    root = topo.root
    assert state.center == root
    A = state.A
    up: C.Env = {}
    order: List[str] = []

    def post(n):
        for c in topo.children[n]:
            post(c)
        order.append(n)

    post(root)
    for n in order:
        if n != root:
            up[(n, topo.parent[n])] = C.env_update(topo, n, topo.parent[n], A[n], W[n], up)
    new: Dict[str, np.ndarray] = {}

    def with_changes(n, t, M):
        for c in topo.children[n]:
            t = _absorb(t, topo.neighbour_index(n, c), M[c].T)
        return t

    def node(n: str, p: str, C_p: np.ndarray, into_p: C.Env):
        q, r = _qr_off(C_p, topo.neighbour_index(p, n), keep=fixed_rank)
        C_n = _absorb(A[n], 0, r)
        E_pn = C.env_update(topo, p, n, q, W[p], into_p)
        into_n = {(p, n): E_pn}
        for c in topo.children[n]:
            into_n[(c, n)] = up[(c, n)]
        M, newenv = {}, {}
        for c in topo.children[n]:
            M[c], newenv[c] = node(c, n, C_n, into_n)
        if topo.children[n]:
            C_n = with_changes(n, C_n, M)
            old = C_n
        else:
            old = A[n]
        envs = {(p, n): E_pn}
        envs.update({(c, n): newenv[c] for c in topo.children[n]})
        evolved = _evolve(topo, n, C_n, W[n], envs, dt)
        if fixed_rank:
            comb, keep = evolved, True
        else:
            comb, keep = np.concatenate([old, evolved], axis=0), False
        qb, _ = qr_split(comb, list(range(1, comb.ndim)), [0], keep=keep)
        basis = _move_to(qb, qb.ndim - 1, 0)
        # M_n[a, b] = sum A[a, c.., s] M_c[c, c'] conj(basis[b, c'.., s])
        t = with_changes(n, A[n], M)
        legs = list(range(1, t.ndim))
        Mn = np.tensordot(t, basis.conj(), axes=(legs, legs))
        env_new = C.env_update(topo, n, p, basis, W[n],
                               {(c, n): newenv[c] for c in topo.children[n]})
        new[n] = basis
        return Mn, env_new

    into_root = {(c, root): up[(c, root)] for c in topo.children[root]}
    M, newenv = {}, {}
    for c in topo.children[root]:
        M[c], newenv[c] = node(c, root, A[root], into_root)
    C_root = with_changes(root, A[root], M)
    new[root] = _evolve(topo, root, C_root, W[root],
                        {(c, root): newenv[c] for c in topo.children[root]}, dt)
    out = OracleState(topo, new)
    out.center = root
    return out


def leaf_to_leaf_path(topo: Topology) -> List[str]:
    leaves = topo.get_leaves()
    best, pair = -1, (leaves[0], leaves[0])
    for i, a in enumerate(leaves):
        dist = topo.distance_to_node(a)
        for b in leaves[i + 1:]:
            if dist[b] > best:
                best, pair = dist[b], (a, b)
    main = topo.path_from_to(*pair)
    visited = set()

    def rec(n):
        if n in visited or n in main:
            return []
        visited.add(n)
        out = []
        p = topo.parent[n]
        if p is not None and p not in main and p not in visited:
            out += rec(p)
        for c in topo.children[n]:
            if c not in main and c not in visited:
                out += rec(c)
        return out + [n]

    path = []
    for o in main:
        p = topo.parent[o]
        if p is not None and p not in main and p not in visited:
            path += rec(p)
        for c in topo.children[o]:
            if c not in main:
                path += rec(c)
        if o not in visited:
            visited.add(o)
            path.append(o)
    return path


def move_center(state: OracleState, target: str, keep: bool):
    topo = state.topo
    path = topo.path_from_to(state.center, target)
    for a, b in zip(path[:-1], path[1:]):
        q, r = _qr_off(state.A[a], topo.neighbour_index(a, b), keep=keep)
        state.A[a] = q
        state.A[b] = _absorb(state.A[b], topo.neighbour_index(b, a), r)
    state.center = target


def svd_truncation(state: OracleState, svd_params: dict):
    topo = state.topo
    path = leaf_to_leaf_path(topo)
    move_center(state, path[0], keep=False)
    for n in path:
        if n == topo.root:
            continue
        move_center(state, n, keep=False)
        p = topo.parent[n]
        t = state.A[n]
        u, sv, _ = svd_split(t, list(range(1, t.ndim)), [0], **svd_params)
        state.A[n] = _move_to(u, u.ndim - 1, 0)
        state.A[p] = _absorb(state.A[p], topo.neighbour_index(p, n), sv)
        state.center = p


def bug_time_step(state: OracleState, W, dt: float, svd_params: Optional[dict] = None,
                  fixed_rank: bool = False) -> OracleState:
    """One BUG (``bug.py:110-117``) or fixed-rank BUG (``fixed_bug.py:93-97``) step."""
    out = bug_step(state, W, dt, fixed_rank=fixed_rank)
    if not fixed_rank:
        svd_truncation(out, svd_params or {})
        move_center(out, out.topo.root, keep=True)
    return out
