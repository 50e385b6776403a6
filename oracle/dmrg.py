"""DMRG sweeps (oracle, test infrastructure only).

Restates ``DMRGAlgorithm`` (``dmrg/dmrg.py:20-299``) on an ``OracleState``:
leaf-to-leaf update path (``update_path.py:256-398``), orthogonalisation paths
between consecutive path nodes, KEEP canonical form to the first node, the
environment cache pointing to it; one-site sweep (``:232-256``) and two-site
sweep (``:258-276``, merge, lowest eigenvector, truncated SVD with S into the
second node).  Every local problem is solved exactly: the lowest eigenpair of
the dense effective Hamiltonian by ``numpy.linalg.eigh``.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np

from . import contract as C
from .bug import heff_matrix, leaf_to_leaf_path
from .splitting import svd_split
from .tdvp import OracleState


class OracleDMRG:
    def __init__(self, state: OracleState, W: Dict[str, np.ndarray], svd_params=None):
        self.s = state.copy()
        self.topo = state.topo
        self.W = W
        self.svd_params = svd_params or {}
        self.update_path = leaf_to_leaf_path(self.topo)
        self.orth_path = [self.topo.path_from_to(a, b) for a, b in
                          zip(self.update_path[:-1], self.update_path[1:])]
        if self.s.center is None:
            self.s.canonical_form(self.update_path[0])
        else:
            self.s.move_center(self.update_path[0])
        self.envs: C.Env = {}
        dist = self.topo.distance_to_node(self.update_path[0])
        for d in range(max(dist.values()), 0, -1):
            for nid in [n for n in dist if dist[n] == d]:
                to = min(self.topo.neighbours(nid), key=lambda x: dist[x])
                self.update_env(nid, to)

    def update_env(self, nid: str, to: str):
        self.envs[(nid, to)] = C.env_update(self.topo, nid, to, self.s.A[nid], self.W[nid],
                                            self.envs)

    def move(self, path: List[str]):
        for a, b in zip(path[:-1], path[1:]):
            self.s.qr_towards(a, b)
            self.update_env(a, b)
        if path:
            self.s.center = path[-1]

    def update_one_site(self, nid: str) -> float:
        x = self.s.A[nid]
        H = heff_matrix(self.topo, nid, self.W[nid], self.envs, x.shape)
        w, v = np.linalg.eigh(0.5 * (H + H.conj().T))
        self.s.A[nid] = v[:, 0].reshape(x.shape)
        return float(w[0])

    def update_two_site(self, n1: str, n2: str) -> float:
        topo = self.topo
        A1, A2 = self.s.A[n1], self.s.A[n2]
        i12 = topo.neighbour_index(n1, n2)
        i21 = topo.neighbour_index(n2, n1)
        x = np.tensordot(A1, A2, axes=([i12], [i21]))
        k1, k2 = A1.ndim - 2, A2.ndim - 2
        perm = list(range(k1)) + list(range(k1 + 1, k1 + 1 + k2)) + [k1, k1 + 1 + k2]
        x = np.transpose(x, perm)
        n = x.size
        cols = []
        for j in range(n):
            e = np.zeros(n, dtype=complex)
            e[j] = 1.0
            cols.append(C.two_site_heff_apply(topo, n1, n2, e.reshape(x.shape), self.W[n1],
                                              self.W[n2], self.envs).reshape(-1))
        H = np.stack(cols, axis=1)
        w, v = np.linalg.eigh(0.5 * (H + H.conj().T))
        x = v[:, 0].reshape(x.shape)
        u_legs = list(range(k1)) + [k1 + k2]
        v_legs = list(range(k1, k1 + k2)) + [k1 + k2 + 1]
        u, sv, _ = svd_split(x, u_legs, v_legs, **self.svd_params)
        outer1 = list(range(k1))
        self.s.A[n1] = np.transpose(u, outer1[:i12] + [k1 + 1] + outer1[i12:] + [k1])
        outer2 = list(range(1, 1 + k2))
        self.s.A[n2] = np.transpose(sv, outer2[:i21] + [0] + outer2[i21:] + [1 + k2])
        self.s.center = n2
        self.update_env(n1, n2)
        return float(w[0])

    def sweep_one_site(self) -> List[float]:
        up, op = self.update_path, self.orth_path
        es = [self.update_one_site(up[0])]
        for i, nid in enumerate(up[1:]):
            self.move(op[i])
            es.append(self.update_one_site(nid))
        es.append(self.update_one_site(up[-1]))
        for i, nid in enumerate(up[::-1][1:]):
            self.move(op[::-1][i][::-1])
            es.append(self.update_one_site(nid))
        return es

    def sweep_two_site(self) -> List[float]:
        up, op = self.update_path, self.orth_path
        es = []
        for i in range(len(up) - 1):
            for j in range(len(op[i]) - 1):
                es.append(self.update_two_site(op[i][j], op[i][j + 1]))
        for i in range(len(up) - 1):
            path = op[::-1][i][::-1]
            for j in range(len(path) - 1):
                es.append(self.update_two_site(path[j], path[j + 1]))
        return es
