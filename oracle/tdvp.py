"""TDVP sweeps (oracle, test infrastructure only).

Restates the sweep schedules of the reference:

* ``TDVPAlgorithm`` (``time_evolution/tdvp_algorithms/tdvp_algorithm.py:59-221``):
  KEEP canonical form to ``update_path[0]``, cache of all environments pointing
  to it, ``_update_site`` (FORWARD Krylov), ``_move_orth_and_update_cache_for_path``.
* ``OneSiteTDVP`` (``onesitetdvp.py:29-134``): QR(KEEP) split towards the next
  site, environment update, BACKWARD bond evolution, R absorbed into the next site.
* ``FirstOrderOneSiteTDVP`` (``firstorderonesite.py:23-104``): full-dt sweep,
  re-canonicalise to the start and rebuild the cache after every step.
* ``SecondOrderOneSiteTDVP`` (``secondorderonesite.py:65-193``): half-dt forward
  sweep, full-dt final site, half-dt backward sweep.
* ``SecondOrderTwoSiteTDVP`` (``twositetdvp.py:89-138``,
  ``secondordertwosite.py:60-193``): two-site Krylov update, truncated SVD with
  S absorbed into V (VCONTR), backward single-site step at -dt/2.

Tensors keep a fixed layout (parent, children..., phys) on a fixed tree; the
transient link/two-site tensors of the reference are formed explicitly here.
"""
from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np

from . import contract as C
from .krylov import krylov_exp_action
from .splitting import qr_split, svd_split
from .tree import Topology


class OracleState:
    def __init__(self, topo: Topology, tensors: Dict[str, np.ndarray]):
        self.topo = topo
        self.A = {k: np.array(v, dtype=complex) for k, v in tensors.items()}
        self.center: Optional[str] = None

    def copy(self) -> "OracleState":
        s = OracleState(self.topo, {k: v.copy() for k, v in self.A.items()})
        s.center = self.center
        return s

    # -- gauge moves ------------------------------------------------------
    def qr_towards(self, nid: str, nb: str):
        """Split ``nid`` by QR(KEEP) off its leg to ``nb`` and absorb R into ``nb``."""
        t = self.A[nid]
        leg = self.topo.neighbour_index(nid, nb)
        q_legs = [i for i in range(t.ndim) if i != leg]
        q, r = qr_split(t, q_legs, [leg], keep=True)
        # q legs (others..., new) -> put new back at ``leg``
        perm = list(range(q.ndim - 1))
        perm.insert(leg, q.ndim - 1)
        self.A[nid] = np.transpose(q, perm)
        self.absorb(nb, nid, r)

    def absorb(self, nid: str, from_nb: str, r: np.ndarray):
        """A_nid[.., new, ..] = sum_old r[new, old] A_nid[.., old, ..] on the leg to from_nb."""
        t = self.A[nid]
        leg = self.topo.neighbour_index(nid, from_nb)
        out = np.tensordot(r, t, axes=([1], [leg]))       # (new, rest...)
        perm = list(range(1, out.ndim))
        perm.insert(leg, 0)
        self.A[nid] = np.transpose(out, perm)

    def canonical_form(self, center: str):
        dist = self.topo.distance_to_node(center)
        for d in range(max(dist.values()), 0, -1):
            for nid in [n for n in dist if dist[n] == d]:
                nb = min(self.topo.neighbours(nid), key=lambda x: dist[x])
                self.qr_towards(nid, nb)
        self.center = center

    def move_center(self, target: str):
        path = self.topo.path_from_to(self.center, target)
        for a, b in zip(path[:-1], path[1:]):
            self.qr_towards(a, b)
        self.center = target

    def normalise(self):
        assert self.center is not None
        self.A[self.center] /= np.linalg.norm(self.A[self.center])


class OracleTDVP:
    """One-site (first / second order) and two-site (second order) TDVP."""

    def __init__(self, state: OracleState, W: Dict[str, np.ndarray], dt: float,
                 kind: str = "second1", krylov_tol: float = 1e-10,
                 max_krylov_dim: Optional[int] = 10, svd_params: Optional[dict] = None,
                 op_topo: Optional[Topology] = None):
        assert kind in ("first1", "second1", "second2", "first2")
        self.s = state.copy()
        self.topo = state.topo
        self.W = W
        # the operator's own tree (its children order may differ from the state's)
        self.op_topo = op_topo
        self.dt = dt
        self.kind = kind
        self.kopts = dict(krylov_tol=krylov_tol, max_krylov_dim=max_krylov_dim)
        self.svd_params = svd_params or {}
        self.update_path = self.topo.update_path()
        self.orth_path = self.topo.orthogonalisation_path(self.update_path)
        self.back_path = list(reversed(self.update_path))
        if kind in ("second2", "first2"):
            self.back_orth = [list(reversed(p)) for p in reversed(self.orth_path)]
        else:
            self.back_orth = self.topo.second_order_one_site_back_orth_path(self.back_path)
        if self.s.center is None:
            self.s.canonical_form(self.update_path[0])
        else:
            self.s.move_center(self.update_path[0])
        self.envs: C.Env = {}
        self._init_cache()
        self.krylov_dims: List[int] = []

    # -- cache --------------------------------------------------------------
    def _init_cache(self):
        c = self.update_path[0]
        dist = self.topo.distance_to_node(c)
        self.envs = {}
        for d in range(max(dist.values()), 0, -1):
            for nid in [n for n in dist if dist[n] == d]:
                to = min(self.topo.neighbours(nid), key=lambda x: dist[x])
                self.update_env(nid, to)

    def update_env(self, nid: str, to: str):
        self.envs[(nid, to)] = C.env_update(self.topo, nid, to, self.s.A[nid],
                                            self.W[nid], self.envs, op_topo=self.op_topo)

    def _krylov(self, x, action, dt, exp_factor):
        st = {}
        y = krylov_exp_action(x, action, dt, exp_factor, stats=st, **self.kopts)
        self.krylov_dims.append(st.get("m", 0))
        return y

    # -- elementary updates ----------------------------------------------------
    def update_site(self, nid: str, factor: float = 1.0):
        assert self.s.center == nid
        act = lambda y: C.heff_apply(self.topo, nid, y, self.W[nid], self.envs,
                                     op_topo=self.op_topo)
        self.s.A[nid] = self._krylov(self.s.A[nid], act, self.dt * factor, -1j)

    def move_orth_and_update_cache(self, path: List[str]):
        if len(path) == 0:
            return
        assert self.s.center == path[0]
        for a, b in zip(path[:-1], path[1:]):
            self.s.move_center(b)
            self.update_env(a, b)

    def update_link(self, nid: str, nxt: str, factor: float = 1.0):
        assert self.s.center == nid
        t = self.s.A[nid]
        leg = self.topo.neighbour_index(nid, nxt)
        q_legs = [i for i in range(t.ndim) if i != leg]
        q, r = qr_split(t, q_legs, [leg], keep=True)
        perm = list(range(q.ndim - 1))
        perm.insert(leg, q.ndim - 1)
        self.s.A[nid] = np.transpose(q, perm)
        self.update_env(nid, nxt)
        env_a, env_b = self.envs[(nid, nxt)], self.envs[(nxt, nid)]
        act = lambda y: C.bond_apply(y, env_a, env_b)
        r = self._krylov(r, act, self.dt * factor, +1j)
        self.s.absorb(nxt, nid, r)
        self.s.center = nxt

    # -- one-site schedules ------------------------------------------------------
    def _first_order_step(self):
        up, op = self.update_path, self.orth_path
        for i, nid in enumerate(up):
            if i == len(up) - 1:
                if len(op) > 0:
                    self.move_orth_and_update_cache(op[-1])
                self.update_site(nid)
                self.s.move_center(up[0])
                self._init_cache()
            else:
                if i > 0:
                    self.move_orth_and_update_cache(op[i - 1])
                self.update_site(nid)
                self.update_link(nid, op[i][0])

    def _second_order_one_site_step(self):
        up, op = self.update_path, self.orth_path
        for i, nid in enumerate(up[:-1]):
            if i > 0:
                self.move_orth_and_update_cache(op[i - 1])
            nxt = op[i][0]
            self.update_site(nid, 0.5)
            self.update_link(nid, nxt, 0.5)
        self.update_site(up[-1], 1.0)
        bp, bo = self.back_path, self.back_orth
        self.update_link(self.s.center, bp[1], 0.5)
        for i, nid in enumerate(bp[1:-1]):
            idx = i + 1
            self.update_site(nid, 0.5)
            self.move_orth_and_update_cache(bo[idx - 1])
            self.update_link(self.s.center, bp[idx + 1], 0.5)
        self.update_site(bp[-1], 0.5)

    # -- two-site ----------------------------------------------------------------
    def two_site_update(self, n1: str, n2: str, factor: float):
        assert self.s.center == n1
        topo = self.topo
        A1, A2 = self.s.A[n1], self.s.A[n2]
        i12 = topo.neighbour_index(n1, n2)
        i21 = topo.neighbour_index(n2, n1)
        x = np.tensordot(A1, A2, axes=([i12], [i21]))
        # x legs: (outer1..., p1, outer2..., p2) -> (outer1, outer2, p1, p2)
        k1, k2 = A1.ndim - 2, A2.ndim - 2
        perm = list(range(k1)) + list(range(k1 + 1, k1 + 1 + k2)) + [k1, k1 + 1 + k2]
        x = np.transpose(x, perm)
        act = lambda y: C.two_site_heff_apply(topo, n1, n2, y, self.W[n1], self.W[n2],
                                              self.envs, op_topo=self.op_topo)
        x = self._krylov(x, act, self.dt * factor, -1j)
        u_legs = list(range(k1)) + [k1 + k2]
        v_legs = list(range(k1, k1 + k2)) + [k1 + k2 + 1]
        u, sv, s_kept = svd_split(x, u_legs, v_legs, **self.svd_params)
        self.last_singular_values = s_kept
        # u legs (outer1..., p1, new): new goes to position i12
        outer1 = list(range(k1))
        order = outer1[:i12] + [k1 + 1] + outer1[i12:] + [k1]
        self.s.A[n1] = np.transpose(u, order)
        # sv legs (new, outer2..., p2): new goes to position i21
        outer2 = list(range(1, 1 + k2))
        order = outer2[:i21] + [0] + outer2[i21:] + [1 + k2]
        self.s.A[n2] = np.transpose(sv, order)
        self.s.center = n2
        self.update_env(n1, n2)

    def _second_order_two_site_step(self):
        up, op = self.update_path, self.orth_path
        for i, nid in enumerate(up[:-2]):
            if i > 0:
                self.move_orth_and_update_cache(op[i - 1])
            nxt = op[i][0]
            self.two_site_update(nid, nxt, 0.5)
            self.update_site(nxt, -0.5)
        self.two_site_update(up[-2], up[-1], 0.5)
        bp, bo = self.back_path, self.back_orth
        for i in range(len(bp) - 1):
            if i == 0:
                self.two_site_update(bp[0], bp[1], 0.5)
            else:
                path = bo[i]
                self.move_orth_and_update_cache(path)
                target = path[-1]
                self.update_site(target, -0.5)
                self.two_site_update(target, bp[i + 1], 0.5)

    def _first_order_two_site_step(self):
        """First-order two-site sweep (``firstordertwosite.py:9-30``, with the pairs
        and moves of the second-order forward sweep: the reference loop pairs
        non-adjacent nodes), then the centre back to the start with the cache."""
        up, op = self.update_path, self.orth_path
        for i, nid in enumerate(up[:-2]):
            if i > 0:
                self.move_orth_and_update_cache(op[i - 1])
            nxt = op[i][0]
            self.two_site_update(nid, nxt, 1.0)
            self.update_site(nxt, -1.0)
        self.two_site_update(up[-2], up[-1], 1.0)
        self.move_orth_and_update_cache(self.topo.path_from_to(self.s.center, up[0]))

    def run_one_time_step(self):
        if self.kind == "first1":
            self._first_order_step()
        elif self.kind == "second1":
            self._second_order_one_site_step()
        elif self.kind == "first2":
            self._first_order_two_site_step()
        else:
            self._second_order_two_site_step()

    # -- observables -------------------------------------------------------------
    def energy(self) -> complex:
        return C.expectation_value(self.topo, self.s.A, self.W, op_topo=self.op_topo)

    def expect_site(self, nid: str, op: np.ndarray) -> complex:
        return C.single_site_expectation(self.topo, self.s.A, nid, op)

    def norm_sq(self) -> float:
        return C.overlap(self.topo, self.s.A, self.s.A).real
