"""DMRG ground-state sweeps on TTNS (``dmrg/dmrg.py:20-299``).

One-site and two-site sweeps along the leaf-to-leaf update path, with the
environment cache of the time evolution.  Each local problem is the lowest
eigenpair of the site (or two-site) effective Hamiltonian.  The reference
forms that matrix densely and runs a Davidson solver on the host; here the
effective Hamiltonian stays a device network:

* local dimension <= ``DENSE_MAX_DIM``: the dense H_eff is written by the
  contraction itself (``engine.dense_operator``) and diagonalised on the
  device (``torch.linalg.eigh``);
* above: restarted Lanczos with full re-orthogonalisation on the device, the
  basis built by the Krylov kernels (``ptn_normalize``,
  ``ptn_krylov_orthonormalize``), only the small tridiagonal problem on the host,
  until the residual ||H x - theta x|| is below ``EIG_TOL`` times |theta|.

Both converge to the lowest eigenpair.  The reference stops its Davidson at an
energy change of 1e-12 and a residual of 1e-6, or after ``max_iter`` cycles in
the two-site form, so its local vectors agree with these to that tolerance.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from . import engine as E
from . import krylov as K
from .bug import find_path_leaf_to_leaf
from .contractions import SandwichCache, heff_network, operator_tensor
from .splitting import SVDParameters, SplitMode, merge_two_sites, split_two_sites
from .time_evolution import TwoSiteTDVP


def lowest_eigenpair_lanczos(x0: torch.Tensor, action, tol: float = 1e-10,
                             max_krylov_dim: int = 32, max_restarts: int = 200
                             ) -> Tuple[float, torch.Tensor]:
    """Lowest eigenpair of the Hermitian map ``action`` from the start vector x0:
    restarted Lanczos (Arnoldi with CGS2 re-orthogonalisation), restarting from the
    Ritz vector.  Returns (theta, unit vector shaped like x0)."""
    E.require_device(x0)
    shape = x0.shape
    dev = x0.device
    n = x0.numel()
    m_max = max(1, min(n, int(max_krylov_dim)))
    lib = _lib.load()
    work = K._work(dev)
    x = K._flat(x0.to(torch.complex128)).clone()
    if float(torch.linalg.vector_norm(x)) == 0.0:
        x = torch.ones_like(x)
    Q = torch.empty((m_max + 1, n), dtype=torch.complex128, device=dev)
    packed = torch.empty(m_max + 3, dtype=torch.complex128, device=dev)
    theta = 0.0
    for _ in range(max_restarts):
        stream = ctypes.c_void_p(E.stream_handle())
        _lib.check(lib.ptn_normalize(stream, n, K._vp(x), K._vp(Q[0]), K._vp(packed),
                                     K._vp(work)), "ptn_normalize")
        T = np.zeros((m_max + 1, m_max), dtype=complex)
        m = 0
        beta = 0.0
        for j in range(m_max):
            w = action(Q[j].view(shape))
            w = K._flat(w.to(torch.complex128)).clone()
            _lib.check(lib.ptn_krylov_orthonormalize(stream, n, K._vp(Q), Q.stride(0), j + 1,
                                                     K._vp(w), K._vp(Q[j + 1]),
                                                     K._vp(packed[1:]), K._vp(work)),
                       "ptn_krylov_orthonormalize")
            host = packed[: j + 3].cpu().numpy()
            T[: j + 1, j] = host[1: j + 2]
            w0 = math.sqrt(max(host[j + 2].real, 0.0))
            beta = math.sqrt(max(host[j + 2].imag, 0.0))
            T[j + 1, j] = beta
            m = j + 1
            if beta <= max(np.finfo(float).eps * n, 1e-12 * w0):
                beta = 0.0
                break
        S = T[:m, :m]
        evals, evecs = np.linalg.eigh(0.5 * (S + S.conj().T))
        theta, y = float(evals[0]), evecs[:, 0]
        resid = abs(beta * y[m - 1])
        out = torch.empty(n, dtype=torch.complex128, device=dev)
        K.combine(Q, y, out, nvec=m)
        x = out
        if resid <= tol * max(abs(theta), 1.0):
            break
    x = x / torch.linalg.vector_norm(x)
    return theta, x.view(shape)


class DMRGAlgorithm:
    """One-site / two-site DMRG (``dmrg/dmrg.py:20-299``).  The state is updated in
    place, as in the reference."""

    DENSE_MAX_DIM = 4096
    EIG_TOL = 1e-10

    def __init__(self, initial_state, hamiltonian, num_sweeps: int, max_iter: int,
                 svd_params: Optional[SVDParameters], site: str):
        assert len(initial_state.nodes) == len(hamiltonian.nodes)
        if site not in ("one-site", "two-site"):
            raise ValueError("site must be 'one-site' or 'two-site'")
        E.require_device(*initial_state.tensors.values())
        self.state = initial_state
        self.hamiltonian = hamiltonian
        self.num_sweeps = num_sweeps
        self.max_iter = max_iter
        self.site = site
        self.svd_params = SVDParameters() if svd_params is None else svd_params
        self.precision = E.FP64
        self.update_path = find_path_leaf_to_leaf(self.state)
        self.orthogonalization_path = [self.state.path_from_to(a, b) for a, b in
                                       zip(self.update_path[:-1], self.update_path[1:])]
        self._orthogonalize_init()
        self.partial_tree_cache = SandwichCache.init_cache_but_one(
            self.state, self.hamiltonian, self.update_path[0])

    def _orthogonalize_init(self, force_new: bool = False):
        if self.state.orthogonality_center_id is None or force_new:
            self.state.canonical_form(self.update_path[0], mode=SplitMode.KEEP)
        else:
            self.state.move_orthogonalization_center(self.update_path[0], mode=SplitMode.KEEP)

    def update_tree_cache(self, node_id: str, next_node_id: str):
        self.partial_tree_cache.update_tree_cache(node_id, next_node_id)

    def _move_orth_and_update_cache_for_path(self, path: List[str]):
        if len(path) == 0:
            return
        assert self.state.orthogonality_center_id == path[0]
        for i, node_id in enumerate(path[1:]):
            self.state.move_orthogonalization_center(node_id, mode=SplitMode.KEEP)
            self.update_tree_cache(path[i], node_id)

    def _lowest(self, x: torch.Tensor, action) -> Tuple[float, torch.Tensor]:
        if x.numel() <= self.DENSE_MAX_DIM:
            H = E.dense_operator(action)
            H = 0.5 * (H + H.conj().transpose(0, 1))
            evals, evecs = torch.linalg.eigh(H)
            return float(evals[0].item()), evecs[:, 0].contiguous().view(x.shape)
        return lowest_eigenpair_lanczos(x, action, tol=self.EIG_TOL)

    def _update_one_site(self, node_id: str) -> float:
        node, tensor = self.state[node_id]
        op = operator_tensor(self.hamiltonian, node_id, self.precision, size=tensor.numel())
        envs = [self.partial_tree_cache.get_entry(nb, node_id)
                for nb in node.neighbouring_nodes()]
        tensors, labels, out = heff_network(node, self.hamiltonian.nodes[node_id], op, envs)
        tensors[0] = tensor
        action = E.CompiledNetwork(tensors, labels, out, precision=self.precision)
        e, v = self._lowest(tensor, action)
        self.state.tensors[node_id] = v
        return e

    def _update_two_site(self, target_node_id: str, next_node_id: str) -> float:
        n1, n2 = target_node_id, next_node_id
        x = merge_two_sites(self.state, n1, n2)
        make, _, _ = TwoSiteTDVP._two_site_action(self, n1, n2)
        e, v = self._lowest(x, make(x))
        split_two_sites(self.state, v, n1, n2, self.svd_params)
        self.state.orthogonality_center_id = n2
        self.update_tree_cache(n1, n2)
        return e

    @staticmethod
    def create_two_site_id(node_id: str, next_node_id: str) -> str:
        return "TwoSite_" + node_id + "_contr_" + next_node_id

    def sweep_one_site(self) -> float:
        """Forward and backward sweep (``dmrg.py:232-256``)."""
        E.refresh_stream()
        up, op = self.update_path, self.orthogonalization_path
        e = self._update_one_site(up[0])
        for i, node_id in enumerate(up[1:]):
            self._move_orth_and_update_cache_for_path(op[i])
            e = self._update_one_site(node_id)
        e = self._update_one_site(up[-1])
        for i, node_id in enumerate(up[::-1][1:]):
            self._move_orth_and_update_cache_for_path(op[::-1][i][::-1])
            e = self._update_one_site(node_id)
        return e

    def sweep_two_site(self) -> float:
        """Forward and backward sweep of two-site updates (``dmrg.py:258-276``)."""
        E.refresh_stream()
        up, op = self.update_path, self.orthogonalization_path
        e = None
        for i, node_id in enumerate(up[:-1]):
            assert node_id == op[i][0], "node is wrong"
            for j, this in enumerate(op[i][:-1]):
                e = self._update_two_site(this, op[i][j + 1])
        for i, node_id in enumerate(up[::-1][:-1]):
            path = op[::-1][i][::-1]
            assert node_id == path[0], "node is wrong"
            for j, this in enumerate(path[:-1]):
                e = self._update_two_site(this, path[j + 1])
        return e

    def sweep(self) -> float:
        return self.sweep_one_site() if self.site == "one-site" else self.sweep_two_site()

    def run(self) -> List[float]:
        return [self.sweep() for _ in range(self.num_sweeps)]
