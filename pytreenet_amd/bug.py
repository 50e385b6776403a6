"""BUG: the basis-update & Galerkin integrators for TTNS.

Counterparts of ``time_evolution/bug.py`` (rank-adaptive BUG, ``:20-117``),
``time_evolution/fixed_bug.py`` (fixed-rank BUG, ``:19-97``), the shared
recursion of ``time_evolution/time_evo_util/common_bug.py:31-378`` with the basis
helpers of ``bug_util.py:30-209``, and the SVD truncation sweep of
``core/truncation/svd_truncation.py:22-118`` over the leaf-to-leaf path of
``time_evo_util/update_path.py:256-398``.

One step, starting from a state whose orthogonality centre is the root:

* every node n below a parent p is visited depth first.  The centre moves from
  p to n by a QR of p's coefficient tensor off its leg to n (REDUCED for BUG,
  KEEP for the fixed rank form); R goes into n, giving n's coefficient C_n, and
  Q_p with the environments pointing into p gives the environment E(p->n);
* the children of n are updated first; each returns its basis change
  M_c = <old basis|new basis> (old bond x new bond) and its new environment.
  C_n with every M_c contracted into its child leg is evolved by exp(-i dt H_eff)
  with E(p->n) and the children's new environments (K-step);
* n's new basis spans [old; evolved] stacked along the parent leg (BUG) or
  the evolved tensor alone (fixed rank): the Q of a QR over (children, phys) x
  (parent), with the new bond first.  M_n contracts n's old basis tensor, the
  children's M_c and the conjugated new basis;
* the root's coefficient tensor with its children's M_c contracted in is
  evolved with the new environments (Galerkin step) and becomes the new root.

BUG then truncates every bond by one-site SVDs along the leaf-to-leaf path
and moves the centre back to the root (KEEP).  Every tensor operation runs on
the device through the engine's contractions, QR and SVD.  Nodes keep their
leg order (parent, children..., phys) throughout; the reference re-inserts
nodes while it splits and merges, which can move a child leg within its
parent, so the states agree up to that leg order.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import engine as E
from .contractions import SandwichCache, env_block, heff_network, operator_tensor
from .krylov import EvoDirection, TimeEvoMode
from .splitting import (ContractionMode, SVDParameters, SplitMode, absorb_into_neighbour,
                        contr_truncated_svd_splitting, qr_split_to_neighbour,
                        tensor_qr_decomposition)
from .time_evolution import (TTNOBasedTimeEvolution, TTNTimeEvolutionConfig,
                             local_time_evolution)


# ---------------------------------------------------------------------------------------
# leaf-to-leaf path and the SVD truncation sweep
# ---------------------------------------------------------------------------------------
def find_path_leaf_to_leaf(tree) -> List[str]:
    """Every node once: from one end of a longest leaf-to-leaf path to the other,
    each off-path subtree visited children first just before the path node it
    hangs from (``update_path.py:256-398``)."""
    leaves = tree.get_leaves(include_root=True)
    best, pair = -1, (leaves[0], leaves[0])
    for i, a in enumerate(leaves):
        dist = tree.distance_to_node(a)
        for b in leaves[i + 1:]:
            if dist[b] > best:
                best, pair = dist[b], (a, b)
    main = tree.path_from_to(*pair)
    on_main = set(main)
    visited: set = set()

    def subtree(nid: str) -> List[str]:
        if nid in visited or nid in on_main:
            return []
        visited.add(nid)
        node = tree.nodes[nid]
        out: List[str] = []
        if node.parent is not None and node.parent not in on_main \
                and node.parent not in visited:
            out += subtree(node.parent)
        for c in node.children:
            if c not in on_main and c not in visited:
                out += subtree(c)
        out.append(nid)
        return out

    path: List[str] = []
    for origin in main:
        node = tree.nodes[origin]
        if node.parent is not None and node.parent not in on_main \
                and node.parent not in visited:
            path += subtree(node.parent)
        for c in node.children:
            if c not in on_main:
                path += subtree(c)
        if origin not in visited:
            visited.add(origin)
            path.append(origin)
    return path


def compress_parent_leg(node_id: str, tree, params: SVDParameters):
    """Truncated SVD of a node (the orthogonality centre) between its parent leg and
    the rest; U stays in the node, S V^H goes into the parent, which becomes the
    centre (``svd_truncation.py:57-81``)."""
    parent = tree.nodes[node_id].parent
    t = tree.tensors[node_id]
    u, sv = contr_truncated_svd_splitting(t, list(range(1, t.dim())), [0],
                                          contr_mode=ContractionMode.VCONTR,
                                          svd_params=params)
    tree.tensors[node_id] = E.permute_copy(u, [u.dim() - 1] + list(range(u.dim() - 1)))
    absorb_into_neighbour(tree, parent, node_id, sv)
    tree.orthogonality_center_id = parent


def svd_truncation(tree, params: SVDParameters):
    """One-site SVD truncation of every bond along the leaf-to-leaf path
    (``svd_truncation.py:22-55``); ends with the centre at the last path node's
    parent side, as the reference does."""
    path = find_path_leaf_to_leaf(tree)
    tree.canonical_form(path[0])
    for nid in path:
        if nid != tree.root_id:
            tree.move_orthogonalization_center(nid)
            compress_parent_leg(nid, tree, params)
    return tree


# ---------------------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------------------
@dataclass
class CommonBUGConfig(TTNTimeEvolutionConfig):
    """``deep`` is accepted for the reference's signature; tensors are never
    mutated in place here, so no copy is ever taken."""
    deep: bool = False
    fixed_rank: bool = False
    time_evo_mode: TimeEvoMode = TimeEvoMode.FASTEST


@dataclass
class BUGConfig(CommonBUGConfig, SVDParameters):
    def __post_init__(self):
        if self.fixed_rank:
            raise AssertionError("Fixed rank is not supported for BUG!")
        self.check_truncation_parameters()


@dataclass
class FixedBUGConfig(CommonBUGConfig):
    def __post_init__(self):
        self.fixed_rank = True


# ---------------------------------------------------------------------------------------
# the recursion
# ---------------------------------------------------------------------------------------
class _Tensors:
    """A node structure with its own tensor dict: the states of the recursion
    share every node's neighbour order and differ only in a few tensors."""

    def __init__(self, nodes, tensors):
        self.nodes = nodes
        self.tensors = tensors


class _BUGStep:
    def __init__(self, state, hamiltonian, dt: float, config: CommonBUGConfig,
                 solver_options: Optional[dict]):
        self.state = state
        self.H = hamiltonian
        self.dt = dt
        self.fixed = bool(config.fixed_rank)
        self.mode = config.time_evo_mode
        self.opts = solver_options
        self.nodes = state.nodes
        self.old = dict(state.tensors)
        self.new: Dict[str, torch.Tensor] = {}
        root = state.root_id
        cache = SandwichCache.init_cache_but_one(state, hamiltonian, root)
        self.up = {k: v for k, v in cache.items()}     # old environments towards the root
        self.ops = {n: operator_tensor(hamiltonian, n, E.FP64) for n in self.nodes}

    def _evolve(self, nid: str, x: torch.Tensor, envs: List[torch.Tensor]) -> torch.Tensor:
        node = self.nodes[nid]
        tensors, labels, out = heff_network(node, self.H.nodes[nid], self.ops[nid], envs)
        tensors[0] = x
        action = E.CompiledNetwork(tensors, labels, out, precision=E.FP64)
        return local_time_evolution(x, action, self.dt, EvoDirection.FORWARD, self.mode,
                                    self.opts)

    def _with_basis_changes(self, nid: str, C: torch.Tensor,
                            M: Dict[str, torch.Tensor]) -> torch.Tensor:
        """C with M_c[old, new] contracted into each child leg (the reference's
        ``contract_all_children`` of the basis change nodes)."""
        node = self.nodes[nid]
        for c in node.children:
            leg = node.neighbour_index(c)
            lab = list(range(C.dim()))
            lab[leg] = "old"
            out = list(range(C.dim()))
            out[leg] = "new"
            C = E.contract(C, lab, M[c], ["old", "new"], out)
        return C

    def _new_basis(self, nid: str, old: torch.Tensor, evolved: torch.Tensor) -> torch.Tensor:
        """Orthonormal basis tensor, new bond first (``bug_util.py:78-141``)."""
        if self.fixed:
            combined, mode = evolved, SplitMode.KEEP
        else:
            combined, mode = torch.cat([old, evolved], dim=0), SplitMode.REDUCED
        q, _ = tensor_qr_decomposition(combined, list(range(1, combined.dim())), [0],
                                       mode=mode)
        return E.permute_copy(q, [q.dim() - 1] + list(range(q.dim() - 1)))

    def _basis_change(self, nid: str, basis: torch.Tensor,
                      M: Dict[str, torch.Tensor]) -> torch.Tensor:
        """M_n[a, b] = sum A_old[a, c..., s] M_c[c, c'] conj(N[b, c'..., s])
        (``bug_util.py:143-170``)."""
        node = self.nodes[nid]
        A = self.old[nid]
        k = A.dim()
        la = ["a"] + [("c", i) for i in range(1, k - 1)] + ["s"]
        lb = ["b"] + [("d", i) for i in range(1, k - 1)] + ["s"]
        tensors, labels, conj = [A, basis], [la, lb], [False, True]
        for i, c in enumerate(node.children, start=1):
            tensors.append(M[c])
            labels.append([("c", i), ("d", i)])
            conj.append(False)
        return E.contract_network(tensors, labels, ["a", "b"], conj=conj)

    def node(self, nid: str, parent: str, C_parent: torch.Tensor,
             down_parent: Optional[torch.Tensor]) -> Tuple[torch.Tensor, torch.Tensor]:
        """Update the subtree at ``nid`` (``common_bug.py:47-309``).  ``C_parent`` is
        the parent's coefficient tensor with the parent as centre, ``down_parent``
        the environment pointing into the parent from its own parent.  Returns
        (basis change, new environment towards the parent)."""
        nodes = self.nodes
        pnode = nodes[parent]
        view = _Tensors(nodes, {parent: C_parent, nid: self.old[nid]})
        mode = SplitMode.KEEP if self.fixed else SplitMode.REDUCED
        q, r = qr_split_to_neighbour(view, parent, nid, mode=mode)
        absorb_into_neighbour(view, nid, parent, r)
        C = view.tensors[nid]
        to = pnode.neighbour_index(nid)
        envs_p = []
        for i, nb in enumerate(pnode.neighbouring_nodes()):
            if i == to:
                envs_p.append(None)
            elif nb == pnode.parent:
                envs_p.append(down_parent)
            else:
                envs_p.append(self.up[(nb, parent)])
        down = env_block(pnode, q, self.H.nodes[parent], self.ops[parent], envs_p, to)
        node = nodes[nid]
        M: Dict[str, torch.Tensor] = {}
        new_envs: Dict[str, torch.Tensor] = {}
        for c in node.children:
            M[c], new_envs[c] = self.node(c, nid, C, down)
        if node.children:
            C = self._with_basis_changes(nid, C, M)
            old = C
        else:
            old = self.old[nid]
        envs = [down] + [new_envs[c] for c in node.children]
        evolved = self._evolve(nid, C, envs)
        basis = self._new_basis(nid, old, evolved)
        Mn = self._basis_change(nid, basis, M)
        env_new = env_block(node, basis, self.H.nodes[nid], self.ops[nid],
                            [None] + [new_envs[c] for c in node.children], 0)
        self.new[nid] = basis
        return Mn, env_new

    def run(self):
        """Root update (``common_bug.py:311-378``); writes the new tensors into the state."""
        root = self.state.root_id
        rnode = self.nodes[root]
        C = self.old[root]
        M: Dict[str, torch.Tensor] = {}
        new_envs: Dict[str, torch.Tensor] = {}
        for c in rnode.children:
            M[c], new_envs[c] = self.node(c, root, C, None)
        C = self._with_basis_changes(root, C, M)
        self.new[root] = self._evolve(root, C, [new_envs[c] for c in rnode.children])
        for nid, t in self.new.items():
            self.state.tensors[nid] = t
        self.state.orthogonality_center_id = root


class _BUGBase(TTNOBasedTimeEvolution):
    def __init__(self, initial_state, hamiltonian, time_step_size: float, final_time: float,
                 operators, config=None, solver_options=None):
        super().__init__(initial_state, hamiltonian, time_step_size, final_time, operators,
                         config=config, solver_options=solver_options)
        E.require_device(*self.state.tensors.values())
        mode = self.config.time_evo_mode
        if mode is not TimeEvoMode.KRYLOV and mode is not TimeEvoMode.FASTEST \
                and not mode.is_dense():
            raise NotImplementedError(f"mode {mode.value}: the device path evolves with "
                                      "TimeEvoMode.KRYLOV or a dense exponential mode")
        self.state.ensure_root_orth_center(mode=SplitMode.KEEP)

    def recursive_update(self):
        E.refresh_stream()
        _BUGStep(self.state, self.hamiltonian, self.time_step_size, self.config,
                 self.solver_options).run()


class BUG(_BUGBase):
    """Rank-adaptive BUG (``time_evolution/bug.py:20-117``)."""
    config_class = BUGConfig

    def truncation(self):
        svd_truncation(self.state, self.config)
        self.state.move_orthogonalization_center(self.state.root_id, mode=SplitMode.KEEP)

    def run_one_time_step(self, **kwargs):
        self.recursive_update()
        self.truncation()


class FixedBUG(_BUGBase):
    """Fixed-rank BUG (``time_evolution/fixed_bug.py:19-97``)."""
    config_class = FixedBUGConfig

    def run_one_time_step(self, **kwargs):
        self.recursive_update()
