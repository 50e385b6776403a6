"""TDVP time evolution on the device: the reference's drop-in call surface.

Classes and methods keep the reference names, arguments and sweep schedules:

* ``TimeEvolution`` run loop and step count -- ``time_evolution/time_evolution.py:31-368``
* ``TTNTimeEvolutionConfig`` / operator evaluation -- ``ttn_time_evolution.py:23-279``
* ``TDVPConfig`` / ``TDVPAlgorithm`` -- ``tdvp_algorithms/tdvp_algorithm.py:24-221``
* ``OneSiteTDVP`` -- ``tdvp_algorithms/onesitetdvp.py:17-134``
* ``FirstOrderOneSiteTDVP`` -- ``tdvp_algorithms/firstorderonesite.py:7-104``
* ``SecondOrderOneSiteTDVP`` -- ``tdvp_algorithms/secondorderonesite.py:11-193``
* ``TwoSiteTDVPConfig`` / ``TwoSiteTDVP`` -- ``tdvp_algorithms/twositetdvp.py:18-146``
* ``SecondOrderTwoSiteTDVP`` -- ``tdvp_algorithms/secondordertwosite.py:11-193``
* ``FirstOrderTwoSiteTDVP`` -- ``tdvp_algorithms/firstordertwosite.py:9-30``
* update / orthogonalisation paths -- ``time_evo_util/update_path.py:63-420``

All tensors stay on the device.  The transient link tensor of a one-site
update is the R factor held as a (node side, next side) matrix rather than a
node inserted into the tree; the two-site tensor has legs
(outer legs of the first node, outer legs of the second, phys1, phys2).
``TDVPConfig.precision`` selects FP64 (complex128, FP64 MFMA), the mixed form
"fp32" (complex64 environments, 3M FP32 MFMA contractions; Krylov
re-orthogonalisation, the small propagator, QR and SVD stay FP64), or
"relaxed": FP64 throughout except the Krylov applications of H_eff whose
contribution to the local result is below ``relax_tol``, which run in 3M FP32
on complex64 copies of the environments.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Union

import numpy as np
import torch

from . import contractions as _C
from . import engine as E
from .contractions import (SandwichCache, bond_apply, contract_any,
                           contract_ket_ham_with_envs, expectation_value, heff_network,
                           operator_tensor, scalar_product)
from . import krylov as _K
from .krylov import EvoDirection, TimeEvoMode, krylov_exp_action
from .splitting import (SVDParameters, SplitMode, SvdRequest, answer_svd,
                        answer_svd_batched, merge_two_sites, split_two_sites,
                        split_two_sites_gen)
from .ttn import TreeTensorNetworkState
from .ttno import TensorProduct, TreeTensorNetworkOperator


# ---------------------------------------------------------------------------------------
# configs and results
# ---------------------------------------------------------------------------------------
@dataclass
class TimeEvoConfig:
    time_dep: bool = False


@dataclass
class TTNTimeEvolutionConfig(TimeEvoConfig):
    record_bond_dim: bool = False
    record_max_bdim: bool = False
    record_average_bdim: bool = False
    record_total_size: bool = False
    record_norm: bool = False
    record_loschmidt_amplitude: bool = False


class PathFinderMode:
    LeafToRoot = "LeafToRoot"


@dataclass
class TDVPConfig(TTNTimeEvolutionConfig):
    time_evo_mode: TimeEvoMode = TimeEvoMode.FASTEST
    main_path_mode: str = PathFinderMode.LeafToRoot
    precision: str = E.FP64
    # precision="relaxed": largest change of a local Krylov result, relative to
    # the norm of its input, that the low-precision applications may cause
    # (engine.RelaxedNetwork)
    relax_tol: float = 1e-12
    # smallest local space (elements) whose applications may drop to FP32: below
    # it the H_eff GEMMs are bound by memory and launches, where the second
    # network costs more than the FP32 arithmetic saves
    relax_min_size: int = 1 << 20
    # precision="relaxed": also run applications of the smallest weights as 3M
    # bf16 products (FP32 accumulation)
    relax_bf16: bool = False
    # precision="relaxed": arithmetic of the FP32-accurate level, "fp32" (FP32 MFMA)
    # or "bf16x6" (six products of three-piece bf16 splits on the bf16 MFMA)
    relax_low: str = "fp32"


@dataclass
class TwoSiteTDVPConfig(TDVPConfig, SVDParameters):
    pass


class Results:
    """Time series of operator results (``time_evolution/results.py``), kept on the host."""

    def __init__(self):
        self._times: List[float] = []
        self._data: Dict[str, List[Any]] = {}

    def initialize(self, keys, length: int):
        self._times = [0.0] * length
        self._data = {k: [None] * length for k in keys}

    def set_element(self, key: str, index: int, value):
        self._data.setdefault(key, [None] * len(self._times))[index] = value

    def set_time(self, index: int, t: float):
        self._times[index] = t

    def times(self) -> np.ndarray:
        return np.asarray(self._times, dtype=float)

    def operator_result(self, key: str, realise: bool = False) -> np.ndarray:
        arr = np.asarray(self._data[key])
        return arr.real if realise else arr

    def operator_results(self, realise: bool = False) -> Dict[str, np.ndarray]:
        return {k: self.operator_result(k, realise) for k in self._data}

    def save_npz(self, path: str):
        np.savez(path, times=self.times(),
                 **{f"op_{k}": np.asarray(v) for k, v in self._data.items()})


# ---------------------------------------------------------------------------------------
# base classes
# ---------------------------------------------------------------------------------------
class TimeEvolution:
    config_class = TimeEvoConfig

    def __init__(self, initial_state, time_step_size: float, final_time: float, operators,
                 solver_options: Optional[Dict[str, Any]] = None, config=None):
        self._initial_state = initial_state.copy()
        self.state = initial_state.copy()
        if time_step_size <= 0:
            raise ValueError("size of one time step has to be positive")
        if final_time <= 0:
            raise ValueError("final time has to be positive")
        self._time_step_size = time_step_size
        self._final_time = final_time
        self._num_time_steps = self._compute_num_time_steps()
        self.operators = self.init_operators(operators)
        self.config = self.config_class() if config is None else config
        self.solver_options = {} if solver_options is None else dict(solver_options)
        self.results = Results()

    def _compute_num_time_steps(self) -> int:
        decimal, integer = math.modf(self._final_time / self._time_step_size)
        return int(integer) if decimal < 0.1 else int(integer + 1)

    @staticmethod
    def init_operators(operators) -> Dict[str, Any]:
        if isinstance(operators, dict) and not isinstance(operators, TensorProduct):
            return dict(operators)
        if isinstance(operators, (list, tuple)):
            return {str(i): op for i, op in enumerate(operators)}
        return {"0": operators}

    @property
    def initial_state(self):
        return self._initial_state

    @property
    def time_step_size(self) -> float:
        return self._time_step_size

    @property
    def num_time_steps(self) -> int:
        return self._num_time_steps

    @property
    def final_time(self) -> float:
        return self._final_time

    def run_one_time_step(self, **kwargs):
        raise NotImplementedError

    def evaluate_operator(self, operator) -> complex:
        raise NotImplementedError

    def evaluate_operators(self, index: int):
        for key, op in self.operators.items():
            self.results.set_element(key, index, self.evaluate_operator(op))

    def run(self, evaluation_time: Union[int, str] = 1, pgbar: bool = False):
        n = self._num_time_steps
        if evaluation_time == "inf":
            idx = {n: 0}
            length = 1
        else:
            idx = {i: i // evaluation_time for i in range(0, n + 1, evaluation_time)}
            length = len(idx)
        self.results.initialize(list(self.operators.keys()), length)
        for i in range(n + 1):
            if i != 0:
                self.run_one_time_step()
            if i in idx:
                self.evaluate_operators(idx[i])
                self.results.set_time(idx[i], i * self._time_step_size)


class TTNTimeEvolution(TimeEvolution):
    """Time evolution of a TTNS with operator evaluation and the optional
    bond-dimension / norm / Loschmidt records (``time_evolution/ttn_time_evolution.py:40-179``)."""
    config_class = TTNTimeEvolutionConfig

    def evaluate_operator(self, operator) -> complex:
        return self.state.operator_expectation_value(operator)

    def evaluate_operators(self, index: int):
        super().evaluate_operators(index)
        cfg = self.config
        if getattr(cfg, "record_max_bdim", False):
            self.results.set_element("max_bond_dim", index, self.state.max_bond_dim())
        if getattr(cfg, "record_norm", False):
            self.results.set_element("norm", index, self.state.norm())
        if getattr(cfg, "record_loschmidt_amplitude", False):
            self.results.set_element("loschmidt_amplitude", index,
                                     scalar_product(self.state, self._initial_state))


class TTNOBasedTimeEvolution(TTNTimeEvolution):
    def __init__(self, initial_state, hamiltonian, time_step_size, final_time, operators,
                 config=None, solver_options=None):
        super().__init__(initial_state, time_step_size, final_time, operators,
                         solver_options=solver_options, config=config)
        self.hamiltonian = hamiltonian


# ---------------------------------------------------------------------------------------
# path finding (update_path.py:63-254, 400-420)
# ---------------------------------------------------------------------------------------
def _branch_rec(ttn, nid: str) -> List[str]:
    node = ttn.nodes[nid]
    if node.is_leaf():
        return [nid]
    out: List[str] = []
    for c in node.children:
        out.extend(_branch_rec(ttn, c))
    out.append(nid)
    return out


def find_update_path_leaf_to_root(ttn) -> List[str]:
    root = ttn.root_id
    dist = ttn.distance_to_node(root)
    start = max(dist, key=dist.get)
    main_path = ttn.find_path_to_root(start)
    path: List[str] = []
    for origin in main_path:
        if origin != root:
            for c in ttn.nodes[origin].children:
                if c not in main_path:
                    path.extend(_branch_rec(ttn, c))
            path.append(origin)
            continue
        if ttn.nodes[root].has_x_children(1):
            path.append(root)
            continue
        leaves = [l for l in ttn.get_leaves() if l not in path]
        leaf_dist = {n: d for n, d in dist.items() if n in leaves}
        final = max(leaf_dist, key=leaf_dist.get)
        down = list(reversed(ttn.find_path_to_root(final)))
        for o in down:
            excluded = (main_path[-2], down[1]) if o == root else tuple(down)
            for c in ttn.nodes[o].children:
                if c not in excluded:
                    path.extend(_branch_rec(ttn, c))
            path.append(o)
    return path


def find_orthogonalisation_path(update_path: List[str], ttn,
                                include_start: bool = False) -> List[List[str]]:
    out = []
    for i in range(len(update_path) - 1):
        sub = ttn.path_from_to(update_path[i], update_path[i + 1])
        out.append(sub if include_start else sub[1:])
    return out


# ---------------------------------------------------------------------------------------
# TDVP
# ---------------------------------------------------------------------------------------
# FASTEST evolves with the dense effective Hamiltonian (like the reference) up to
# this local dimension and with the Krylov recurrence above it, where a dense
# matrix (dim^2 complex) is out of reach
DENSE_MAX_DIM = 4096


def local_time_evolution(x: torch.Tensor, action, dt: float, direction: EvoDirection,
                         mode: TimeEvoMode, solver_options: Optional[Dict[str, Any]] = None,
                         dense_max_dim: int = DENSE_MAX_DIM) -> torch.Tensor:
    """exp(f dt H_eff) x for a site, bond or two-site action ``action``
    (``time_evo_util/effective_time_evolution.py:28-140``).

    KRYLOV (and FASTEST above ``dense_max_dim``) runs the matrix-free Krylov
    exponential; the dense modes build H_eff as a matrix (the network without x,
    ``effective_hamiltonians.py:20-470``) and exponentiate it exactly
    (``time_evolution.py:617-655``, ``util/std_utils.py:134-175``).
    """
    if mode is TimeEvoMode.FASTEST and x.numel() > dense_max_dim:
        mode = TimeEvoMode.KRYLOV
    opts = dict(TimeEvoMode.default_solver_options(mode))
    opts.update(solver_options or {})
    relaxed = None
    if isinstance(action, E.RelaxedNetwork):
        # only the matrix-free Krylov loop uses the low-precision form
        relaxed, action = action, action.high
    if mode is TimeEvoMode.KRYLOV:
        if x.numel() <= _K.SMALL_N and isinstance(action, E.CompiledNetwork):
            # short local spaces: the recurrence runs in one launch on the dense
            # H_eff (one contraction instead of one network replay per step)
            y = _K.small_time_evolve(x, E.dense_operator(action, transposed=True), dt,
                                     direction, opts, transposed=True)
            if y is not None:
                return y
        return mode.time_evolve_action(x, lambda t, y: action(y), dt, forward=direction,
                                       relaxed=relaxed, **opts)
    H = E.dense_operator(action)
    return mode.time_evolve(x, H, dt, forward=direction,
                            **{k: v for k, v in opts.items() if k == "exp_factor"})


class KrylovRequest:
    """One pending exp(f dt H) x of a sweep: x, the H action, dt, direction."""
    __slots__ = ("x", "action", "dt", "direction", "result", "applies")

    def __init__(self, x, action, dt, direction):
        self.x, self.action, self.dt, self.direction = x, action, dt, direction
        self.result = None
        self.applies = 0


class TDVPAlgorithm(TTNOBasedTimeEvolution):
    config_class = TDVPConfig

    def __init__(self, initial_state: TreeTensorNetworkState,
                 hamiltonian: TreeTensorNetworkOperator, time_step_size: float,
                 final_time: float, operators, config: Optional[TDVPConfig] = None,
                 solver_options: Optional[Dict[str, Any]] = None):
        assert len(initial_state.nodes) == len(hamiltonian.nodes)
        super().__init__(initial_state, hamiltonian, time_step_size, final_time, operators,
                         config=config, solver_options=solver_options)
        mode = self.config.time_evo_mode
        if mode is not TimeEvoMode.KRYLOV and not mode.is_dense():
            raise NotImplementedError(f"mode {mode.value}: the device path evolves with "
                                      "TimeEvoMode.KRYLOV or a dense exponential mode")
        self._mode = mode
        self.precision = getattr(self.config, "precision", E.FP64)
        if self.precision not in (E.FP64, E.FP32, E.RELAXED):
            raise ValueError(f"unknown precision {self.precision!r}")
        # relaxed: FP64 storage and environments; site applications in the Krylov
        # loop drop to 3M FP32 where their contribution is small
        self.relaxed = self.precision == E.RELAXED
        if self.relaxed:
            self.precision = E.FP64
        self.apply_log: List[tuple] = []   # (kind, node or edge, Krylov applies)
        self._last_applies = 0
        # dense TTNO tensors in the site networks of local spaces below this many
        # elements (a BatchedTDVP sets it: its grouped replay needs dense operators;
        # the spaces it solves one instance at a time keep the block-sparse ones)
        self.dense_site_below = 0
        self.update_path = self._finds_update_path()
        self.orthogonalization_path = find_orthogonalisation_path(self.update_path, self.state)
        self._orthogonalize_init()
        self.partial_tree_cache = self._init_partial_tree_cache()

    def _finds_update_path(self) -> List[str]:
        return find_update_path_leaf_to_root(self.state)

    def _orthogonalize_init(self, force_new: bool = False):
        if self.state.orthogonality_center_id is None or force_new:
            self.state.canonical_form(self.update_path[0], mode=SplitMode.KEEP)
        else:
            self.state.move_orthogonalization_center(self.update_path[0], mode=SplitMode.KEEP)

    def _init_partial_tree_cache(self) -> SandwichCache:
        # the state is canonical around update_path[0], and the sweeps keep every
        # block's subtree an isometry towards the node it points to
        return SandwichCache.init_cache_but_one(self.state, self.hamiltonian,
                                                self.update_path[0], precision=self.precision,
                                                canonical=True)

    def _site_identity(self, node, tensor: torch.Tensor, op) -> Optional[Dict[int, tuple]]:
        """Identity environment blocks of a site network [x, W, E_0, ...] (slots 2...)
        when W is block sparse and the site tensor is large enough."""
        if not (_C.IDENTITY_BLOCKS and isinstance(op, E.BlockSparse)
                and tensor.numel() >= _C.IDENTITY_MIN_SIZE):
            return None
        return _C.identity_slots(node.identifier, node.neighbouring_nodes(), self.hamiltonian, 2)

    def update_tree_cache(self, node_id: str, next_node_id: str):
        self.partial_tree_cache.update_tree_cache(node_id, next_node_id)

    DENSE_MAX_DIM = 4096

    def _krylov(self, x: torch.Tensor, action, dt: float, direction: EvoDirection):
        return local_time_evolution(x, action, dt, direction, self._mode, self.solver_options,
                                    self.DENSE_MAX_DIM)

    def krylov_options(self) -> Dict[str, Any]:
        opts = dict(TimeEvoMode.default_solver_options(self._mode))
        opts.update(self.solver_options)
        return opts

    # The sweep steps are generators: every Krylov exponential is handed out as a
    # KrylovRequest (``y = yield from self._evolve(...)``).  ``_drive`` answers
    # the requests one by one; a BatchedTDVP answers those of many instances at
    # once.
    def _evolve(self, x: torch.Tensor, action, dt: float, direction: EvoDirection):
        req = KrylovRequest(x, action, dt, direction)
        yield req
        self._last_applies = req.applies
        return req.result

    def _drive(self, gen):
        from .krylov import STATS
        try:
            req = next(gen)
            while True:
                if isinstance(req, SvdRequest):
                    answer_svd(req)
                else:
                    a0, p0 = STATS.applies, _K.pending_count()
                    req.result = self._krylov(req.x, req.action, req.dt, req.direction)
                    # a one-launch solve of a deferred block hands out its Krylov
                    # dimension later (resolved at the end of the time step)
                    req.applies = (_K.PendingApplies(p0) if _K.pending_count() > p0
                                   else STATS.applies - a0)
                req = gen.send(None)
        except StopIteration as stop:
            return stop.value

    def _time_step_gen(self):
        """One time step as a generator of KrylovRequests."""
        raise NotImplementedError
        yield

    # A time step runs with the CholeskyQR pivot checks deferred to its end (no
    # device -> host stall after every split).  If a split turned out rank
    # deficient, the step is restored from its start and run again with every
    # split checked as it happens (and the Householder fallback where needed).
    # The state, the cache and the logs are only ever rebound to new tensors,
    # so the snapshot holds references, not copies.
    DEFER_QR_CHECKS = True
    steps_redone = 0

    def _snapshot(self):
        cache = self.partial_tree_cache
        return (dict(self.state.tensors), self.state.orthogonality_center_id, cache,
                dict(cache), len(self.apply_log), len(getattr(self, "kept_dims", [])))

    def _restore(self, snap):
        tensors, center, cache, entries, nlog, nkept = snap
        for k, v in tensors.items():
            self.state.tensors[k] = v
        self.state.orthogonality_center_id = center
        cache.clear()
        cache.update(entries)
        self.partial_tree_cache = cache
        del self.apply_log[nlog:]
        if hasattr(self, "kept_dims"):
            del self.kept_dims[nkept:]

    def run_one_time_step(self, **kwargs):
        E.refresh_stream()
        if not self.DEFER_QR_CHECKS:
            self._drive(self._time_step_gen())
            return
        from .splitting import deferred_qr_checks
        snap = self._snapshot()
        nlog = len(self.apply_log)
        with deferred_qr_checks() as checks, _K.deferred_small_info() as small:
            try:
                self._drive(self._time_step_gen())
            except Exception:
                # work downstream of a collapsed split may fail in its own way
                if not checks.failed():
                    raise
                failed = True
            else:
                failed = checks.failed()
            if not failed:
                # the Krylov dimensions of the step's one-launch solves, read once
                ms = _K.resolve_pending(small)
                self.apply_log[nlog:] = [
                    (k, n, ms[a.index] if isinstance(a, _K.PendingApplies) else a)
                    for k, n, a in self.apply_log[nlog:]]
                if isinstance(self._last_applies, _K.PendingApplies):
                    self._last_applies = ms[self._last_applies.index]
        if failed:
            TDVPAlgorithm.steps_redone += 1
            self._restore(snap)
            self._drive(self._time_step_gen())

    def _update_site(self, node_id: str, time_step_factor: float = 1):
        return self._drive(self._update_site_gen(node_id, time_step_factor))

    def _update_site_gen(self, node_id: str, time_step_factor: float = 1):
        node, tensor = self.state[node_id]
        op_node = self.hamiltonian.nodes[node_id]
        dense = tensor.numel() < self.dense_site_below
        op = operator_tensor(self.hamiltonian, node_id, self.precision,
                             block=False if dense else None, size=tensor.numel())
        cache = self.partial_tree_cache
        envs = [cache.get_entry(nb, node_id) for nb in node.neighbouring_nodes()]
        tensors, labels, out_l = heff_network(node, op_node, op, envs)
        tensors[0] = tensor
        ident = self._site_identity(node, tensor, op)
        # recorded on the first Krylov apply, replayed for the others
        action = E.CompiledNetwork(tensors, labels, out_l, precision=self.precision,
                                   identity=ident)
        if getattr(self, "relaxed", False) and \
                tensor.numel() >= getattr(self.config, "relax_min_size", 0):
            op32 = operator_tensor(self.hamiltonian, node_id, E.FP32,
                                   block=False if dense else None, size=tensor.numel())
            t32, l32, o32 = heff_network(node, op_node, op32,
                                         [e.to(torch.complex64) for e in envs])
            t32[0] = tensor
            # complex64 output: the last product then has complex64 operands
            # and result, like the others
            action = self._relaxed(action, E.CompiledNetwork(t32, l32, o32,
                                                             precision=self._low_precision(),
                                                             out_dtype=torch.complex64,
                                                             identity=ident),
                                   (t32, l32, o32), ident)

        self.state.tensors[node_id] = yield from self._evolve(
            tensor, action, self.time_step_size * time_step_factor, EvoDirection.FORWARD)
        self.apply_log.append(("site", node_id, self._last_applies))

    def _low_precision(self) -> str:
        low = getattr(self.config, "relax_low", E.FP32)
        if low not in (E.FP32, E.BF16X6):
            raise ValueError(f"relax_low must be 'fp32' or 'bf16x6', not {low!r}")
        return low

    def _relaxed(self, high, low, net32, ident=None):
        lower = None
        if getattr(self.config, "relax_bf16", False):
            lower = E.CompiledNetwork(*net32, precision=E.BF16, out_dtype=torch.complex64,
                                      identity=ident)
        return E.RelaxedNetwork(high, low, getattr(self.config, "relax_tol", 1e-12),
                                lower=lower)

    def _move_orth_and_update_cache_for_path(self, path: List[str]):
        if len(path) == 0:
            return
        assert self.state.orthogonality_center_id == path[0]
        for i, node_id in enumerate(path[1:]):
            self.state.move_orthogonalization_center(node_id, mode=SplitMode.KEEP)
            self.partial_tree_cache.update_tree_cache(path[i], node_id)


class OneSiteTDVP(TDVPAlgorithm):
    """Site + link updates; the link is the R factor of a QR(KEEP) split."""

    def _update_link(self, node_id: str, next_node_id: str, time_step_factor: float = 1):
        return self._drive(self._update_link_gen(node_id, next_node_id, time_step_factor))

    def _update_link_gen(self, node_id: str, next_node_id: str, time_step_factor: float = 1):
        assert self.state.orthogonality_center_id == node_id
        from .splitting import qr_split_to_neighbour, absorb_into_neighbour
        q, r = qr_split_to_neighbour(self.state, node_id, next_node_id)
        self.state.tensors[node_id] = q
        self.partial_tree_cache.update_tree_cache(node_id, next_node_id)
        env_a = self.partial_tree_cache.get_entry(node_id, next_node_id)
        env_b = self.partial_tree_cache.get_entry(next_node_id, node_id)
        bond = E.CompiledNetwork([r, env_a, env_b],
                                 [["a0", "a1"], ["a0", "w", "b0"], ["a1", "w", "b1"]],
                                 ["b0", "b1"], precision=self.precision)
        r = yield from self._evolve(r, bond, self.time_step_size * time_step_factor,
                                    EvoDirection.BACKWARD)
        self.apply_log.append(("bond", (node_id, next_node_id), self._last_applies))
        absorb_into_neighbour(self.state, next_node_id, node_id, r)
        self.state.orthogonality_center_id = next_node_id

    @staticmethod
    def create_link_id(node_id: str, next_node_id: str) -> str:
        return "link_" + node_id + "_with_" + next_node_id


class FirstOrderOneSiteTDVP(OneSiteTDVP):
    def _update_site_and_link(self, node_id: str, update_index: int):
        return self._drive(self._update_site_and_link_gen(node_id, update_index))

    def _update_site_and_link_gen(self, node_id: str, update_index: int):
        yield from self._update_site_gen(node_id)
        yield from self._update_link_gen(node_id, self.orthogonalization_path[update_index][0])

    def _reset_for_next_time_step(self):
        self.state.move_orthogonalization_center(self.update_path[0], mode=SplitMode.KEEP)
        self.partial_tree_cache = self._init_partial_tree_cache()

    def _time_step_gen(self):
        up, op = self.update_path, self.orthogonalization_path
        for i, node_id in enumerate(up):
            if i == len(up) - 1:
                if len(op) > 0:
                    self._move_orth_and_update_cache_for_path(op[-1])
                yield from self._update_site_gen(node_id)
                self._reset_for_next_time_step()
            elif i == 0:
                yield from self._update_site_and_link_gen(node_id, 0)
            else:
                self._move_orth_and_update_cache_for_path(op[i - 1])
                yield from self._update_site_and_link_gen(node_id, i)


class SecondOrderOneSiteTDVP(OneSiteTDVP):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.backwards_update_path = list(reversed(self.update_path))
        self.backwards_orth_path = self._init_second_order_orth_path()

    def _init_second_order_orth_path(self) -> List[List[str]]:
        bp = self.backwards_update_path
        out = []
        for i, node_id in enumerate(bp[1:-1]):
            out.append(self.state.path_from_to(node_id, bp[i + 2])[:-1])
        out.append([bp[-1]])
        return out

    def forward_sweep(self):
        return self._drive(self._forward_sweep_gen())

    def _forward_sweep_gen(self):
        for i, node_id in enumerate(self.update_path[:-1]):
            if i > 0:
                self._move_orth_and_update_cache_for_path(self.orthogonalization_path[i - 1])
            nxt = self.orthogonalization_path[i][0]
            yield from self._update_site_gen(node_id, time_step_factor=0.5)
            yield from self._update_link_gen(node_id, nxt, time_step_factor=0.5)

    def _final_forward_update(self):
        return self._drive(self._final_forward_update_gen())

    def _final_forward_update_gen(self):
        node_id = self.update_path[-1]
        assert self.state.orthogonality_center_id == node_id
        yield from self._update_site_gen(node_id)

    def backward_sweep(self):
        return self._drive(self._backward_sweep_gen())

    def _backward_sweep_gen(self):
        bp = self.backwards_update_path
        yield from self._update_link_gen(self.state.orthogonality_center_id, bp[1],
                                         time_step_factor=0.5)
        for i, node_id in enumerate(bp[1:-1]):
            idx = i + 1
            assert self.state.orthogonality_center_id == node_id
            yield from self._update_site_gen(node_id, time_step_factor=0.5)
            self._move_orth_and_update_cache_for_path(self.backwards_orth_path[idx - 1])
            yield from self._update_link_gen(self.state.orthogonality_center_id, bp[idx + 1],
                                             time_step_factor=0.5)
        yield from self._update_site_gen(bp[-1], time_step_factor=0.5)

    def _time_step_gen(self):
        yield from self._forward_sweep_gen()
        yield from self._final_forward_update_gen()
        yield from self._backward_sweep_gen()


class TwoSiteTDVP(TDVPAlgorithm):
    config_class = TwoSiteTDVPConfig

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        if self.state.num_nodes() < 2:
            raise ValueError("Two-site TDVP cannot be used for systems with less than 2 sites!")
        self.kept_dims: List[int] = []

    def _two_site_action(self, n1: str, n2: str):
        """H_eff on the merged tensor with legs (outer1..., outer2..., p1, p2)."""
        s, h, cache = self.state, self.hamiltonian, self.partial_tree_cache
        dt = E.storage_dtype(self.precision)
        node1, node2 = s.nodes[n1], s.nodes[n2]
        op1n, op2n = h.nodes[n1], h.nodes[n2]
        nb1 = [n for n in node1.neighbouring_nodes() if n != n2]
        nb2 = [n for n in node2.neighbouring_nodes() if n != n1]
        k1, k2 = len(nb1), len(nb2)
        # operator labels: per operator leg, the matching ket-side label
        def op_lab(op_node, own_nbs, offset, other, phys_out, phys_in):
            lab = [None] * (op_node.nneighbours() + 2)
            for i, nb in enumerate(own_nbs):
                lab[op_node.neighbour_index(nb)] = ("w", offset + i)
            lab[op_node.neighbour_index(other)] = "wl"
            lab[-2], lab[-1] = phys_out, phys_in
            return lab
        envs = [cache.get_entry(nb, n1) for nb in nb1] + [cache.get_entry(nb, n2) for nb in nb2]
        env_l = [[("a", i), ("w", i), ("b", i)] for i in range(k1 + k2)]
        x_l = [("a", i) for i in range(k1 + k2)] + ["p1", "p2"]
        out = [("b", i) for i in range(k1 + k2)] + ["q1", "q2"]
        dense_l = [x_l, op_lab(op1n, nb1, 0, n2, "q1", "p1"),
                   op_lab(op2n, nb2, k1, n1, "q2", "p2")] + env_l
        # In the canonical form of the sweep the outer environments have identity
        # blocks: the two operator tensors then meet as one block-sparse two-site
        # tensor, and the network splits at those blocks (engine._split_identity)
        split = getattr(cache, "canonical", False) and _C.IDENTITY_BLOCKS
        ident = None
        if split:
            ident = {2 + i: h.identity_states(nb, n1) for i, nb in enumerate(nb1)}
            ident.update({2 + k1 + i: h.identity_states(nb, n2) for i, nb in enumerate(nb2)})
            split_l = [x_l, [("w", i) for i in range(k1 + k2)] + ["q1", "q2", "p1", "p2"]] + env_l
        split_min = max(_C.IDENTITY_MIN_SIZE, getattr(self, "dense_site_below", 0))

        def net(y, dtype, precision, out_dtype):
            if split and y.numel() >= split_min:
                w12 = h.two_site_block_tensor(n1, n2, nb1, nb2, dtype)
                es = envs if dtype == torch.complex128 else [e.to(dtype) for e in envs]
                return E.CompiledNetwork([y, w12] + es, split_l, out, precision=precision,
                                         out_dtype=out_dtype, identity=ident)
            es = envs if dtype == torch.complex128 else [e.to(dtype) for e in envs]
            return E.CompiledNetwork([y, h.tensor_in(n1, dtype), h.tensor_in(n2, dtype)] + es,
                                     dense_l, out, precision=precision, out_dtype=out_dtype)
        # (DMRG borrows this method: its objects have no relaxed form)
        if getattr(self, "relaxed", False):
            min_size = getattr(self.config, "relax_min_size", 0)

            def make(y):
                high = net(y, dt, self.precision, torch.complex128)
                if y.numel() < min_size:
                    return high
                low = net(y, torch.complex64, self._low_precision(), torch.complex64)
                return self._relaxed(high, low, (low.tensors, low.labels, out), low.identity)
            return make, k1, k2
        return (lambda y: net(y, dt, self.precision, torch.complex128)), k1, k2

    def _update_two_site_nodes(self, target_node_id: str, next_node_id: str,
                               time_step_factor: float = 1):
        return self._drive(self._update_two_site_nodes_gen(target_node_id, next_node_id,
                                                           time_step_factor))

    def _update_two_site_nodes_gen(self, target_node_id: str, next_node_id: str,
                                   time_step_factor: float = 1):
        s = self.state
        n1, n2 = target_node_id, next_node_id
        assert s.orthogonality_center_id == n1
        x = merge_two_sites(s, n1, n2)
        make, _, _ = self._two_site_action(n1, n2)
        x = yield from self._evolve(x, make(x), self.time_step_size * time_step_factor,
                                    EvoDirection.FORWARD)
        self.apply_log.append(("two_site", (n1, n2), self._last_applies))
        # truncated SVD, S absorbed into V (VCONTR); the dense SVD is handed out as
        # an SvdRequest (a BatchedTDVP answers all instances' in one launch)
        s_new = yield from split_two_sites_gen(s, x, n1, n2, self.config)
        self.kept_dims.append(len(s_new))
        s.orthogonality_center_id = n2
        self.update_tree_cache(n1, n2)

    def _single_site_backwards_update(self, node_id: str, time_step_factor: float = 1):
        return self._drive(self._single_site_backwards_update_gen(node_id, time_step_factor))

    def _single_site_backwards_update_gen(self, node_id: str, time_step_factor: float = 1):
        yield from self._update_site_gen(node_id, -1 * time_step_factor)

    @staticmethod
    def create_two_site_id(node_id: str, next_node_id: str) -> str:
        return "TwoSite_" + node_id + "_contr_" + next_node_id


class FirstOrderTwoSiteTDVP(TwoSiteTDVP):
    """First-order two-site TDVP (``tdvp_algorithms/firstordertwosite.py:9-30``): one
    sweep along the update path at the full time step -- per pair of the path the
    two-site forward update, then the backward single-site update of the next
    site, the last pair updated in one go -- after which the orthogonality centre
    returns to the start of the path (KEEP QR).

    The reference loop names its pairs as (update_path[i],
    orthogonalization_path[i+1][0]) and moves the centre along
    orthogonalization_path[i] after the pair's update; on every tree that pairs
    nodes which are not neighbours, so it cannot run as written.  This class runs
    the sweep its docstring describes, with the pairs and moves of the
    second-order two-site forward sweep, and updates the environment cache on
    the way back so the next step starts from a complete cache."""

    def _time_step_gen(self):
        up, op = self.update_path, self.orthogonalization_path
        for i, node_id in enumerate(up[:-2]):
            if i > 0:
                self._move_orth_and_update_cache_for_path(op[i - 1])
            nxt = op[i][0]
            yield from self._update_two_site_nodes_gen(node_id, nxt)
            yield from self._single_site_backwards_update_gen(nxt)
        yield from self._update_two_site_nodes_gen(up[-2], up[-1])
        self._move_orth_and_update_cache_for_path(
            self.state.path_from_to(self.state.orthogonality_center_id, up[0]))


class SecondOrderTwoSiteTDVP(TwoSiteTDVP):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.backwards_update_path = list(reversed(self.update_path))
        self.backwards_orth_path = [list(reversed(p))
                                    for p in reversed(self.orthogonalization_path)]

    def forward_sweep(self):
        return self._drive(self._forward_sweep_gen())

    def _forward_sweep_gen(self):
        up, op = self.update_path, self.orthogonalization_path
        for i, node_id in enumerate(up[:-2]):
            if i > 0:
                self._move_orth_and_update_cache_for_path(op[i - 1])
            nxt = op[i][0]
            yield from self._update_two_site_nodes_gen(node_id, nxt, time_step_factor=0.5)
            yield from self._single_site_backwards_update_gen(nxt, time_step_factor=0.5)
        yield from self._update_two_site_nodes_gen(up[-2], up[-1], time_step_factor=0.5)

    def backwards_sweep(self):
        return self._drive(self._backwards_sweep_gen())

    def _backwards_sweep_gen(self):
        bp, bo = self.backwards_update_path, self.backwards_orth_path
        for i in range(len(bp) - 1):
            if i == 0:
                yield from self._update_two_site_nodes_gen(bp[0], bp[1], time_step_factor=0.5)
            else:
                path = bo[i]
                self._move_orth_and_update_cache_for_path(path)
                target = path[-1]
                yield from self._single_site_backwards_update_gen(target, time_step_factor=0.5)
                yield from self._update_two_site_nodes_gen(target, bp[i + 1],
                                                           time_step_factor=0.5)

    def _time_step_gen(self):
        yield from self._forward_sweep_gen()
        yield from self._backwards_sweep_gen()


# ---------------------------------------------------------------------------------------
# batched TDVP: independent instances stepped together
# ---------------------------------------------------------------------------------------
def _high(action):
    """The FP64 network of a relaxed action (the action itself otherwise)."""
    return action.high if isinstance(action, E.RelaxedNetwork) else action


class BatchedTDVP:
    """Many TDVP runs on the same tree (disorder realisations, parameter sweeps)
    advanced in lockstep.

    Every instance is an ordinary TDVP algorithm object (its own state, TTNO,
    environment cache and truncation decisions, so bond dimensions may differ
    between instances).  All instances share the tree, hence the update and
    orthogonalisation paths, the time step and the solver options: their sweeps
    reach the same Krylov exponentials in the same order.  Each such exponential
    runs for all instances at once (``krylov_exp_action_batched``): the H_eff
    applies as grouped launches over instances with equal shapes
    (``engine.BatchedNetwork``), the Arnoldi vector work as batched launches, one
    host round trip per Krylov step.  Splits, truncation and environment updates
    run per instance between those points.  Each instance follows exactly the
    single-instance algorithm.  ``lockstep_group`` steps the instances in groups of
    that many, one group after another, to bound the peak memory (512 instances of
    the 31-node chi = 128 configuration need groups of 256 on one 288 GB GPU).
    """

    # Krylov requests with local spaces of at least SEQUENTIAL_N elements are solved
    # instance by instance: at the 31-node random-field Heisenberg configuration
    # (chi <= 128, 32 instances) the batched solves of the 4.2M-element two-site
    # spaces took 61 s per step against 15 s one by one
    SEQUENTIAL_N = 1 << 18

    def __init__(self, algorithms: Sequence[TDVPAlgorithm]):
        algs = list(algorithms)
        if not algs:
            raise ValueError("BatchedTDVP needs at least one instance")
        a0 = algs[0]
        for a in algs[1:]:
            if type(a) is not type(a0) or a.update_path != a0.update_path or \
                    a.time_step_size != a0.time_step_size or a.precision != a0.precision or \
                    a.krylov_options() != a0.krylov_options():
                raise ValueError("batched instances must share the algorithm, tree, time "
                                 "step, precision and solver options")
        if len({a.relaxed for a in algs}) > 1:
            raise ValueError("batched instances must share the precision")
        self.algorithms = algs
        self.krylov_dims: List[int] = []
        self.memory_fraction = 0.5   # share of free device memory a batched solve may use
        self.max_batch: Optional[int] = None
        # called as progress(points_done, kind) after every lockstep point of a step
        # ("krylov" or "svd"): for callers that report long steps as they go
        self.progress = None
        # instances stepped together: every instance of a lockstep group holds its
        # merged two-site tensor and QR factors while the group's SVDs are answered,
        # so groups bound the peak memory (None: all instances in one group)
        self.lockstep_group: Optional[int] = None

    def __len__(self):
        return len(self.algorithms)

    def _instance_bytes(self, req: KrylovRequest, max_dim: Optional[int]) -> int:
        """Device bytes one instance needs inside a batched Krylov solve: the
        basis, the vectors of one step and the live intermediates of its H_eff
        network (planned order)."""
        from .krylov import KRYLOV_START_CAP
        net = _high(req.action)
        n = req.x.numel()
        # the basis starts at KRYLOV_START_CAP vectors and doubles only while the
        # recurrence runs on; with no max_krylov_dim the budget takes twice that
        m = min(n, int(max_dim) if max_dim is not None else 2 * KRYLOV_START_CAP)
        dims: Dict[Any, int] = {}
        for t, lab in zip(net.tensors, net.labels):
            t = req.x if t is None else t
            for l, d in zip(lab, t.shape):
                dims[l] = d
        live = E.network_live_peak(net.labels, dims, net.out_labels)
        esize = torch.empty((), dtype=E.storage_dtype(net.precision)).element_size()
        return 16 * n * (m + 5) + esize * live

    def _solve(self, reqs: List[KrylovRequest],
               algorithms: Optional[Sequence[TDVPAlgorithm]] = None):
        from .krylov import krylov_exp_action_batched
        algorithms = self.algorithms if algorithms is None else algorithms
        a0 = algorithms[0]
        if a0._mode is not TimeEvoMode.KRYLOV and not (
                a0._mode is TimeEvoMode.FASTEST and reqs[0].x.numel() > a0.DENSE_MAX_DIM):
            # dense exponential modes: one dense solve per instance
            for a, r in zip(algorithms, reqs):
                r.result, r.applies = a._krylov(r.x, r.action, r.dt, r.direction), 0
            return
        # large local spaces: one instance at a time on the single-instance path
        # (its LARGE GEMMs fill the chip; the grouped launches and the batched
        # buffers only add memory traffic and allocations)
        small = []
        for a, r in zip(algorithms, reqs):
            if r.x.numel() >= self.SEQUENTIAL_N:
                from .krylov import STATS
                n0, p0 = STATS.applies, _K.pending_count()
                r.result = a._krylov(r.x, r.action, r.dt, r.direction)
                # a one-launch solve inside a deferred step hands out its Krylov
                # dimension when the step ends (resolved by _lockstep_checked)
                r.applies = (_K.PendingApplies(p0) if _K.pending_count() > p0
                             else STATS.applies - n0)
                self.krylov_dims.append(r.applies)
                # every instance's network stays referenced until all are answered:
                # drop this one's recorded intermediates now
                if hasattr(r.action, "release"):
                    r.action.release()
            else:
                small.append(r)
        reqs = small
        if not reqs:
            return
        opts = dict(self.algorithms[0].krylov_options())
        direction = reqs[0].direction
        override = opts.pop("exp_factor", None)
        exp_factor = (direction.exp_factor() if override is None
                      else (override if direction is EvoDirection.FORWARD else -override))
        max_dim = opts.get("max_krylov_dim", None)
        # instances per launch: bounded by a share of the free device memory, so the
        # batched buffers never crowd out the factorisation libraries' workspaces
        free, total = torch.cuda.mem_get_info()
        cached = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
        budget = self.memory_fraction * (free + cached)
        per = max(self._instance_bytes(r, max_dim) for r in reqs)
        chunk = max(1, min(len(reqs), int(budget // per)))
        if self.max_batch is not None:
            chunk = min(chunk, self.max_batch)
        dims: List[int] = []
        for c0 in range(0, len(reqs), chunk):
            part = reqs[c0:c0 + chunk]
            # relaxed instances: the local spaces solved here (below SEQUENTIAL_N)
            # lie below relax_min_size, where the relaxed form runs FP64 anyway
            net = E.BatchedNetwork([_high(r.action) for r in part])
            res = krylov_exp_action_batched([r.x for r in part], net, part[0].dt, exp_factor,
                                            krylov_tol=opts.get("krylov_tol", 1e-10),
                                            max_krylov_dim=max_dim,
                                            hermitian=opts.get("hermitian", True), stats=dims)
            for r, y in zip(part, res):
                r.result = y
            del net, res
            if chunk < len(reqs) or per * len(part) > 0.25 * total:
                torch.cuda.empty_cache()
        for r, m in zip(reqs, dims):
            r.applies = m
        self.krylov_dims.extend(dims)

    def run_one_time_step(self):
        E.refresh_stream()
        for a in self.algorithms:
            a.dense_site_below = self.SEQUENTIAL_N
        size = self.lockstep_group or len(self.algorithms)
        for g0 in range(0, len(self.algorithms), size):
            self._lockstep_checked(self.algorithms[g0:g0 + size])

    def _lockstep_checked(self, algorithms: Sequence[TDVPAlgorithm]):
        """One lockstep group's time step with the CholeskyQR pivot checks and the
        one-launch Krylov infos read once at its end, as a single instance's step
        does (``TDVPAlgorithm.run_one_time_step``).  Each instance's split statuses
        are kept apart: an instance with a collapsed split is restored and steps
        again on its own (the single-instance path, which checks its splits as it
        goes when it has to); the other instances keep their batched step."""
        from .splitting import deferred_qr_checks
        if not TDVPAlgorithm.DEFER_QR_CHECKS:
            self._lockstep(algorithms)
            return
        snaps = [a._snapshot() for a in algorithms]
        nlogs = [len(a.apply_log) for a in algorithms]
        nk = len(self.krylov_dims)
        flags: List[List[torch.Tensor]] = [[] for _ in algorithms]
        err = None
        with deferred_qr_checks(), _K.deferred_small_info() as small:
            try:
                self._lockstep(algorithms, flags)
            except Exception as exc:      # downstream of a collapsed split, possibly
                err = exc
            bad = [False] * len(algorithms)
            allf = [(i, f) for i, fl in enumerate(flags) for f in fl]
            if allf:
                vals = torch.cat([f for _, f in allf]).view(len(allf), -1).sum(1).cpu()
                for (i, _), v in zip(allf, vals.tolist()):
                    bad[i] = bad[i] or v != 0
            if err is not None and not any(bad):
                raise err
            if err is None:
                ms = _K.resolve_pending(small)
                for a, nlog in zip(algorithms, nlogs):
                    a.apply_log[nlog:] = [
                        (k, n, ms[x.index] if isinstance(x, _K.PendingApplies) else x)
                        for k, n, x in a.apply_log[nlog:]]
                    if isinstance(a._last_applies, _K.PendingApplies):
                        a._last_applies = ms[a._last_applies.index]
                self.krylov_dims[nk:] = [ms[x.index] if isinstance(x, _K.PendingApplies)
                                         else x for x in self.krylov_dims[nk:]]
        if err is not None:
            # which instance failed downstream is unknown: the group steps again with
            # every split checked as it happens
            TDVPAlgorithm.steps_redone += 1
            for a, snap in zip(algorithms, snaps):
                a._restore(snap)
            del self.krylov_dims[nk:]
            self._lockstep(algorithms)
            return
        for a, snap, b in zip(algorithms, snaps, bad):
            if b:
                TDVPAlgorithm.steps_redone += 1
                a._restore(snap)
                a.run_one_time_step()

    def _lockstep(self, algorithms: Sequence[TDVPAlgorithm],
                  flags: Optional[List[List[torch.Tensor]]] = None):
        """Advance all instances' sweeps together.  ``flags``: per instance, the
        list its CholeskyQR splits append their deferred statuses to."""
        from .splitting import _DEFER
        gens = [a._time_step_gen() for a in algorithms]

        def advance(i, first=False):
            if flags is not None:
                _DEFER.flags = flags[i]
            try:
                return next(gens[i]) if first else gens[i].send(None)
            except StopIteration:
                return None
        reqs = [advance(i, True) for i in range(len(gens))]
        done = 0
        while True:
            if all(r is None for r in reqs):
                return
            if any(r is None for r in reqs):
                raise RuntimeError("batched instances left the common sweep schedule")
            if isinstance(reqs[0], SvdRequest):
                if not all(isinstance(r, SvdRequest) for r in reqs):
                    raise RuntimeError("batched instances left the common sweep schedule")
                answer_svd_batched(reqs)
                kind = "svd"
            else:
                self._solve(reqs, algorithms)
                kind = "krylov"
            done += 1
            if self.progress is not None:
                self.progress(done, kind)
            reqs = [advance(i) for i in range(len(gens))]

    def run(self, steps: int):
        for _ in range(steps):
            self.run_one_time_step()
