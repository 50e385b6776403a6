"""TEBD: Trotterised time evolution of a TTNS by one- and two-site gates.

Counterparts of ``time_evolution/trotter.py`` (``SWAPlist`` ``:15-103``,
``TrotterStep`` ``:105-194``, ``TrotterSplitting`` ``:196-322``) and
``time_evolution/tebd.py`` (``TEBD`` ``:15-155``).

The gates are small host matrices, exponentiated once at construction as the
reference does, and copied to the device once.  Every time step then runs on
the device: a one-site gate is one contraction into the node's physical leg; a
two-site gate contracts the two node tensors over their bond, applies the gate
to both physical legs, and splits the result again by the device SVD with the
reference truncation rules (S absorbed into the second node's tensor).  Each
node keeps its leg order (parent, children..., phys) through the split, where
the reference re-inserts the split nodes and may move a child leg to the end
of its parent's children; the states are equal up to that leg order.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import engine as E
from .splitting import SVDParameters, merge_two_sites, split_two_sites
from .time_evolution import TTNTimeEvolution
from .ttno import NumericOperator, TensorProduct


def swap_gate(dimension: int = 2) -> np.ndarray:
    """The SWAP of two d-dimensional sites as a (d^2, d^2) matrix: output
    (x, y) = input (y, x) (``operators/common_operators.py:89-112``)."""
    if dimension <= 0:
        raise ValueError("dimension has to be positive")
    d = dimension
    eye = np.eye(d * d, dtype=complex).reshape(d, d, d * d)
    return eye.transpose(1, 0, 2).reshape(d * d, d * d)


class SWAPlist(list):
    """Pairs of neighbouring nodes whose states are swapped, in order
    (``trotter.py:15-103``)."""

    def __init__(self, swap_list: Optional[Sequence[Tuple[str, str]]] = None):
        swap_list = [] if swap_list is None else list(swap_list)
        for pair in swap_list:
            if len(pair) != 2:
                raise ValueError("SWAPs can only happen between exactly two nodes!")
        super().__init__(swap_list)

    def is_compatible_with_ttn(self, ttn) -> bool:
        """Every pair is two neighbouring nodes of ``ttn`` with equal open dimension."""
        for a, b in self:
            if a not in ttn.nodes:
                return False
            node_a = ttn.nodes[a]
            if b not in node_a.neighbouring_nodes():
                return False
            if node_a.open_dimension() != ttn.nodes[b].open_dimension():
                return False
        return True

    def into_operators(self, ttn=None, dim: Optional[int] = None) -> List[NumericOperator]:
        """The SWAP gates as (out_a, out_b, in_a, in_b) tensors."""
        if ttn is None and dim is None and len(self) != 0:
            raise ValueError("`ttn` and `dim` cannot both be `None`!")
        out = []
        for pair in self:
            d = ttn.nodes[pair[0]].open_dimension() if ttn is not None else dim
            out.append(NumericOperator(swap_gate(d), list(pair)).to_tensor(dim=d, ttn=ttn))
        return out


class TrotterStep:
    """One factor exp(-i factor dt op) of a splitting with the SWAPs around it
    (``trotter.py:105-194``)."""

    def __init__(self, operator: TensorProduct, factor: float,
                 swaps_before: Union[SWAPlist, Sequence, None] = None,
                 swaps_after: Union[SWAPlist, Sequence, None] = None):
        self.operator = operator
        self.factor = factor
        self.swaps_before = SWAPlist() if swaps_before is None else swaps_before
        self.swaps_after = SWAPlist() if swaps_after is None else swaps_after

    def exponentiate_operator(self, delta_time: float, ttn=None,
                              dim: Optional[int] = None) -> NumericOperator:
        return self.operator.exp(-1j * self.factor * delta_time).to_tensor(dim=dim, ttn=ttn)

    def realise_swaps(self, ttn=None, dim: Optional[int] = None):
        before = SWAPlist(self.swaps_before).into_operators(ttn, dim)
        after = SWAPlist(self.swaps_after).into_operators(ttn, dim)
        return before, after


class TrotterSplitting(list):
    """An ordered list of Trotter steps (``trotter.py:196-322``)."""

    def __init__(self, trotter_steps: Optional[Sequence[TrotterStep]] = None):
        super().__init__([] if trotter_steps is None else list(trotter_steps))

    @classmethod
    def from_lists(cls, tensor_products: Sequence[TensorProduct],
                   splitting: Optional[Sequence[Union[Tuple[int, float], int]]] = None,
                   swaps_before: Optional[Sequence[SWAPlist]] = None,
                   swaps_after: Optional[Sequence[SWAPlist]] = None) -> "TrotterSplitting":
        """Steps from (index, factor) pairs into ``tensor_products`` (a bare index
        means factor 1; no splitting means every product once, factor 1).  The
        SWAP lists are indexed like ``tensor_products``."""
        if splitting is None:
            splitting = [(i, 1) for i in range(len(tensor_products))]
        steps = []
        for entry in splitting:
            index, factor = (entry, 1) if isinstance(entry, int) else (entry[0], entry[1])
            before = SWAPlist() if swaps_before is None else swaps_before[index]
            after = SWAPlist() if swaps_after is None else swaps_after[index]
            steps.append(TrotterStep(tensor_products[index], factor, swaps_before=before,
                                     swaps_after=after))
        return cls(steps)

    def exponentiate_splitting(self, delta_time: float, ttn=None,
                               dim: Optional[int] = None) -> List[NumericOperator]:
        """Every gate of one time step in application order, SWAPs included."""
        out: List[NumericOperator] = []
        for step in self:
            gate = step.exponentiate_operator(delta_time, dim=dim, ttn=ttn)
            before, after = step.realise_swaps(dim=dim, ttn=ttn)
            out.extend(before)
            out.append(gate)
            out.extend(after)
        return out


class TEBD(TTNTimeEvolution):
    """TEBD on a TTNS (``time_evolution/tebd.py:15-155``)."""

    def __init__(self, initial_state, trotter_splitting: TrotterSplitting,
                 time_step_size: float, final_time: float, operators,
                 svd_parameters: Optional[SVDParameters] = None, config=None):
        super().__init__(initial_state, time_step_size, final_time, operators, config=config)
        self._trotter_splitting = trotter_splitting
        self.svd_parameters = SVDParameters() if svd_parameters is None else svd_parameters
        self._exponents = trotter_splitting.exponentiate_splitting(self._time_step_size,
                                                                   self.state)
        E.require_device(*self.state.tensors.values())
        dev = self.state.device
        self._gates: Dict[int, torch.Tensor] = {}
        for exp_op in self._exponents:
            self._gates[id(exp_op)] = torch.as_tensor(
                np.asarray(exp_op.operator, dtype=np.complex128)).to(dev)
        self.kept_dims: List[int] = []

    @property
    def exponents(self) -> List[NumericOperator]:
        return self._exponents

    @property
    def trotter_splitting(self) -> TrotterSplitting:
        return self._trotter_splitting

    def _gate(self, unitary: NumericOperator) -> torch.Tensor:
        g = self._gates.get(id(unitary))
        if g is None:
            g = torch.as_tensor(np.asarray(unitary.operator, dtype=np.complex128)).to(
                self.state.device)
        return g

    def _apply_one_trotter_step_single_site(self, single_site_exponent: NumericOperator):
        """state <- U_n state: the gate's input leg contracted with the node's
        physical leg (``TTN.absorb_into_open_legs``, ``core/ttn.py:929-964``)."""
        nid = single_site_exponent.node_identifiers[0]
        node, t = self.state[nid]
        g = self._gate(single_site_exponent)
        nv = t.dim() - 1
        if tuple(g.shape) != (t.shape[-1], t.shape[-1]):
            raise AssertionError("the operator's legs do not match the node's open legs")
        lab = list(range(nv)) + ["p"]
        self.state.tensors[nid] = E.contract(t, lab, g, ["q", "p"], list(range(nv)) + ["q"])
        if self.state.orthogonality_center_id != nid:
            self.state.orthogonality_center_id = None

    def _apply_one_trotter_step_two_site(self, two_site_exponent: NumericOperator):
        """Merge the two nodes, apply the gate to both physical legs, split by the
        truncated SVD (``tebd.py:98-120``)."""
        n1, n2 = two_site_exponent.node_identifiers
        g = self._gate(two_site_exponent)
        d1 = self.state.tensors[n1].shape[-1]
        d2 = self.state.tensors[n2].shape[-1]
        if tuple(g.shape) != (d1, d2, d1, d2):
            raise AssertionError("the operator's legs do not match the nodes' open legs")
        x = merge_two_sites(self.state, n1, n2)
        nv = x.dim() - 2
        outer = list(range(nv))
        x = E.contract(x, outer + ["p1", "p2"], g, ["q1", "q2", "p1", "p2"],
                       outer + ["q1", "q2"])
        s_kept = split_two_sites(self.state, x, n1, n2, self.svd_parameters)
        self.kept_dims.append(len(s_kept))
        self.state.orthogonality_center_id = None

    def _apply_one_trotter_step(self, unitary: NumericOperator):
        n = len(unitary.node_identifiers)
        if n == 0:
            return
        if n == 1:
            self._apply_one_trotter_step_single_site(unitary)
        elif n == 2:
            self._apply_one_trotter_step_two_site(unitary)
        else:
            raise NotImplementedError("More than two-site interactions are not implemented.")

    def run_one_time_step(self, **kwargs):
        E.refresh_stream()
        for unitary in self._exponents:
            self._apply_one_trotter_step(unitary)
