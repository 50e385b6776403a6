"""Dense effective Hamiltonians (oracle, test infrastructure only).

Restates ``contractions/effective_hamiltonians.py:20-470`` on plain neighbour
lists (the reference reads them from ``Node`` objects):

* ``find_tensor_leg_permutation`` (``:20-52``)
* ``contract_all_except_node`` (``:54-104``): the cached blocks are contracted
  into the Hamiltonian tensor one by one in the Hamiltonian node's neighbour
  order, each over (Hamiltonian leg 0, block leg 1), then permuted;
* ``effective_single_site_hamiltonian`` (``:106-149``), the half matricisation
  of ``util/tensor_util.py`` (``tensor_matricisation_half``);
* ``effective_bond_hamiltonian_tensor`` (``:151-202``);
* ``contract_all_but_one_block_to_hamiltonian``
  (``contractions/contraction_util.py:341-432``, ``determine_index_with_ignored_leg``);
* ``determine_two_site_leg_permutation`` (``:300-392``) and
  ``contract_all_except_two_nodes`` (``:242-298``).

A "Hamiltonian node" is given as its neighbour list (parent first, then its
children in the TTNO's own order), a "state node" as its neighbour list and
its number of open legs.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

Env = Dict[Tuple[str, str], np.ndarray]


def find_tensor_leg_permutation(state_nbs: Sequence[str], ham_nbs: Sequence[str],
                                n_open: int) -> Tuple[int, ...]:
    ham_index = [list(ham_nbs).index(nb) for nb in state_nbs]
    out_legs, in_legs = [], []
    for h in ham_index:
        out_legs.append(2 * n_open + 2 * h + 1)
        in_legs.append(2 * n_open + 2 * h)
    out_legs += list(range(n_open))
    in_legs += list(range(n_open, 2 * n_open))
    return tuple(out_legs + in_legs)


def contract_all_except_node(nid: str, state_nbs: Sequence[str], ham_nbs: Sequence[str],
                             n_open: int, W: np.ndarray, envs: Env) -> np.ndarray:
    t = W
    for nb in ham_nbs:
        t = np.tensordot(t, envs[(nb, nid)], axes=(0, 1))
    return np.transpose(t, find_tensor_leg_permutation(state_nbs, ham_nbs, n_open))


def half_matrix(t: np.ndarray) -> np.ndarray:
    half = t.ndim // 2
    return t.reshape(int(np.prod(t.shape[:half])), -1)


def effective_single_site_hamiltonian(nid, state_nbs, ham_nbs, n_open, W, envs) -> np.ndarray:
    return half_matrix(contract_all_except_node(nid, state_nbs, ham_nbs, n_open, W, envs))


def effective_bond_hamiltonian_tensor(parent: str, child: str, envs: Env) -> np.ndarray:
    t = np.tensordot(envs[(parent, child)], envs[(child, parent)], axes=(1, 1))
    return np.transpose(t, (1, 3, 0, 2))


def _index_with_ignored_leg(ham_nbs: Sequence[str], nb: str, ignored: str) -> int:
    """Current leg of neighbour ``nb`` once the blocks before it are contracted: 1 if
    the ignored neighbour's leg (still present) precedes it, else 0."""
    return 1 if list(ham_nbs).index(ignored) < list(ham_nbs).index(nb) else 0


def contract_all_but_one_block_to_hamiltonian(nid: str, ham_nbs: Sequence[str], W: np.ndarray,
                                              nxt: str, envs: Env) -> np.ndarray:
    """Legs (leg to ``nxt``, out, in, then (ket, bra) per contracted block)."""
    t = W
    for nb in ham_nbs:
        if nb == nxt:
            continue
        t = np.tensordot(t, envs[(nb, nid)],
                         axes=(_index_with_ignored_leg(ham_nbs, nb, nxt), 1))
    return t


def _block_leg_target(ham_nbs: Sequence[str], nxt: str, nb: str) -> int:
    index_next = list(ham_nbs).index(nxt)
    idx = list(ham_nbs).index(nb)
    return 2 * (idx + int(idx < index_next))


def determine_two_site_leg_permutation(state_nbs: Sequence[str], target: str,
                                       target_nbs: Sequence[str], nxt: str,
                                       next_nbs: Sequence[str]) -> Tuple[int, ...]:
    in_legs: List[int] = []
    for nb in state_nbs:
        if nb in target_nbs:
            in_legs.append(_block_leg_target(target_nbs, nxt, nb))
        elif nb in next_nbs:
            in_legs.append(2 * len(target_nbs) + _block_leg_target(next_nbs, target, nb))
        else:
            raise ValueError("The two-site Hamiltonian has a neighbour that is not a "
                             "neighbour of the two sites.")
    out_legs = [leg + 1 for leg in in_legs]
    nt = len(target_nbs)
    return tuple(out_legs + [0, 2 * nt] + in_legs + [1, 2 * nt + 1])


def contract_all_except_two_nodes(state_nbs, target: str, target_nbs, W_t: np.ndarray,
                                  nxt: str, next_nbs, W_n: np.ndarray, envs: Env) -> np.ndarray:
    bt = contract_all_but_one_block_to_hamiltonian(target, target_nbs, W_t, nxt, envs)
    bn = contract_all_but_one_block_to_hamiltonian(nxt, next_nbs, W_n, target, envs)
    h = np.tensordot(bt, bn, axes=(0, 0))
    return h.transpose(determine_two_site_leg_permutation(state_nbs, target, target_nbs, nxt,
                                                          next_nbs))


def effective_two_site_hamiltonian(state_nbs, target, target_nbs, W_t, nxt, next_nbs, W_n,
                                   envs) -> np.ndarray:
    return half_matrix(contract_all_except_two_nodes(state_nbs, target, target_nbs, W_t, nxt,
                                                     next_nbs, W_n, envs))
