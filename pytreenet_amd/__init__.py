"""pytreenet_amd -- MI355X-native TDVP hot path with PyTreeNet's call surface.

The public names mirror the reference package's star exports
(``pytreenet/__init__.py``) for the time-evolution path: TTNS/TTNO objects,
the environment cache and contractions, the Krylov exponential, QR/SVD splits,
the TDVP algorithms, TEBD, the BUG integrators and DMRG.  All arithmetic runs in libptnhip.so (HIP, gfx950);
there is no CPU fallback.
"""
from ._lib import PtnBackendError
from .ttn import Node, TreeTensorNetwork, TreeTensorNetworkState, TTNS
from .ttno import (NumericOperator, TensorProduct, TreeTensorNetworkOperator, TTNO, pauli_matrices,
                   transverse_field_ising, xyz_model, xxz_model, random_xyz_model,
                   random_field_heisenberg)
from .contractions import (PartialTreeCachDict, SandwichCache, contract_any,
                           contract_ket_ham_with_envs, contract_bond_tensor, expectation_value,
                           contract_two_ttns, scalar_product)
from .krylov import EvoDirection, TimeEvoMode, krylov_exp_action
from .splitting import (SplitMode, SVDParameters, ContractionMode, tensor_qr_decomposition,
                        tensor_svd, truncated_tensor_svd, truncate_singular_values,
                        contr_truncated_svd_splitting, value_truncation, sum_truncation,
                        renormalise_singular_values)
from .time_evolution import (TDVPConfig, TwoSiteTDVPConfig, TDVPAlgorithm, OneSiteTDVP,
                             FirstOrderOneSiteTDVP, SecondOrderOneSiteTDVP, TwoSiteTDVP,
                             FirstOrderTwoSiteTDVP,
                             SecondOrderTwoSiteTDVP, Results, find_orthogonalisation_path,
                             BatchedTDVP, TTNTimeEvolution)
from .tebd import TEBD, SWAPlist, TrotterSplitting, TrotterStep, swap_gate
from .bug import (BUG, FixedBUG, BUGConfig, FixedBUGConfig, CommonBUGConfig, svd_truncation,
                  find_path_leaf_to_leaf)
from .dmrg import DMRGAlgorithm
from .effective_hamiltonians import (find_tensor_leg_permutation, contract_all_except_node,
                                     get_effective_single_site_hamiltonian,
                                     get_effective_two_site_hamiltonian,
                                     get_effective_bond_hamiltonian)
from .random_states import random_tree_state, complete_tree, build_state
from . import engine

__all__ = [n for n in dir() if not n.startswith("_")]
