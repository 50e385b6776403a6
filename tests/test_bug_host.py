"""BUG on the host side: the leaf-to-leaf path of the truncation sweep against
the reference tests' expected paths, and the oracle's BUG pinned by exactness at
full bond dimensions, norm and energy conservation of the rank-adaptive
Galerkin step, and the Eckart-Young optimality of the truncation."""
import json
import os

import numpy as np
import pytest
import torch
from scipy.linalg import expm

import pytreenet_amd as ptn
from oracle import bug as OB
from oracle import contract as OC
from oracle import dense as OD
from oracle.tdvp import OracleState
from oracle.tree import Topology
from pytreenet_amd.bug import find_path_leaf_to_leaf
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import build_edge_ttno_tensors
from tests.helpers import oracle_problem

GOLD = os.path.join(os.path.dirname(__file__), "golden")
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
CASES = json.load(open(os.path.join(GOLD, "leaf_to_leaf_paths.json")))["cases"]


def _tree_from_edges(case):
    topo = Topology.from_parent_map(case["root"], [tuple(e) for e in case["edges"]])
    st = ptn.TreeTensorNetworkState()
    st.add_root(ptn.Node(identifier=case["root"]), torch.zeros([1] * (len(topo.children[case["root"]]) + 1)))
    for child, parent in case["edges"]:
        nch = len(topo.children[child])
        leg = st.nodes[parent].nvirt_legs()
        st.add_child_to_parent(ptn.Node(identifier=child), torch.zeros([1] * (nch + 2)), 0,
                               parent, leg)
    return topo, st


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_leaf_to_leaf_paths_match_reference_kats(case):
    topo, st = _tree_from_edges(case)
    assert OB.leaf_to_leaf_path(topo) == case["path"]
    assert find_path_leaf_to_leaf(st) == case["path"]
    assert topo.path_from_to(*case["diameter_leaves"]) == case["main_path"]


@pytest.mark.parametrize("fixed", [False, True])
def test_oracle_bug_exact_at_full_bond_dims(fixed):
    """Complete bases (7-site binary tree, bonds at their caps): both integrators
    equal exact evolution."""
    topo, st, W, H = oracle_problem(2, 2, 64, 5)
    psi0 = OD.state_to_vector(topo, st.A)
    for _ in range(2):
        st = OB.bug_time_step(st, W, 0.05, dict(max_bond_dim=100, rel_tol=1e-15,
                                                total_tol=1e-15), fixed_rank=fixed)
    psi = OD.state_to_vector(topo, st.A)
    np.testing.assert_allclose(psi, expm(-1j * 0.1 * H) @ psi0, atol=1e-11)
    assert {st.A[n].shape[0] for n in topo.node_ids if topo.parent[n] == topo.root} == {8}


def test_oracle_bug_galerkin_step_conserves_norm_and_energy():
    """Below full rank the rank-adaptive Galerkin step (before truncation) stays in
    a subspace that contains the initial state: norm and energy are conserved,
    and the bonds grow (at most doubled)."""
    topo, st, W, H = oracle_problem(2, 2, 2, 7)    # 7 sites, level-1 bonds 2 of 8
    psi0 = OD.state_to_vector(topo, st.A)
    e0 = np.vdot(psi0, H @ psi0).real
    out = OB.bug_step(st, W, 0.1)
    psi = OD.state_to_vector(topo, out.A)
    assert abs(np.linalg.norm(psi) - 1) < 1e-12
    assert abs(np.vdot(psi, H @ psi).real - e0) < 1e-11
    grown = [out.A[n].shape[0] for n in topo.node_ids if topo.parent[n] is not None]
    before = [st.A[n].shape[0] for n in topo.node_ids if topo.parent[n] is not None]
    assert all(b <= g <= 2 * b for b, g in zip(before, grown)) and grown != before
    # it is closer to exact evolution than the fixed rank step
    exact = expm(-0.1j * H) @ psi0
    fx = OD.state_to_vector(topo, OB.bug_step(st, W, 0.1, fixed_rank=True).A)
    assert np.linalg.norm(psi - exact) < np.linalg.norm(fx - exact)


def test_oracle_svd_truncation_is_best_low_rank():
    """Two sites: the truncation sweep keeps the best rank-r approximation."""
    rng = np.random.default_rng(3)
    topo = Topology.from_parent_map("a", [("b", "a")])
    A = {"a": rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3)),
         "b": rng.standard_normal((4, 5)) + 1j * rng.standard_normal((4, 5))}
    st = OracleState(topo, A)
    st.canonical_form("a")
    psi = OD.state_to_vector(topo, st.A).reshape(3, 5)
    OB.svd_truncation(st, dict(max_bond_dim=2, rel_tol=-np.inf, total_tol=-np.inf))
    OB.move_center(st, "a", keep=True)
    u, s, vh = np.linalg.svd(psi)
    best = (u[:, :2] * s[:2]) @ vh[:2]
    np.testing.assert_allclose(OD.state_to_vector(topo, st.A).reshape(3, 5), best, atol=1e-12)
    assert st.A["b"].shape == (2, 5)
