"""BUG and fixed-rank BUG on the device against the oracle's restatement, and
against exact evolution where the bases are complete."""
import numpy as np
import pytest
import torch
from scipy.linalg import expm

import pytreenet_amd as ptn
from oracle import bug as OB
from oracle import dense as OD
from oracle.tdvp import OracleState
from pytreenet_amd.bug import BUG, BUGConfig, FixedBUG, FixedBUGConfig
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import PAULI_Z
from tests.helpers import host_tensors, topology_from_ttn

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _setup(branching, depth, chi, seed):
    ids, parent, children = complete_tree(branching, depth)
    tensors = random_state_tensors(ids, parent, children, chi, seed)
    st = build_state(ids, parent, children, tensors, device="cuda")
    st.canonical_form(st.root_id, mode=ptn.SplitMode.KEEP)
    c = st.root_id
    st.tensors[c] = st.tensors[c] / st.norm()
    H = ptn.random_xyz_model(st, seed)
    topo = topology_from_ttn(st)
    ost = OracleState(topo, host_tensors(st))
    ost.center = c
    W = host_tensors(H)
    return st, H, topo, ost, W


def _dense(st):
    return OD.state_to_vector(topology_from_ttn(st), host_tensors(st))


def _shapes(st):
    return {k: tuple(v.shape) for k, v in st.tensors.items()}


@pytest.mark.parametrize("fixed", [False, True])
def test_bug_exact_at_full_bond_dims(gpu, fixed):
    _say(f"bug full bond dims fixed={fixed}")
    st, H, topo, ost, W = _setup(2, 2, 64, 3)
    psi0 = _dense(st)
    Hd = OD.ttno_to_matrix(topo, W)
    alg = (FixedBUG(st, H, 0.05, 0.1, [], config=FixedBUGConfig()) if fixed else
           BUG(st, H, 0.05, 0.1, [], config=BUGConfig()))
    for _ in range(2):
        alg.run_one_time_step()
    psi = _dense(alg.state)
    np.testing.assert_allclose(psi, expm(-0.1j * Hd) @ psi0, rtol=0, atol=1e-10)
    assert alg.state.orthogonality_center_id == alg.state.root_id


@pytest.mark.parametrize("tree,params", [((2, 2, 2, 11), (3, 1e-10, 1e-10)),
                                         ((2, 3, 3, 12), (4, 1e-8, 1e-8)),
                                         ((3, 2, 3, 13), (5, 1e-8, 1e-8))])
def test_bug_matches_oracle_below_full_rank(gpu, tree, params):
    """Rank-adaptive BUG: bases grow, the truncation sweep cuts them back; the
    device state equals the oracle's (dense vectors) with the same bond dims."""
    _say(f"bug tree={tree} params={params}")
    st, H, topo, ost, W = _setup(*tree)
    cfg = BUGConfig(max_bond_dim=params[0], rel_tol=params[1], total_tol=params[2])
    alg = BUG(st, H, 0.05, 0.1, [], config=cfg)
    opar = dict(max_bond_dim=params[0], rel_tol=params[1], total_tol=params[2])
    for step in range(2):
        alg.run_one_time_step()
        ost = OB.bug_time_step(ost, W, 0.05, opar)
        ref = OD.state_to_vector(topo, ost.A)
        got = _dense(alg.state)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max(),
                                   err_msg=f"step {step}")
        assert _shapes(alg.state) == {k: v.shape for k, v in ost.A.items()}


@pytest.mark.parametrize("tree", [(2, 2, 2, 21), (3, 2, 3, 22)])
def test_fixed_bug_matches_oracle_below_full_rank(gpu, tree):
    _say(f"fixed bug tree={tree}")
    st, H, topo, ost, W = _setup(*tree)
    alg = FixedBUG(st, H, 0.05, 0.1, [], config=FixedBUGConfig())
    shapes0 = _shapes(alg.state)
    for step in range(2):
        alg.run_one_time_step()
        ost = OB.bug_time_step(ost, W, 0.05, fixed_rank=True)
        ref = OD.state_to_vector(topo, ost.A)
        np.testing.assert_allclose(_dense(alg.state), ref, rtol=0,
                                   atol=1e-9 * np.abs(ref).max(), err_msg=f"step {step}")
    assert _shapes(alg.state) == shapes0


def test_bug_run_records_results(gpu):
    """BUG.run evaluates the operators through the device state; energies stay
    close to the initial energy (small dt, truncation at 1e-10)."""
    _say("bug run")
    st, H, topo, ost, W = _setup(2, 2, 4, 31)
    z = ptn.TensorProduct({"site1": PAULI_Z})
    alg = BUG(st, H, 0.02, 0.06, {"z": z, "H": H}, config=BUGConfig(max_bond_dim=8))
    alg.run()
    e = alg.results.operator_result("H")
    assert len(e) == 4 and np.max(np.abs(e - e[0])) < 1e-6
    psi = _dense(alg.state)
    zfull = OD.kron_op(len(topo.node_ids), {topo.node_ids.index("site1"): PAULI_Z})
    assert abs(alg.results.operator_result("z")[-1] - np.vdot(psi, zfull @ psi)) < 1e-10


def test_reduced_qr_moves_shrink_oversized_bonds(gpu):
    """A bond wider than the product of a node's other dimensions: REDUCED moves
    (the reference's default) shrink it, KEEP moves keep it; the state is the same."""
    _say("reduced qr moves")
    rng = np.random.default_rng(0)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    ids = ["r", "a", "b"]
    parent = {"r": None, "a": "r", "b": "r"}
    children = {"r": ["a", "b"], "a": [], "b": []}
    tensors = {"r": c(4, 3, 2), "a": c(4, 2), "b": c(3, 2)}
    psi = None
    for mode, shapes in [(ptn.SplitMode.REDUCED, {"a": (2, 2), "b": (2, 2), "r": (2, 2, 2)}),
                         (ptn.SplitMode.KEEP, {"a": (4, 2), "b": (3, 2), "r": (4, 3, 2)})]:
        st = build_state(ids, parent, children, tensors, device="cuda")
        st.canonical_form("r", mode=mode)
        assert _shapes(st) == shapes
        v = _dense(st)
        if psi is None:
            psi = v
        np.testing.assert_allclose(v, psi, rtol=0, atol=1e-13 * np.abs(psi).max())
        np.testing.assert_allclose(v, OD.state_to_vector(topology_from_ttn(st), tensors),
                                   rtol=0, atol=1e-13 * np.abs(psi).max())
