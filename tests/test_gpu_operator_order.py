"""Device path with a TTNO whose children order differs from the state's, the dense
effective Hamiltonians, the interchange files through the oracle, and the public
entry points under a non-default torch stream.

Bar: 1e-12 relative on gauge-invariant results against the oracle (FP64)."""
import os

import numpy as np
import pytest
import torch

import pytreenet_amd as ptn
from oracle import contract as OC
from oracle import effective as OE
from oracle.io import load_ttn, save_ttn
from oracle.tdvp import OracleState, OracleTDVP
from pytreenet_amd import effective_hamiltonians as EH
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import PAULI_Z
from tests.helpers import host_tensors, reversed_children_operator, topology_from_ttn

pytestmark = pytest.mark.gpu
REL = 1e-12


def _say(msg):
    print(msg, flush=True)


def _setup(branching, depth, chi, seed):
    ids, parent, children = complete_tree(branching, depth)
    tensors = random_state_tensors(ids, parent, children, chi, seed)
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = ptn.random_xyz_model(st, seed)
    topo = topology_from_ttn(st)
    op_topo, Wp = reversed_children_operator(topo, host_tensors(H))
    Hp = ptn.TTNO.from_tensors(st, Wp, children=op_topo.children)
    return st, H, Hp, topo, op_topo, tensors, Wp


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_reordered_operator_environments_heff_energy(gpu):
    _say("reordered TTNO: environments, H_eff, energy; ternary depth 2 chi 6")
    st, H, Hp, topo, op_topo, tensors, Wp = _setup(3, 2, 6, 31)
    assert Hp.nodes["site0"].children == list(reversed(st.nodes["site0"].children))
    e = ptn.expectation_value(st, Hp)
    e_ref = OC.expectation_value(topo, tensors, Wp, op_topo=op_topo)
    assert _rel(e, e_ref) < REL
    assert _rel(ptn.expectation_value(st, H), e_ref) < REL
    cache = ptn.PartialTreeCachDict()
    oenv = {}
    c = "site1"
    dist = topo.distance_to_node(c)
    for d in range(max(dist.values()), 0, -1):
        for n in [m for m in dist if dist[m] == d]:
            to = min(topo.neighbours(n), key=lambda x: dist[x])
            cache.add_entry(n, to, ptn.contract_any(n, to, st, Hp, cache))
            oenv[(n, to)] = OC.env_update(topo, n, to, tensors[n], Wp[n], oenv, op_topo=op_topo)
            ref = oenv[(n, to)]
            np.testing.assert_allclose(cache.get_entry(n, to).cpu().numpy(), ref, rtol=0,
                                       atol=REL * np.abs(ref).max())
    node, x = st[c]
    y = ptn.contract_ket_ham_with_envs(node, x, Hp.nodes[c], Hp.tensors[c], cache)
    ref = OC.heff_apply(topo, c, tensors[c], Wp[c], oenv, op_topo=op_topo)
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=0, atol=REL * np.abs(ref).max())


def test_dense_effective_hamiltonians_match_oracle(gpu):
    """Site, two-site and bond matrices from the blocks a TDVP cache holds, against the
    oracle's restatement of the reference functions (reordered TTNO)."""
    _say("dense effective site / bond / two-site Hamiltonians, reordered TTNO")
    st, H, Hp, topo, op_topo, tensors, Wp = _setup(2, 2, 4, 32)
    alg = ptn.SecondOrderOneSiteTDVP(st, Hp, 0.05, 0.05, [],
                                     config=ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV))
    c = alg.update_path[0]
    p = topo.parent[c]
    cache, s = alg.partial_tree_cache, alg.state
    A = host_tensors(s)
    # every block pointing towards c (the TDVP cache); it also holds the blocks into
    # p from p's other neighbours
    oenv = {}
    dist = topo.distance_to_node(c)
    for d in range(max(dist.values()), 0, -1):
        for n in [m for m in dist if dist[m] == d]:
            to = min(topo.neighbours(n), key=lambda x: dist[x])
            oenv[(n, to)] = OC.env_update(topo, n, to, A[n], Wp[n], oenv, op_topo=op_topo)
    got = EH.get_effective_single_site_hamiltonian(c, s, Hp, cache).cpu().numpy()
    ref = OE.effective_single_site_hamiltonian(c, topo.neighbours(c), op_topo.neighbours(c), 1,
                                               Wp[c], oenv)
    np.testing.assert_allclose(got, ref, rtol=0, atol=REL * np.abs(ref).max())
    # the dense matrix is the action of contract_ket_ham_with_envs
    node, x = s[c]
    y = ptn.contract_ket_ham_with_envs(node, x, Hp.nodes[c], Hp.tensors[c], cache)
    np.testing.assert_allclose(got @ A[c].reshape(-1), y.cpu().numpy().reshape(-1),
                               rtol=0, atol=1e-12 * np.abs(got).max() * np.abs(A[c]).sum())
    merged = EH.two_site_node(s, c, p)
    got2 = EH.get_effective_two_site_hamiltonian(c, p, s, Hp, cache).cpu().numpy()
    ref2 = OE.effective_two_site_hamiltonian(merged.neighbouring_nodes(), c,
                                             op_topo.neighbours(c), Wp[c], p,
                                             op_topo.neighbours(p), Wp[p], oenv)
    assert got2.shape == ref2.shape
    np.testing.assert_allclose(got2, ref2, rtol=0, atol=REL * np.abs(ref2).max())
    # bond Hamiltonian across (p, c) from the two directed blocks
    oenv[(c, p)] = OC.env_update(topo, c, p, A[c], Wp[c], oenv, op_topo=op_topo)
    cache.add_entry(c, p, ptn.contract_any(c, p, s, Hp, cache))
    link = ptn.Node(identifier="link")
    link.parent, link.children = p, [c]
    gotb = EH.get_effective_bond_hamiltonian_nodes(link, cache).cpu().numpy()
    refb = OE.half_matrix(OE.effective_bond_hamiltonian_tensor(p, c, oenv))
    np.testing.assert_allclose(gotb, refb, rtol=0, atol=REL * np.abs(refb).max())


@pytest.mark.parametrize("kind", ["second1", "second2"])
def test_tdvp_with_reordered_operator_matches_oracle(gpu, kind):
    _say(f"tdvp {kind} with reordered TTNO, ternary depth 2 chi 6")
    st, H, Hp, topo, op_topo, tensors, Wp = _setup(3, 2, 6, 33)
    opts = {"krylov_tol": 1e-10, "max_krylov_dim": 10}
    svd = dict(max_bond_dim=8, rel_tol=1e-10, total_tol=1e-10)
    if kind == "second2":
        cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, **svd)
        alg = ptn.SecondOrderTwoSiteTDVP(st, Hp, 0.05, 0.1, [], config=cfg, solver_options=opts)
        ref_alg = ptn.SecondOrderTwoSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                             solver_options=opts)
    else:
        cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
        alg = ptn.SecondOrderOneSiteTDVP(st, Hp, 0.05, 0.1, [], config=cfg, solver_options=opts)
        ref_alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                             solver_options=opts)
    otd = OracleTDVP(OracleState(topo, tensors), Wp, 0.05, kind=kind, krylov_tol=1e-10,
                     max_krylov_dim=10, svd_params=svd if kind == "second2" else None,
                     op_topo=op_topo)
    for step in range(2):
        alg.run_one_time_step()
        ref_alg.run_one_time_step()
        otd.run_one_time_step()
        e, e_same, e_ref = (ptn.expectation_value(alg.state, Hp),
                            ptn.expectation_value(ref_alg.state, H), otd.energy())
        _say(f"  step {step}: E={e.real:.15f} oracle={e_ref.real:.15f}")
        assert _rel(e, e_ref) < REL and _rel(e_same, e_ref) < REL
        n2 = abs(alg.state.scalar_product())
        for nid in ("site0", "site5", "site12"):
            z = alg.state.single_site_operator_expectation_value(nid, PAULI_Z)
            assert abs(z - otd.expect_site(nid, PAULI_Z)) < REL * n2
    if kind == "second2":
        assert alg.kept_dims == ref_alg.kept_dims


def test_interchange_files_device_to_oracle_and_reordered_ttno(gpu, tmp_path):
    _say("save/load: device state -> oracle, oracle-written TTNO with its own order -> device")
    st, H, Hp, topo, op_topo, tensors, Wp = _setup(2, 3, 8, 34)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.05, [],
                                     config=ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV),
                                     solver_options={"max_krylov_dim": 10})
    alg.run_one_time_step()
    path = str(tmp_path / "state")
    alg.state.save(path)
    otopo, A, center = load_ttn(path)
    assert center == alg.state.orthogonality_center_id
    assert all(otopo.children[n] == topo.children[n] for n in topo.node_ids)
    e_dev = ptn.expectation_value(alg.state, H)
    assert _rel(e_dev, OC.expectation_value(otopo, A, host_tensors(H))) < REL
    ov_dev = ptn.contract_two_ttns(alg.state, st, conj_second=True)
    assert _rel(ov_dev, OC.overlap(topo, tensors, A)) < REL
    # a TTNO written in the reference format with its own children order
    opath = str(tmp_path / "ttno")
    save_ttn(opath, op_topo, Wp)
    Hl = ptn.TTNO.load(opath, device="cuda")
    assert Hl.nodes["site0"].children == op_topo.children["site0"]
    assert _rel(ptn.expectation_value(alg.state, Hl), e_dev) < REL
    # the device state read back from the files
    back = ptn.TTNS.load(path, device="cuda")
    assert _rel(ptn.expectation_value(back, Hl), e_dev) < REL


def test_default_solver_options_on_large_local_space(gpu):
    """No max_krylov_dim: the recurrence runs until the Saad estimate stops it, and the
    buffers grow with the steps taken, not with the local dimension (131072 here)."""
    _say("default solver options, binary depth 4 chi 64 (local spaces up to 131072)")
    ids, parent, children = complete_tree(2, 4)
    tensors = random_state_tensors(ids, parent, children, 64, 35)
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = ptn.transverse_field_ising(st, 1.0, 0.5)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.01, 0.01, [],
                                     config=ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV))
    # normalised: the Saad stop compares an absolute estimate (it carries ||psi||)
    # with krylov_tol, so an unnormalised random state would run to breakdown
    c = alg.state.orthogonality_center_id
    alg.state.tensors[c] = alg.state.tensors[c] / alg.state.norm()
    assert max(t.numel() for t in alg.state.tensors.values()) >= 100000
    e0 = ptn.expectation_value(alg.state, H) / alg.state.scalar_product()
    alg.run_one_time_step()
    e1 = ptn.expectation_value(alg.state, H) / alg.state.scalar_product()
    assert abs(e1 - e0) < 1e-9 * abs(e0)
    assert abs(alg.state.norm() - 1.0) < 1e-10


def test_public_entry_points_follow_the_current_stream(gpu):
    _say("expectation values and contractions inside torch.cuda.stream")
    st, H, Hp, topo, op_topo, tensors, Wp = _setup(2, 3, 8, 36)
    e_ref = OC.expectation_value(topo, tensors, host_tensors(H))
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        e = ptn.expectation_value(st, H)
        z = st.single_site_operator_expectation_value("site3", PAULI_Z)
        n2 = st.scalar_product(use_orthogonal_center=False)
    assert _rel(e, e_ref) < REL
    assert abs(z - OC.single_site_expectation(topo, tensors, "site3", PAULI_Z)) < REL * abs(n2)
    with torch.cuda.stream(side):
        e2 = ptn.expectation_value(st, Hp)
    torch.cuda.synchronize()
    assert _rel(e2, e_ref) < REL


def test_multi_single_site_expectation_values_match_oracle(gpu):
    _say("<Z_i>, <X_i> for every site in two passes, binary depth 3 chi 8")
    from pytreenet_amd.contractions import multi_single_site_expectation_value
    from pytreenet_amd.ttno import PAULI_X
    st, H, Hp, topo, op_topo, tensors, Wp = _setup(2, 3, 8, 37)
    ops = {nid: [PAULI_Z, ptn.TensorProduct({nid: PAULI_X})] for nid in st.nodes}
    got = multi_single_site_expectation_value(st, ops)
    n2 = abs(OC.overlap(topo, tensors, tensors))
    for nid in st.nodes:
        assert abs(got[nid][0] - OC.single_site_expectation(topo, tensors, nid, PAULI_Z)) \
            < REL * n2
        assert abs(got[nid][1] - OC.single_site_expectation(topo, tensors, nid, PAULI_X)) \
            < REL * n2
    assert st.orthogonality_center_id is None
