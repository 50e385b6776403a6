"""Environment blocks that are identity matrices (canonical form), skipped in the
effective Hamiltonians and environment updates of the time evolution.

In the canonical form a TDVP sweep keeps, the block of an environment in the
bond state where the operator on its side is the identity
(``TTNO.identity_states``) equals the identity matrix, and
``engine._split_identity`` contracts the network as a sum of smaller networks
without it.  Checked here with every network split (size thresholds 0) and every
identity block verified on the device (``engine.CHECK_IDENTITY``):

* environments and H_eff actions of a canonical state equal the unsplit ones;
* one-site TDVP (first and second order, three models) equals the oracle within
  1e-12 relative, as in the unsplit parity tests;
* the relaxed form and the two-site form (whose environment updates split)
  equal the oracle at their usual bars;
* a TDVP step on a local space of 2^18 elements (default thresholds) equals the
  same step with the split switched off.
"""
import numpy as np
import pytest
import torch

import pytreenet_amd as ptn
from oracle.tdvp import OracleState, OracleTDVP
from pytreenet_amd import contractions as PC
from pytreenet_amd import engine as E
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import PAULI_Z
from tests.helpers import host_tensors, topology_from_ttn

pytestmark = pytest.mark.gpu
REL = 1e-12


def _say(msg):
    print(msg, flush=True)


@pytest.fixture
def split_everywhere():
    saved = (PC.BLOCK_SPARSE_MIN_SIZE, PC.IDENTITY_MIN_SIZE, E.CHECK_IDENTITY,
             PC.IDENTITY_BLOCKS)
    PC.BLOCK_SPARSE_MIN_SIZE = 0
    PC.IDENTITY_MIN_SIZE = 0
    E.CHECK_IDENTITY = True
    PC.IDENTITY_BLOCKS = True
    E.clear_caches()
    yield
    (PC.BLOCK_SPARSE_MIN_SIZE, PC.IDENTITY_MIN_SIZE, E.CHECK_IDENTITY,
     PC.IDENTITY_BLOCKS) = saved
    E.clear_caches()


def _setup(branching, depth, chi, seed, model):
    ids, parent, children = complete_tree(branching, depth)
    tensors = random_state_tensors(ids, parent, children, chi, seed)
    st = build_state(ids, parent, children, tensors, device="cuda")
    if model == "tfim":
        H = ptn.transverse_field_ising(st, 1.0, 0.5)
    elif model == "xxz":
        H = ptn.xxz_model(st, 1.0, 0.8, field=0.3)
    elif model == "heis":
        H = ptn.random_field_heisenberg(st, 2.0, seed)
    else:
        H = ptn.random_xyz_model(st, seed)
    return st, H, tensors


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_identity_states_of_edge_term_ttno(gpu):
    st, H, _ = _setup(3, 2, 4, 1, "xyz")
    for nid, node in H.nodes.items():
        for nb in node.neighbouring_nodes():
            # state 0 (nothing placed) towards the parent, 1 (term complete) downwards
            want = (0,) if node.parent == nb else (1,)
            assert H.identity_states(nid, nb) == want, (nid, nb)


@pytest.mark.parametrize("model", ["xyz", "tfim"])
def test_split_networks_equal_unsplit(gpu, split_everywhere, model):
    _say(f"identity split vs unsplit networks, ternary depth 2 chi 6, {model}")
    st, H, _ = _setup(3, 2, 6, 5, model)
    for center in ["site0", "site1", "site5"]:
        st.canonical_form(center, mode=ptn.SplitMode.KEEP)
        split = PC.SandwichCache.init_cache_but_one(st, H, center, canonical=True)
        plain = PC.SandwichCache.init_cache_but_one(st, H, center, canonical=False)
        # the small environment networks are recorded on their first build and
        # replayed on this one (the split parts read slices of the new inputs)
        again = PC.SandwichCache.init_cache_but_one(st, H, center, canonical=True)
        for key, blk in plain.items():
            ref = blk.cpu().numpy()
            for c in (split, again):
                got = c[key].cpu().numpy()
                np.testing.assert_allclose(got, ref, rtol=0, atol=REL * np.abs(ref).max())
        node, x = st[center]
        op = H.block_tensor(center, x.dtype)
        envs = [split.get_entry(nb, center) for nb in node.neighbouring_nodes()]
        tensors, labels, out = PC.heff_network(node, H.nodes[center], op, envs)
        tensors[0] = x
        ident = PC.identity_slots(center, node.neighbouring_nodes(), H, 2)
        net = E.CompiledNetwork(tensors, labels, out, identity=ident)
        ref = E.contract_network(tensors, labels, out).cpu().numpy()
        for rep in range(3):   # recorded once, replayed twice (new input each time)
            xin = x * (1.0 + 0.5j * rep)
            got = net(xin).cpu().numpy()
            np.testing.assert_allclose(got, ref * (1.0 + 0.5j * rep), rtol=0,
                                       atol=REL * np.abs(ref).max() * 2)


@pytest.mark.parametrize("kind,model,tol", [("second1", "xyz", 1e-10), ("second1", "tfim", 0.0),
                                            ("first1", "xxz", 1e-10)])
def test_tdvp_one_site_with_split_matches_oracle(gpu, split_everywhere, kind, model, tol):
    _say(f"tdvp {kind} {model} with identity split, krylov_tol={tol}")
    st, H, tensors = _setup(2, 3, 8, 11, model)
    topo, W = topology_from_ttn(st), host_tensors(H)
    cls = ptn.FirstOrderOneSiteTDVP if kind == "first1" else ptn.SecondOrderOneSiteTDVP
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
    opts = {"krylov_tol": tol, "max_krylov_dim": 10}
    alg = cls(st, H, 0.05, 0.15, [], config=cfg, solver_options=opts)
    otd = OracleTDVP(OracleState(topo, tensors), W, 0.05, kind=kind, krylov_tol=tol,
                     max_krylov_dim=10)
    for step in range(3):
        alg.run_one_time_step()
        otd.run_one_time_step()
        e = ptn.expectation_value(alg.state, alg.hamiltonian)
        e_ref = otd.energy()
        assert _rel(e, e_ref) < REL, (step, e, e_ref)
        for nid in ["site0", "site3", "site14"]:
            z = alg.state.single_site_operator_expectation_value(nid, PAULI_Z)
            assert abs(z - otd.expect_site(nid, PAULI_Z)) < REL * abs(e_ref), nid


def test_two_site_tdvp_with_split_environments_matches_oracle(gpu, split_everywhere):
    _say("tdvp second2 heisenberg truncation with identity split environments")
    st, H, tensors = _setup(2, 3, 4, 17, "heis")
    topo, W = topology_from_ttn(st), host_tensors(H)
    svd = dict(max_bond_dim=6, rel_tol=1e-10, total_tol=1e-10)
    cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, **svd)
    alg = ptn.SecondOrderTwoSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                     solver_options={"krylov_tol": 1e-10, "max_krylov_dim": 10})
    otd = OracleTDVP(OracleState(topo, tensors), W, 0.05, kind="second2", krylov_tol=1e-10,
                     max_krylov_dim=10, svd_params=svd)
    for _ in range(2):
        alg.run_one_time_step()
        otd.run_one_time_step()
        assert _rel(ptn.expectation_value(alg.state, alg.hamiltonian), otd.energy()) < REL
    dims = {k: tuple(v.shape) for k, v in alg.state.tensors.items()}
    assert dims == {k: tuple(v.shape) for k, v in otd.s.A.items()}


def test_relaxed_tdvp_with_split_matches_oracle(gpu, split_everywhere):
    _say("relaxed tdvp with identity split (relax_tol 0: every apply FP64)")
    st, H, tensors = _setup(2, 3, 8, 21, "xyz")
    topo, W = topology_from_ttn(st), host_tensors(H)
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision="relaxed",
                         relax_tol=0.0, relax_min_size=0)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                     solver_options={"krylov_tol": 1e-10, "max_krylov_dim": 10})
    otd = OracleTDVP(OracleState(topo, tensors), W, 0.05, kind="second1", krylov_tol=1e-10,
                     max_krylov_dim=10)
    for _ in range(2):
        alg.run_one_time_step()
        otd.run_one_time_step()
        assert _rel(ptn.expectation_value(alg.state, alg.hamiltonian), otd.energy()) < REL
    # a loose tolerance runs the split FP32 network too; it stays near the oracle
    cfg2 = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision="relaxed",
                          relax_tol=1e-6, relax_min_size=0)
    st2 = build_state(*complete_tree(2, 3), tensors, device="cuda")
    alg2 = ptn.SecondOrderOneSiteTDVP(st2, H, 0.05, 0.1, [], config=cfg2,
                                      solver_options={"krylov_tol": 1e-10, "max_krylov_dim": 10})
    for _ in range(2):
        alg2.run_one_time_step()
    assert _rel(ptn.expectation_value(alg2.state, H), otd.energy()) < 1e-6


def test_large_site_split_equals_unsplit_step(gpu):
    """Default thresholds: the ternary tree of 40 nodes at chi 64 (root 2^19
    elements, level-1 nodes 2^18); one second-order step with and without the
    split, from the same state."""
    _say("identity split at default thresholds, ternary depth 3 chi 64")
    ids, parent, children = complete_tree(3, 3)
    tensors = random_state_tensors(ids, parent, children, 64, seed=7)
    res = []
    for on in (True, False):
        PC.IDENTITY_BLOCKS = on
        E.clear_caches()
        try:
            st = build_state(ids, parent, children, tensors, device="cuda")
            H = ptn.random_xyz_model(st, seed=3)
            cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
            alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.05, [], config=cfg,
                                             solver_options={"krylov_tol": 1e-10,
                                                             "max_krylov_dim": 10})
            c = alg.state.orthogonality_center_id
            alg.state.tensors[c] = alg.state.tensors[c] / alg.state.norm()
            assert max(t.numel() for t in alg.state.tensors.values()) >= 1 << 16
            alg.run_one_time_step()
            z = [alg.state.single_site_operator_expectation_value(n, PAULI_Z).real
                 for n in ["site0", "site1", "site13", "site39"]]
            res.append((ptn.expectation_value(alg.state, H), z, alg.apply_log))
        finally:
            PC.IDENTITY_BLOCKS = True
    (e1, z1, log1), (e0, z0, log0) = res
    assert _rel(e1, e0) < REL
    assert np.max(np.abs(np.array(z1) - np.array(z0))) < REL
    assert [a for _, _, a in log1] == [a for _, _, a in log0]
