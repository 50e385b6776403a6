"""Parity of the device path with the CPU oracle on the same seeded inputs.

Bar (FP64): gauge-invariant results (energies, overlaps, <Z_i>) within 1e-12
relative; kept bond dimensions identical.  Mixed precision: stated bound.
"""
import numpy as np
import pytest
import torch

import pytreenet_amd as ptn
from oracle import contract as OC
from oracle.tdvp import OracleState, OracleTDVP
from pytreenet_amd import engine as E
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import PAULI_Z, build_edge_ttno_tensors
from tests.helpers import host_tensors, topology_from_ttn

pytestmark = pytest.mark.gpu
REL = 1e-12


def _say(msg):
    print(msg, flush=True)


def _setup(branching, depth, chi, seed, model="xyz"):
    ids, parent, children = complete_tree(branching, depth)
    tensors = random_state_tensors(ids, parent, children, chi, seed)
    st = build_state(ids, parent, children, tensors, device="cuda")
    if model == "tfim":
        H = ptn.transverse_field_ising(st, 1.0, 0.5)
    elif model == "xxz":
        H = ptn.xxz_model(st, 1.0, 0.8)
    elif model == "heis":
        H = ptn.random_field_heisenberg(st, 2.0, seed)
    else:
        H = ptn.random_xyz_model(st, seed)
    topo = topology_from_ttn(st)
    W = host_tensors(H)
    return st, H, topo, tensors, W


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_env_heff_bond_match_oracle(gpu):
    _say("env/heff/bond parity, ternary depth 2 chi 6")
    st, H, topo, tensors, W = _setup(3, 2, 6, 3)
    cache = ptn.PartialTreeCachDict()
    oenv = {}
    c = "site1"
    dist = topo.distance_to_node(c)
    for d in range(max(dist.values()), 0, -1):
        for n in [m for m in dist if dist[m] == d]:
            to = min(topo.neighbours(n), key=lambda x: dist[x])
            cache.add_entry(n, to, ptn.contract_any(n, to, st, H, cache))
            oenv[(n, to)] = OC.env_update(topo, n, to, tensors[n], W[n], oenv)
            got = cache.get_entry(n, to).cpu().numpy()
            np.testing.assert_allclose(got, oenv[(n, to)], rtol=0,
                                       atol=REL * np.abs(oenv[(n, to)]).max())
    node, x = st[c]
    y = ptn.contract_ket_ham_with_envs(node, x, H.nodes[c], H.tensors[c], cache)
    ref = OC.heff_apply(topo, c, tensors[c], W[c], oenv)
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=0, atol=REL * np.abs(ref).max())
    # bond action across (site1, site0): E[(site1, site0)] and E[(site0, site1)]
    rng = np.random.default_rng(0)
    dim = tensors["site1"].shape[0]
    r = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
    e_a = OC.env_update(topo, "site1", "site0", tensors["site1"], W["site1"], oenv)
    got_a = ptn.contract_any("site1", "site0", st, H, cache).cpu().numpy()
    np.testing.assert_allclose(got_a, e_a, rtol=0, atol=REL * np.abs(e_a).max())
    e_b = oenv[("site0", "site1")]
    yb = ptn.contractions.bond_apply(torch.as_tensor(r).cuda(), torch.as_tensor(e_a).cuda(),
                                     torch.as_tensor(e_b).cuda()).cpu().numpy()
    refb = OC.bond_apply(r, e_a, e_b)
    np.testing.assert_allclose(yb, refb, rtol=0, atol=REL * np.abs(refb).max())


def test_observables_match_oracle(gpu):
    _say("observables parity, binary depth 3 chi 8")
    st, H, topo, tensors, W = _setup(2, 3, 8, 5)
    e = ptn.expectation_value(st, H)
    e_ref = OC.expectation_value(topo, tensors, W)
    assert _rel(e, e_ref) < REL
    n2 = st.scalar_product(use_orthogonal_center=False)
    assert _rel(n2, OC.overlap(topo, tensors, tensors)) < REL
    for nid in ["site0", "site4", "site14"]:
        z = st.single_site_operator_expectation_value(nid, PAULI_Z)
        zr = OC.single_site_expectation(topo, tensors, nid, PAULI_Z)
        assert abs(z - zr) < REL * abs(n2)


def _run_pair(kind, branching, depth, chi, seed, model, dt, steps, krylov_tol=1e-10,
              svd=None, precision="fp64"):
    st, H, topo, tensors, W = _setup(branching, depth, chi, seed, model)
    ops = {nid: ptn.TensorProduct({nid: PAULI_Z}) for nid in list(st.nodes)[:4]}
    opts = {"krylov_tol": krylov_tol, "max_krylov_dim": 10}
    if kind in ("second2", "first2"):
        cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision=precision,
                                    **(svd or {}))
        cls = ptn.SecondOrderTwoSiteTDVP if kind == "second2" else ptn.FirstOrderTwoSiteTDVP
        alg = cls(st, H, dt, dt * steps, ops, config=cfg, solver_options=opts)
    else:
        cls = ptn.FirstOrderOneSiteTDVP if kind == "first1" else ptn.SecondOrderOneSiteTDVP
        cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision=precision)
        alg = cls(st, H, dt, dt * steps, ops, config=cfg, solver_options=opts)
    ost = OracleState(topo, tensors)
    otd = OracleTDVP(ost, W, dt, kind=kind, krylov_tol=krylov_tol, max_krylov_dim=10,
                     svd_params=svd)
    return alg, otd, ops


@pytest.mark.parametrize("kind,model,tol", [("second1", "xyz", 1e-10), ("second1", "tfim", 0.0),
                                            ("first1", "xxz", 1e-10)])
def test_tdvp_one_site_matches_oracle(gpu, kind, model, tol):
    _say(f"tdvp {kind} {model} krylov_tol={tol}")
    alg, otd, ops = _run_pair(kind, 2, 3, 8, 11, model, 0.05, 3, krylov_tol=tol)
    for step in range(3):
        alg.run_one_time_step()
        otd.run_one_time_step()
        e = ptn.expectation_value(alg.state, alg.hamiltonian)
        e_ref = otd.energy()
        assert _rel(e, e_ref) < REL, (step, e, e_ref)
        for nid in ops:
            z = alg.state.single_site_operator_expectation_value(nid, PAULI_Z)
            assert abs(z - otd.expect_site(nid, PAULI_Z)) < REL * abs(e_ref), nid


@pytest.mark.parametrize("kind", ["second2", "first2"])
def test_tdvp_two_site_truncation_matches_oracle(gpu, kind):
    _say(f"tdvp {kind} heisenberg truncation")
    svd = dict(max_bond_dim=6, rel_tol=1e-10, total_tol=1e-10)
    alg, otd, ops = _run_pair(kind, 2, 3, 4, 17, "heis", 0.05, 2, svd=svd)
    otd_dims = []
    for step in range(2):
        alg.run_one_time_step()
        n0 = len(otd.krylov_dims)
        otd.run_one_time_step()
        e = ptn.expectation_value(alg.state, alg.hamiltonian)
        assert _rel(e, otd.energy()) < REL
    dims = {k: v.shape for k, v in alg.state.tensors.items()}
    odims = {k: v.shape for k, v in otd.s.A.items()}
    assert dims == {k: tuple(v) for k, v in odims.items()}


def test_run_loop_config1_tfim(gpu):
    """Binary tree of 15 nodes, chi <= 16, TFIM J=1 g=0.5, dt=0.01, 20 steps, Krylov dim 10."""
    _say("config C1 full run vs oracle")
    st, H, topo, tensors, W = _setup(2, 3, 16, 2024, "tfim")
    ops = {nid: ptn.TensorProduct({nid: PAULI_Z}) for nid in st.nodes}
    ops["energy"] = H
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.01, 0.2, ops, config=cfg,
                                     solver_options={"max_krylov_dim": 10})
    alg.run()
    otd = OracleTDVP(OracleState(topo, tensors), W, 0.01, kind="second1", max_krylov_dim=10)
    e_series = alg.results.operator_result("energy")
    assert len(e_series) == 21
    for i in range(21):
        if i:
            otd.run_one_time_step()
        e_ref = otd.energy()
        assert _rel(e_series[i], e_ref) < REL, i
        for nid in ["site0", "site7", "site14"]:
            assert abs(alg.results.operator_result(nid)[i] - otd.expect_site(nid, PAULI_Z)) \
                < REL * abs(e_ref)


def test_mixed_precision_error_bound(gpu):
    _say("mixed precision fp32 3M vs fp64")
    a64, _, _ = _run_pair("second1", 2, 3, 16, 5, "xyz", 0.05, 2)
    a32, _, _ = _run_pair("second1", 2, 3, 16, 5, "xyz", 0.05, 2, precision="fp32")
    for _ in range(2):
        a64.run_one_time_step()
        a32.run_one_time_step()
    e64 = ptn.expectation_value(a64.state, a64.hamiltonian)
    e32 = ptn.expectation_value(a32.state, a32.hamiltonian)
    assert abs(e64 - e32) < 1e-4 * abs(e64)


def test_block_sparse_ttno_matches_dense(gpu):
    """H_eff, environments and energy with the block-sparse TTNO equal the dense ones."""
    _say("block-sparse vs dense TTNO")
    from pytreenet_amd import contractions as PC
    st, H, topo, tensors, W = _setup(3, 2, 6, 9)
    c = "site1"
    envs = {}
    dist = topo.distance_to_node(c)
    cache_b, cache_d = ptn.PartialTreeCachDict(), ptn.PartialTreeCachDict()
    min_size = PC.BLOCK_SPARSE_MIN_SIZE
    PC.BLOCK_SPARSE_MIN_SIZE = 0     # block sparse at every size of this small tree
    try:
        for d in range(max(dist.values()), 0, -1):
            for n in [m for m in dist if dist[m] == d]:
                to = min(topo.neighbours(n), key=lambda x: dist[x])
                PC.USE_BLOCK_SPARSE = True
                cache_b.add_entry(n, to, ptn.contract_any(n, to, st, H, cache_b))
                PC.USE_BLOCK_SPARSE = False
                cache_d.add_entry(n, to, ptn.contract_any(n, to, st, H, cache_d))
                eb, ed = cache_b.get_entry(n, to).cpu().numpy(), cache_d.get_entry(n, to).cpu().numpy()
                np.testing.assert_allclose(eb, ed, rtol=0, atol=REL * np.abs(ed).max())
        node, x = st[c]
        yb = ptn.contract_ket_ham_with_envs(node, x, H.nodes[c], H.block_tensor(c, x.dtype),
                                            cache_b).cpu().numpy()
        yd = ptn.contract_ket_ham_with_envs(node, x, H.nodes[c], H.tensors[c],
                                            cache_d).cpu().numpy()
        np.testing.assert_allclose(yb, yd, rtol=0, atol=REL * np.abs(yd).max())
        PC.USE_BLOCK_SPARSE = True
        e_b = ptn.expectation_value(st, H)
        PC.USE_BLOCK_SPARSE = False
        e_d = ptn.expectation_value(st, H)
        assert _rel(e_b, e_d) < REL
    finally:
        PC.USE_BLOCK_SPARSE = True
        PC.BLOCK_SPARSE_MIN_SIZE = min_size


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_block_apply_path_matches_dense(gpu, precision):
    """Root of a ternary tree at chi=16: the TTNO step runs through ptn_block_apply."""
    _say(f"block apply vs dense {precision}")
    from pytreenet_amd import contractions as PC
    st, H, topo, tensors, W = _setup(3, 2, 16, 21)
    c = "site0"
    cache = ptn.PartialTreeCachDict()
    PC.USE_BLOCK_SPARSE = False
    for n, to in ptn.contractions.caching_path(st, c):
        cache.add_entry(n, to, ptn.contract_any(n, to, st, H, cache))
    PC.USE_BLOCK_SPARSE = True
    node, x = st[c]
    dt = E.storage_dtype(precision)
    envs = {k: v.to(dt) for k, v in cache.items()}
    c2 = ptn.PartialTreeCachDict(envs)
    yb = ptn.contract_ket_ham_with_envs(node, x, H.nodes[c], H.block_tensor(c, dt), c2,
                                        precision=precision).cpu().numpy()
    yd = ptn.contract_ket_ham_with_envs(node, x, H.nodes[c], H.tensor_in(c, dt), c2,
                                        precision=precision).cpu().numpy()
    tol = REL if precision == "fp64" else 1e-5
    np.testing.assert_allclose(yb, yd, rtol=0, atol=tol * np.abs(yd).max())
    for to in range(3):
        eb = PC.env_block(node, x, H.nodes[c], H.block_tensor(c, dt),
                          [None if i == to else envs[(nb, c)] for i, nb in
                           enumerate(node.neighbouring_nodes())], to,
                          precision=precision).cpu().numpy()
        ed = PC.env_block(node, x, H.nodes[c], H.tensor_in(c, dt),
                          [None if i == to else envs[(nb, c)] for i, nb in
                           enumerate(node.neighbouring_nodes())], to,
                          precision=precision).cpu().numpy()
        np.testing.assert_allclose(eb, ed, rtol=0, atol=tol * np.abs(ed).max())


def test_tdvp_block_apply_shared_structure_matches_oracle(gpu):
    """Binary 31-node tree at chi=32, random XYZ: both level-1 nodes (bonds 32,32,32)
    take the block-apply TTNO step with the same block structure but different
    couplings (regression: operator values must not be shared between them)."""
    _say("tdvp second1 binary depth 4 chi32 xyz")
    alg, otd, ops = _run_pair("second1", 2, 4, 32, 31, "xyz", 0.05, 1)
    for step in range(1):
        alg.run_one_time_step()
        otd.run_one_time_step()
        e = ptn.expectation_value(alg.state, alg.hamiltonian)
        assert _rel(e, otd.energy()) < REL, step
        for nid in ops:
            z = alg.state.single_site_operator_expectation_value(nid, PAULI_Z)
            assert abs(z - otd.expect_site(nid, PAULI_Z)) < REL * abs(otd.energy()), nid


@pytest.mark.parametrize("padded", [False, True])
def test_deferred_qr_checks_match_synchronous_checks(gpu, padded):
    """A time step runs with the CholeskyQR pivot checks deferred to its end.  With
    zero-padded (rank-deficient) bonds every CholeskyQR split of the root collapses:
    the step is redone with synchronous checks and the Householder fallback, and
    equals the run that checks synchronously from the start.  Full-rank states are
    not redone and give the same result either way."""
    _say(f"deferred QR checks padded={padded}")
    from pytreenet_amd.time_evolution import TDVPAlgorithm
    ids, parent, children = complete_tree(3, 2)
    tensors = random_state_tensors(ids, parent, children, 16, 41)
    if padded:
        for nid, t in tensors.items():
            t = t.copy()
            for ax in range(t.ndim - 1):
                idx = [slice(None)] * t.ndim
                idx[ax] = slice(3, None)
                t[tuple(idx)] = 0
            tensors[nid] = t
    from pytreenet_amd.splitting import QR_STATS
    results = []
    redone = []
    collapsed = 0  # steps of the synchronous run with a CholeskyQR3 split that fell back
    for defer in (True, False):
        st = build_state(ids, parent, children, tensors, device="cuda")
        H = ptn.random_xyz_model(st, 41)
        cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
        alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                         solver_options={"max_krylov_dim": 10})
        alg.DEFER_QR_CHECKS = defer
        n0 = TDVPAlgorithm.steps_redone
        for _ in range(2):
            h0 = QR_STATS["householder"]
            alg.run_one_time_step()
            # the only tensor above the one-launch QR's size is the root, which is tall
            collapsed += (not defer) and QR_STATS["householder"] > h0
        redone.append(TDVPAlgorithm.steps_redone - n0)
        results.append({k: v.cpu().numpy() for k, v in alg.state.tensors.items()})
    assert redone[1] == 0
    assert redone[0] == collapsed
    assert (collapsed >= 1) == padded
    for k in ids:
        np.testing.assert_allclose(results[0][k], results[1][k], rtol=0,
                                   atol=REL * np.abs(results[1][k]).max())


def test_config2_tree_matches_oracle_at_reduced_bond_dim(gpu):
    """The 63-node binary tree of the plain FP64 configuration (XXZ, Jxy=1, Jz=0.8,
    second-order one-site TDVP, dt=0.02) at bond dimension 8, where the oracle runs
    in seconds: energies, overlaps with the initial state and <Z_i> after each of 3
    steps, to 1e-12."""
    _say("config C2 tree, chi 8, vs oracle")
    alg, otd, ops = _run_pair("second1", 2, 5, 8, 63, "xxz", 0.02, 3)
    psi0 = {k: v.copy() for k, v in otd.s.A.items()}
    st0 = alg.state.copy()
    for step in range(3):
        alg.run_one_time_step()
        otd.run_one_time_step()
        e_ref = otd.energy()
        assert _rel(ptn.expectation_value(alg.state, alg.hamiltonian), e_ref) < REL, step
        ov = ptn.scalar_product(st0, alg.state)
        ov_ref = OC.overlap(otd.topo, otd.s.A, psi0)     # sum psi0 * conj(psi_t)
        assert abs(ov - ov_ref) < REL * abs(ov_ref), step
        for nid in ["site0", "site5", "site40", "site62"]:
            z = alg.state.single_site_operator_expectation_value(nid, PAULI_Z)
            assert abs(z - otd.expect_site(nid, PAULI_Z)) < REL * abs(e_ref), (step, nid)


def test_config2_full_size_conserves_norm_and_energy(gpu):
    """The plain FP64 configuration at full size (63-node binary tree, bond dimension
    64, XXZ, dt=0.02), checked through what one-site TDVP conserves: the norm and
    the energy stay at their initial values over the configuration's 50 steps (up to
    the Krylov tolerance), and |<psi_0|psi_t>| <= 1 decreases from 1."""
    _say("config C2 full size, 50 steps")
    st, H, topo, tensors, W = _setup(2, 5, 64, 63, "xxz")
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.02, 1.0, [], config=cfg,
                                     solver_options={"max_krylov_dim": 10})
    c = alg.state.orthogonality_center_id
    alg.state.tensors[c] = alg.state.tensors[c] / alg.state.norm()
    st0 = alg.state.copy()
    e0 = ptn.expectation_value(alg.state, H).real
    assert max(alg.state.bond_dims().values()) == 64 and alg.num_time_steps == 50
    for step in range(50):
        alg.run_one_time_step()
        assert abs(alg.state.norm() - 1.0) < 1e-12, step
        assert abs(ptn.expectation_value(alg.state, H).real - e0) < 1e-10 * max(1, abs(e0)), step
    ov = abs(ptn.scalar_product(st0, alg.state))
    assert ov <= 1 + 1e-12 and ov < 1 - 1e-8


def test_config3_full_size_conserves_norm_and_energy(gpu):
    """The chi=256 configuration at full size (ternary tree N=40, root (256, 256, 256, 2),
    random XYZ, dt=0.05): one second-order one-site step keeps the norm and the
    energy (the FP64 path conserves both up to the Krylov tolerance)."""
    _say("config C3 full size, 1 step")
    import bench
    st, H, _ = bench.make_problem(256, torch.device("cuda"))
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.05, [], config=cfg,
                                     solver_options={"max_krylov_dim": 10, "krylov_tol": 1e-10})
    c = alg.state.orthogonality_center_id
    alg.state.tensors[c] = alg.state.tensors[c] / alg.state.norm()
    assert tuple(alg.state.tensors[alg.state.root_id].shape) == (256, 256, 256, 2)
    e0 = ptn.expectation_value(alg.state, H).real
    alg.run_one_time_step()
    assert abs(alg.state.norm() - 1.0) < 1e-12
    assert abs(ptn.expectation_value(alg.state, H).real - e0) < 1e-10 * max(1, abs(e0))
    del alg, st, H
    E.clear_caches()
    torch.cuda.empty_cache()


def test_tdvp_autotuned_launches_match_oracle(gpu):
    """A second-order one-site run with every non-LARGE gather-GEMM launched on its
    autotuned configuration stays at the oracle's energies (1e-12)."""
    _say("tdvp second1 with autotuned GEMM configurations")
    alg, otd, _ = _run_pair("second1", 2, 3, 16, 21, "xxz", 0.05, 2)
    E.AUTOTUNE = True
    n0 = len(E.TUNE_LOG)
    try:
        for step in range(2):
            alg.run_one_time_step()
            otd.run_one_time_step()
            e, e_ref = ptn.expectation_value(alg.state, alg.hamiltonian), otd.energy()
            assert _rel(e, e_ref) < REL, (step, e, e_ref)
    finally:
        E.AUTOTUNE = False
    assert len(E.TUNE_LOG) > n0
