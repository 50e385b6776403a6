"""Pins of the oracle's dense effective Hamiltonians, of its operator-leg map for
TTNOs whose children order differs from the state's, and of its reference-order
flop counts (all on the CPU)."""
import json
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import contract as OC
from oracle import dense as OD
from oracle import effective as OE
from oracle import flops as OF
from oracle.tdvp import OracleState, OracleTDVP
from oracle.tree import Topology
from pytreenet_amd.random_states import bond_dims, complete_tree
from tests.helpers import oracle_problem, reversed_children_operator, small_tree, \
    topology_from_spec

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gold():
    with open(os.path.join(GOLD, "effective_hamiltonian.json")) as f:
        return json.load(f)


def _crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ---- dense effective Hamiltonians: the reference tests' expected values ---------------
@pytest.mark.parametrize("case", _gold()["leg_permutations"], ids=lambda c: c["name"])
def test_oracle_leg_permutation_kats(case):
    got = OE.find_tensor_leg_permutation(case["state_nbs"], case["ham_nbs"], case["n_open"])
    assert got == tuple(case["perm"])


@pytest.mark.parametrize("case", _gold()["contract_all_except_node"], ids=lambda c: c["name"])
def test_oracle_contract_all_except_node_matches_reference_route(case):
    """The reference test builds the expected tensor by tensordots in the
    Hamiltonian node's neighbour order followed by a fixed transpose."""
    rng = np.random.default_rng(11)
    W = _crandn(rng, *case["ham_shape"])
    envs = {(nb, "node"): _crandn(rng, d, W.shape[case["ham_nbs"].index(nb)], d)
            for nb, d in case["env_dims"].items()}
    expect = W
    for nb, leg_w, leg_e in case["route"]:
        expect = np.tensordot(expect, envs[(nb, "node")], axes=(leg_w, leg_e))
    expect = expect.transpose(case["transpose"])
    got = OE.contract_all_except_node("node", case["state_nbs"], case["ham_nbs"], 1, W, envs)
    assert got.shape == tuple(case["shape"])
    np.testing.assert_allclose(got, expect, atol=1e-12)


def _small_ttns_and_ttno_blocks(rng):
    """Shapes of the reference's small_ttns_and_ttno (root with children c1, c2;
    bonds 5 and 6; physical dims 2, 3, 4) with operator bond dims 3 and 4."""
    D1, D2 = 3, 4
    W = {"root": _crandn(rng, D1, D2, 2, 2), "c1": _crandn(rng, D1, 3, 3),
         "c2": _crandn(rng, D2, 4, 4)}
    envs = {("c1", "root"): _crandn(rng, 5, D1, 5), ("c2", "root"): _crandn(rng, 6, D2, 6),
            ("root", "c1"): _crandn(rng, 5, D1, 5), ("root", "c2"): _crandn(rng, 6, D2, 6)}
    return W, envs


def test_oracle_single_site_effective_hamiltonians_match_reference_routes():
    g = _gold()["small_ttns_and_ttno"]
    W, envs = _small_ttns_and_ttno_blocks(np.random.default_rng(12))
    nbs = {"root": ["c1", "c2"], "c1": ["root"], "c2": ["root"]}
    routes = {
        "c1": np.tensordot(envs[("root", "c1")], W["c1"], axes=(1, 0)),
        "c2": np.tensordot(W["c2"], envs[("root", "c2")], axes=(0, 1)),
        "root": np.tensordot(np.tensordot(W["root"], envs[("c1", "root")], axes=(0, 1)),
                             envs[("c2", "root")], axes=(0, 1)),
    }
    for case in g["single_site"]:
        nid = case["node"]
        expect = routes[nid].transpose(case["transpose"]).reshape(case["matrix"])
        got = OE.effective_single_site_hamiltonian(nid, nbs[nid], nbs[nid], 1, W[nid], envs)
        np.testing.assert_allclose(got, expect, atol=1e-12)


def test_oracle_two_site_effective_hamiltonian_matches_reference_route():
    g = _gold()["small_ttns_and_ttno"]["two_site_c2_root"]
    W, envs = _small_ttns_and_ttno_blocks(np.random.default_rng(13))
    expect = np.tensordot(np.tensordot(envs[("c1", "root")], W["root"], axes=(1, 0)),
                          W["c2"], axes=(2, 0)).transpose(g["transpose"]).reshape(g["matrix"])
    got = OE.effective_two_site_hamiltonian(g["merged_nbs"], "c2", ["root"], W["c2"],
                                            "root", ["c1", "c2"], W["root"], envs)
    np.testing.assert_allclose(got, expect, atol=1e-12)


def test_oracle_effective_hamiltonian_is_projected_hamiltonian():
    """On a tree in canonical form the dense H_eff equals P^H H P, and equals the
    matrix of heff_apply, also with the operator's children in reverse order."""
    topo, st, W, H = oracle_problem(2, 2, 4, seed=21)
    op_topo, Wp = reversed_children_operator(topo, W)
    c = topo.node_ids[1]
    st.move_center(c)
    dist = topo.distance_to_node(c)
    envs = {}
    for d in range(max(dist.values()), 0, -1):
        for n in [m for m in dist if dist[m] == d]:
            to = min(topo.neighbours(n), key=lambda x: dist[x])
            envs[(n, to)] = OC.env_update(topo, n, to, st.A[n], Wp[n], envs, op_topo=op_topo)
    x = st.A[c]
    basis = np.eye(x.size, dtype=complex)
    P = np.stack([OD.state_to_vector(topo, {**st.A, c: basis[i].reshape(x.shape)})
                  for i in range(x.size)], axis=1)
    heff = OE.effective_single_site_hamiltonian(c, topo.neighbours(c), op_topo.neighbours(c), 1,
                                                Wp[c], envs)
    np.testing.assert_allclose(heff, P.conj().T @ H @ P, atol=1e-11)
    y = OC.heff_apply(topo, c, x, Wp[c], envs, op_topo=op_topo).reshape(-1)
    np.testing.assert_allclose(y, heff @ x.reshape(-1), atol=1e-11)


# ---- operator whose children order differs from the state's ------------------------------
def test_oracle_operator_leg_map_expectation_value():
    topo, st, W, H = oracle_problem(2, 2, 4, seed=22)
    op_topo, Wp = reversed_children_operator(topo, W)
    assert op_topo.children[topo.root] != topo.children[topo.root]
    psi = OD.state_to_vector(topo, st.A)
    np.testing.assert_allclose(OD.ttno_to_matrix(op_topo, Wp), H, atol=1e-12)
    e = OC.expectation_value(topo, st.A, Wp, op_topo=op_topo)
    assert abs(e - np.vdot(psi, H @ psi)) < 1e-11 * abs(e)
    # a plausible slip -- the W legs taken in the state's order -- is detectable
    e_bad = OC.expectation_value(topo, st.A, Wp)
    assert abs(e_bad - e) > 1e-6 * abs(e)


@pytest.mark.parametrize("kind", ["second1", "second2"])
def test_oracle_tdvp_with_reordered_operator_exact(kind):
    """At full bond dimensions TDVP with the reordered operator equals exact evolution."""
    ids, parent, children, tensors = small_tree(3, 1, 64, seed=23)
    topo = topology_from_spec(ids, parent, children)
    _, _, W, H = oracle_problem(3, 1, 64, seed=23)
    op_topo, Wp = reversed_children_operator(topo, W)
    st = OracleState(topo, tensors)
    tdvp = OracleTDVP(st, Wp, 0.05, kind=kind, krylov_tol=1e-13, max_krylov_dim=40,
                      svd_params=dict(max_bond_dim=1000, rel_tol=1e-15, total_tol=1e-15),
                      op_topo=op_topo)
    tdvp.s.normalise()
    psi0 = OD.state_to_vector(topo, tdvp.s.A)
    for _ in range(2):
        tdvp.run_one_time_step()
    psi = OD.state_to_vector(topo, tdvp.s.A)
    np.testing.assert_allclose(psi, expm(-1j * 0.1 * H) @ psi0, atol=1e-8)


# ---- reference-order flop counts ----------------------------------------------------------
def _cfg_shapes(branching, depth, chi, D):
    ids, parent, children = complete_tree(branching, depth)
    b = bond_dims(ids, parent, children, chi)
    topo = topology_from_spec(ids, parent, children)
    sh = {n: tuple(([b[n]] if parent[n] else []) + [b[c] for c in children[n]] + [2])
          for n in ids}
    wsh = {n: tuple(([D] if parent[n] else []) + [D] * len(children[n]) + [2, 2]) for n in ids}
    return topo, sh, wsh


def _closed_form_heff(chis, Ds, d=2):
    """Reference order: environments absorbed one at a time (8 |current| D_i chi_i
    each, the current tensor growing by D_i), then W over (all D_i, d)."""
    cur = math.prod(chis) * d
    total = 0.0
    for c, D in zip(chis, Ds):
        total += 8.0 * cur * D * c
        cur = cur * D
    return total + 8.0 * math.prod(chis) * math.prod(Ds) * d * d


@pytest.mark.parametrize("cfg,node,expect,unit", [
    ((3, 3, 256, 5), "site0", 10718.6, 1e9),   # C3 root
    ((3, 3, 256, 5), "site1", 250.5, 1e9),     # C3 level 1
    ((2, 5, 64, 5), "site1", 42.66, 1e9),      # C2 level 1
    ((2, 3, 16, 3), "site1", 6.39, 1e6),       # C1 level 1
])
def test_flops_reference_order_heff(cfg, node, expect, unit):
    topo, sh, wsh = _cfg_shapes(*cfg)
    f = OF.heff_flops(topo, node, sh, wsh)
    assert abs(f / unit - expect) < 0.01 * max(1.0, expect / 100)
    nb = len(topo.neighbours(node))
    assert f == _closed_form_heff(sh[node][:nb], wsh[node][:nb])


def test_flops_counter_matches_real_tensordots():
    """The shape-only counter equals 8 M N K summed over the oracle's real tensordot
    calls on actual arrays (env update, H_eff apply and bond action)."""
    topo, st, W, _ = oracle_problem(2, 2, 4, seed=24)
    real = np.tensordot
    count = [0.0]

    def counting(a, b, axes):
        out = real(a, b, axes)
        ax_a = [axes[0]] if isinstance(axes[0], int) else list(axes[0])
        k = math.prod(a.shape[i] for i in ax_a)
        count[0] += 8.0 * (a.size // k) * k * (b.size // k)
        return out

    envs = OC.build_envs_to_root(topo, st.A, W)
    sh = {n: st.A[n].shape for n in topo.node_ids}
    wsh = {n: W[n].shape for n in topo.node_ids}
    nid = topo.node_ids[1]
    OC.np.tensordot = counting
    try:
        count[0] = 0.0
        OC.heff_apply(topo, topo.root, st.A[topo.root], W[topo.root], envs)
        heff_real = count[0]
        count[0] = 0.0
        OC.env_update(topo, nid, topo.root, st.A[nid], W[nid], envs)
        env_real = count[0]
        count[0] = 0.0
        x = np.ones((4, 4), dtype=complex)
        e = np.ones((4, 5, 4), dtype=complex)
        OC.bond_apply(x, e, e)
        bond_real = count[0]
    finally:
        OC.np.tensordot = real
    assert OF.heff_flops(topo, topo.root, sh, wsh) == heff_real
    assert OF.env_flops(topo, nid, topo.root, sh, wsh) == env_real
    assert OF.bond_flops(4, 5) == bond_real
