"""Pins of the CPU oracle: golden KATs from the reference tests, closed forms, brute force."""
import json
import os
import warnings

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import contract as OC
from oracle import dense as OD
from oracle.krylov import krylov_exp_action
from oracle.splitting import (qr_split, svd_split, sum_truncation_index,
                              truncate_singular_values, value_truncation)
from oracle.tdvp import OracleState, OracleTDVP
from oracle.tree import Topology
from tests.helpers import small_tree, topology_from_spec

GOLD = os.path.join(os.path.dirname(__file__), "golden")
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _load("paths.json")["cases"], ids=lambda c: c["name"])
def test_oracle_paths_match_reference_kats(case):
    topo = Topology.from_parent_map(case["root"], [tuple(e) for e in case["edges"]])
    up = topo.update_path()
    assert up == case["update_path"]
    assert topo.orthogonalisation_path(up) == case["orth_path"]
    back = list(reversed(up))
    assert back == case["back_update_path"]
    assert topo.second_order_one_site_back_orth_path(back) == case["back_orth_path_one_site"]


def test_oracle_truncation_kats():
    g = _load("truncation.json")
    s = np.array(g["s"])
    for c in g["value_truncation"]:
        assert len(value_truncation(s, c["total_tol"], c["rel_tol"])) == c["kept"]
    for c in g["sum_truncation"]:
        assert sum_truncation_index(s, c["total_tol"], True) == c["kept"]
    for c in g["truncate_singular_values"]:
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            kept, trunc = truncate_singular_values(s, max_bond_dim=c["max_bond_dim"],
                                                   rel_tol=c["rel_tol"],
                                                   total_tol=c["total_tol"])
        assert len(kept) == c["kept"]
        np.testing.assert_allclose(trunc, c["truncated"])
        assert bool(w) == c.get("warns", False)


def test_oracle_renorm_restores_two_norm():
    s = np.concatenate([[1.0], np.full(200, 1e-4)])
    kept, _ = truncate_singular_values(s, max_bond_dim=1, rel_tol=0, total_tol=0, renorm=True)
    assert abs(np.linalg.norm(kept) - np.linalg.norm(s)) < 1e-12


def _rand_herm(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return (a + a.conj().T) / 2


@pytest.mark.parametrize("sign", [-1j, 1j])
def test_oracle_krylov_to_breakdown_equals_expm(sign):
    H = _rand_herm(24, 1)
    rng = np.random.default_rng(2)
    psi = rng.standard_normal(24) + 1j * rng.standard_normal(24)
    got = krylov_exp_action(psi, lambda y: H @ y, 0.3, sign, krylov_tol=0)
    np.testing.assert_allclose(got, expm(sign * 0.3 * H) @ psi, atol=1e-11)


def test_oracle_krylov_adaptive_and_nonhermitian():
    H = _rand_herm(40, 3)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    st = {}
    got = krylov_exp_action(psi, lambda y: H @ y, 0.05, -1j, krylov_tol=1e-10, stats=st)
    assert st["m"] < 40
    np.testing.assert_allclose(got, expm(-0.05j * H) @ psi, atol=1e-9)
    A = rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20))
    v = rng.standard_normal(20) + 0j
    got = krylov_exp_action(v, lambda y: A @ y, 0.1, 1.0, krylov_tol=0, hermitian=False)
    np.testing.assert_allclose(got, expm(0.1 * A) @ v, atol=1e-10)


def test_oracle_qr_keep_and_svd_split():
    rng = np.random.default_rng(5)
    t = rng.standard_normal((2, 3, 8)) + 1j * rng.standard_normal((2, 3, 8))
    q, r = qr_split(t, [0, 1], [2], keep=True)
    assert q.shape == (2, 3, 8) and r.shape == (8, 8)   # 6 < 8: zero padded
    qm = q.reshape(6, 8)
    np.testing.assert_allclose(qm.conj().T @ qm, np.diag([1.0] * 6 + [0.0] * 2), atol=1e-13)
    np.testing.assert_allclose(np.tensordot(q, r, axes=([2], [0])), t, atol=1e-13)
    u, sv, s = svd_split(t, [0, 2], [1], max_bond_dim=2, rel_tol=0, total_tol=0)
    assert u.shape == (2, 8, 2) and sv.shape == (2, 3)
    full = np.linalg.svd(t.transpose(0, 2, 1).reshape(16, 3), compute_uv=False)
    rec = np.tensordot(u, sv, axes=([2], [0])).transpose(0, 2, 1)
    assert abs(np.linalg.norm(rec - t) - full[2]) < 1e-12


def _ttno_dense_pair(topo, seed=0):
    rng = np.random.default_rng(seed)
    site = {n: rng.uniform(-1, 1) * Z for n in topo.node_ids}
    edges = []
    for n in topo.node_ids:
        p = topo.parent[n]
        if p is not None:
            j = rng.uniform(-1, 1, 3)
            edges += [(p, n, -j[0], X, X), (p, n, -j[1], Y, Y), (p, n, -j[2], Z, Z)]
    return site, edges


def _product_ttno(ids, parent, children, tensors, site, edges):
    from pytreenet_amd.random_states import build_state
    from pytreenet_amd.ttno import build_edge_ttno_tensors
    st = build_state(ids, parent, children, tensors, device="cpu")
    return build_edge_ttno_tensors(st, site, edges)


def test_ttno_builder_matches_dense_hamiltonian():
    ids, parent, children, tensors = small_tree(2, 2, 4, seed=1)
    topo = topology_from_spec(ids, parent, children)
    site, edges = _ttno_dense_pair(topo)
    W = _product_ttno(ids, parent, children, tensors, site, edges)
    H = OD.edge_hamiltonian(topo, site, edges)
    np.testing.assert_allclose(OD.ttno_to_matrix(topo, W), H, atol=1e-13)
    assert all(w.shape[i] == 5 for n, w in W.items() for i in range(w.ndim - 2))


def test_oracle_contractions_brute_force():
    ids, parent, children, tensors = small_tree(2, 2, 4, seed=3)
    topo = topology_from_spec(ids, parent, children)
    site, edges = _ttno_dense_pair(topo, 1)
    W = _product_ttno(ids, parent, children, tensors, site, edges)
    H = OD.edge_hamiltonian(topo, site, edges)
    psi = OD.state_to_vector(topo, tensors)
    e = OC.expectation_value(topo, tensors, W)
    assert abs(e - np.vdot(psi, H @ psi)) < 1e-10 * abs(e)
    assert abs(OC.overlap(topo, tensors, tensors) - np.vdot(psi, psi)) < 1e-10 * np.vdot(psi, psi).real
    nid = ids[3]
    zi = OC.single_site_expectation(topo, tensors, nid, Z)
    Zd = OD.kron_op(len(ids), {ids.index(nid): Z})
    assert abs(zi - np.vdot(psi, Zd @ psi)) < 1e-10 * np.vdot(psi, psi).real


def test_oracle_heff_is_projected_hamiltonian():
    """At the orthogonality centre H_eff = P^H H P with P the isometric embedding."""
    ids, parent, children, tensors = small_tree(2, 2, 4, seed=4)
    topo = topology_from_spec(ids, parent, children)
    site, edges = _ttno_dense_pair(topo, 2)
    W = _product_ttno(ids, parent, children, tensors, site, edges)
    H = OD.edge_hamiltonian(topo, site, edges)
    st = OracleState(topo, tensors)
    c = ids[1]
    st.canonical_form(c)
    dist = topo.distance_to_node(c)
    envs = {}
    for d in range(max(dist.values()), 0, -1):
        for n in [m for m in dist if dist[m] == d]:
            to = min(topo.neighbours(n), key=lambda x: dist[x])
            envs[(n, to)] = OC.env_update(topo, n, to, st.A[n], W[n], envs)
    x = st.A[c]
    # P maps the centre tensor to the full vector
    basis = np.eye(x.size, dtype=complex)
    P = np.stack([OD.state_to_vector(topo, {**st.A, c: basis[i].reshape(x.shape)})
                  for i in range(x.size)], axis=1)
    np.testing.assert_allclose(P.conj().T @ P, np.eye(x.size), atol=1e-12)
    heff = P.conj().T @ H @ P
    y = OC.heff_apply(topo, c, x, W[c], envs).reshape(-1)
    np.testing.assert_allclose(y, heff @ x.reshape(-1), atol=1e-11)


@pytest.mark.parametrize("kind", ["first1", "second1", "second2", "first2"])
def test_oracle_tdvp_exact_at_full_bond_dims(kind):
    """With uncapped bond dims (no truncation), TDVP equals exact evolution."""
    ids, parent, children, tensors = small_tree(2, 2, 64, seed=5)   # 7 sites, caps not hit
    topo = topology_from_spec(ids, parent, children)
    site, edges = _ttno_dense_pair(topo, 3)
    W = _product_ttno(ids, parent, children, tensors, site, edges)
    H = OD.edge_hamiltonian(topo, site, edges)
    st = OracleState(topo, tensors)
    params = dict(max_bond_dim=1000, rel_tol=1e-15, total_tol=1e-15)
    tdvp = OracleTDVP(st, W, 0.05, kind=kind, krylov_tol=1e-13, max_krylov_dim=40,
                      svd_params=params)
    tdvp.s.normalise()
    psi0 = OD.state_to_vector(topo, tdvp.s.A)
    for _ in range(3):
        tdvp.run_one_time_step()
    psi = OD.state_to_vector(topo, tdvp.s.A)
    exact = expm(-1j * 0.15 * H) @ psi0
    assert abs(abs(np.vdot(exact, psi)) - 1) < 1e-9
    np.testing.assert_allclose(psi, exact, atol=1e-8)
    e0 = np.vdot(psi0, H @ psi0).real
    assert abs(tdvp.energy().real - e0) < 1e-8
