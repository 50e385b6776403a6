"""TEBD on the host side: Trotter splittings, SWAP gates and gate exponentials
(the reference's ``tests/time_evolution/test_trotter.py`` cases), and the
oracle's TEBD sweep pinned against brute force on dense state vectors."""
import numpy as np
import pytest
from scipy.linalg import expm

import pytreenet_amd as ptn
from oracle import dense as OD
from oracle import tebd as OT
from oracle.tdvp import OracleState
from oracle.tree import Topology
from pytreenet_amd.ttno import PAULI_X, PAULI_Y, PAULI_Z

# tests/time_evolution/test_trotter.py:17-22: the SWAP of two qubits
REF_SWAP = np.asarray([[1, 0, 0, 0],
                       [0, 0, 1, 0],
                       [0, 1, 0, 0],
                       [0, 0, 0, 1]]).reshape(2, 2, 2, 2)


def _crandn(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def _herm(rng, d):
    a = _crandn(rng, d, d)
    return a + a.conj().T


def _small_ttn():
    """The reference's random_small_ttns layout: root (5, 6, 2), c1 (5, 3), c2 (6, 4)
    (CPU tensors: only the structure is used here)."""
    st = ptn.TreeTensorNetworkState()
    rng = np.random.default_rng(0)
    import torch
    st.add_root(ptn.Node(identifier="root"), torch.as_tensor(_crandn(rng, 5, 6, 2)))
    st.add_child_to_parent(ptn.Node(identifier="c1"), torch.as_tensor(_crandn(rng, 5, 3)),
                           0, "root", 0)
    st.add_child_to_parent(ptn.Node(identifier="c2"), torch.as_tensor(_crandn(rng, 6, 4)),
                           0, "root", 1)
    return st


def test_swap_gate_matches_reference_fixture_and_swaps():
    assert np.array_equal(ptn.swap_gate(2).reshape(2, 2, 2, 2), REF_SWAP)
    rng = np.random.default_rng(1)
    for d in (2, 3, 5):
        S = ptn.swap_gate(d)
        u, v = _crandn(rng, d, d), _crandn(rng, d, d)
        a, b = _crandn(rng, d), _crandn(rng, d)
        np.testing.assert_allclose(S @ np.kron(a, b), np.kron(b, a), atol=1e-15)
        np.testing.assert_allclose(S @ np.kron(u, v) @ S, np.kron(v, u), atol=1e-14)
        assert np.array_equal(S @ S, np.eye(d * d))


def test_swaplist_rules():
    swaps = [("c1", "root"), ("root", "c2")]
    with pytest.raises(ValueError):
        ptn.SWAPlist(swaps + [("c1",)])
    with pytest.raises(ValueError):
        ptn.SWAPlist(swaps + [("c1", "c2", "root")])
    sl = ptn.SWAPlist(swaps)
    assert sl == swaps and len(ptn.SWAPlist()) == 0
    ops = sl.into_operators(dim=2)
    for pair, op in zip(swaps, ops):
        assert op.node_identifiers == list(pair)
        assert np.array_equal(op.operator, REF_SWAP)
    with pytest.raises(ValueError):
        sl.into_operators()
    assert ptn.SWAPlist().into_operators() == []
    st = _small_ttn()  # phys dims 2, 3, 4: no swap is compatible
    assert not sl.is_compatible_with_ttn(st)
    assert not ptn.SWAPlist([("c1", "c2")]).is_compatible_with_ttn(st)   # not neighbours
    assert not ptn.SWAPlist([("x", "root")]).is_compatible_with_ttn(st)  # unknown node


def test_trotter_step_exponentials_closed_form():
    """exp(-i f dt X(x)X) = cos(f dt) I - i sin(f dt) X(x)X, legs (out1, out2, in1, in2)
    in the tensor product's key order."""
    f, dt = 0.5, 0.2
    step = ptn.TrotterStep(ptn.TensorProduct({"c1": PAULI_X, "root": PAULI_X}), f)
    g = step.exponentiate_operator(dt, dim=2)
    assert g.node_identifiers == ["c1", "root"]
    th = f * dt
    ref = np.cos(th) * np.eye(4) - 1j * np.sin(th) * np.kron(PAULI_X, PAULI_X)
    np.testing.assert_allclose(g.operator, ref.reshape(2, 2, 2, 2), atol=1e-15)
    # non-commuting factors: Z (x) Y is ordered
    step = ptn.TrotterStep(ptn.TensorProduct({"a": PAULI_Z, "b": PAULI_Y}), 1.0)
    g = step.exponentiate_operator(0.3, dim=2).operator.reshape(4, 4)
    ref = np.cos(0.3) * np.eye(4) - 1j * np.sin(0.3) * np.kron(PAULI_Z, PAULI_Y)
    np.testing.assert_allclose(g, ref, atol=1e-15)
    # one site, dimension from a TTN
    rng = np.random.default_rng(2)
    h = _herm(rng, 3)
    step = ptn.TrotterStep(ptn.TensorProduct({"c1": h}), 0.7)
    g = step.exponentiate_operator(0.1, ttn=_small_ttn())
    w, v = np.linalg.eigh(h)
    np.testing.assert_allclose(g.operator, (v * np.exp(-1j * 0.07 * w)) @ v.conj().T,
                               atol=1e-14)
    before, after = ptn.TrotterStep(ptn.TensorProduct({"c1": h}), 1).realise_swaps(dim=2)
    assert before == [] and after == []


def test_trotter_splitting_from_lists():
    """tests/time_evolution/test_trotter.py:217-345: splitting order, factors and
    the SWAPs around each factor."""
    rng = np.random.default_rng(3)
    ops = [_herm(rng, 2) for _ in range(4)]
    tp = [ptn.TensorProduct({"c1": ops[0], "root": ops[1]}),
          ptn.TensorProduct({"c1": ops[0], "root": ops[1]}),
          ptn.TensorProduct({"root": ops[2], "c2": ops[3]})]
    swaps = ptn.SWAPlist([("c1", "root"), ("root", "c2")])
    sp = ptn.TrotterSplitting.from_lists(tp, [(0, 0.5), 1, (2, 0.2)],
                                         [swaps, ptn.SWAPlist(), swaps],
                                         [swaps, ptn.SWAPlist(), swaps])
    assert [s.factor for s in sp] == [0.5, 1, 0.2]
    assert [s.operator is t for s, t in zip(sp, tp)] == [True] * 3
    assert [len(s.swaps_before) for s in sp] == [2, 0, 2]
    sp0 = ptn.TrotterSplitting.from_lists(tp)
    assert [s.factor for s in sp0] == [1, 1, 1] and len(sp0[0].swaps_after) == 0
    gates = sp.exponentiate_splitting(0.1, dim=2)
    # swaps before, the gate, swaps after for each step
    assert [g.node_identifiers for g in gates] == [
        ["c1", "root"], ["root", "c2"], ["c1", "root"], ["c1", "root"], ["root", "c2"],
        ["c1", "root"],
        ["c1", "root"], ["root", "c2"], ["root", "c2"], ["c1", "root"], ["root", "c2"]]
    ref = expm(-1j * 0.2 * 0.1 * np.kron(ops[2], ops[3]))
    np.testing.assert_allclose(gates[8].operator.reshape(4, 4), ref, atol=1e-14)
    assert np.array_equal(gates[0].operator, REF_SWAP)


def _oracle_small_state(rng):
    topo = Topology()
    topo.add_root("root")
    topo.add_child("c1", "root")
    topo.add_child("c2", "root")
    A = {"root": _crandn(rng, 5, 6, 2), "c1": _crandn(rng, 5, 3), "c2": _crandn(rng, 6, 4)}
    return OracleState(topo, A)


def _binary7(rng, chi=4):
    topo = Topology()
    topo.add_root("s0")
    for c, p in [("s1", "s0"), ("s2", "s0"), ("s3", "s1"), ("s4", "s1"), ("s5", "s2"),
                 ("s6", "s2")]:
        topo.add_child(c, p)
    dims = {"s0": (chi, chi, 2), "s1": (chi, 2, 2, 2), "s2": (chi, 2, 2, 2)}
    A = {n: _crandn(rng, *dims.get(n, (2, 2))) for n in topo.node_ids}
    return OracleState(topo, A)


NO_TRUNC = dict(max_bond_dim=np.inf, rel_tol=-np.inf, total_tol=-np.inf)


def test_oracle_tebd_equals_dense_gate_product():
    """Without truncation the oracle's gate sweep equals the gates applied to the
    dense state vector, for gates given in and against node order, with SWAPs."""
    rng = np.random.default_rng(4)
    st = _oracle_small_state(rng)
    ids = st.topo.node_ids
    dims = [2, 3, 4]
    g1 = expm(-0.3j * _herm(rng, 2))
    g2 = expm(-0.2j * _herm(rng, 6)).reshape(2, 3, 2, 3)          # (root, c1)
    g3 = expm(-0.2j * _herm(rng, 8)).reshape(4, 2, 4, 2)          # (c2, root)
    gates = [(["root"], g1), (["root", "c1"], g2), (["c2", "root"], g3), ([], None)]
    psi = OD.state_to_vector(st.topo, st.A)
    ref = OT.apply_gates_dense(psi, ids, dims, [g for g in gates if g[0]])
    OT.run_one_time_step(st, gates, NO_TRUNC)
    np.testing.assert_allclose(OD.state_to_vector(st.topo, st.A), ref,
                               atol=1e-12 * np.abs(ref).max())
    # 7-site tree, TFIM-like Trotter gates on every edge with SWAPs
    st = _binary7(rng)
    ids = st.topo.node_ids
    gates = []
    for c in ids[1:]:
        p = st.topo.parent[c]
        h = np.kron(PAULI_X, PAULI_X) + 0.5 * np.kron(PAULI_Z, np.eye(2))
        gates.append(([p, c], expm(-0.05j * h).reshape(2, 2, 2, 2)))
        gates.append(([c, p], ptn.swap_gate(2).reshape(2, 2, 2, 2)))
    psi = OD.state_to_vector(st.topo, st.A)
    ref = OT.apply_gates_dense(psi, ids, [2] * 7, gates)
    kept = OT.run_one_time_step(st, gates, NO_TRUNC)
    np.testing.assert_allclose(OD.state_to_vector(st.topo, st.A), ref,
                               atol=1e-12 * np.abs(ref).max())
    assert len(kept) == 12
    with pytest.raises(NotImplementedError):
        OT.run_one_time_step(st, [(["s0", "s1", "s2"], np.eye(8).reshape([2] * 6))], NO_TRUNC)


def test_oracle_tebd_truncation_is_best_low_rank_split():
    """Two sites: the truncated gate step is the best rank-r approximation of the
    gate-applied (d1*chi) x (d2) state matrix (Eckart-Young), and the kept rank
    follows the truncation rules."""
    rng = np.random.default_rng(5)
    topo = Topology()
    topo.add_root("a")
    topo.add_child("b", "a")
    A = {"a": _crandn(rng, 6, 3), "b": _crandn(rng, 6, 5)}
    st = OracleState(topo, A)
    g = expm(-0.4j * _herm(rng, 15)).reshape(3, 5, 3, 5)
    psi = OT.apply_gates_dense(OD.state_to_vector(topo, st.A), ["a", "b"], [3, 5],
                               [(["a", "b"], g)])
    kept = OT.run_one_time_step(st, [(["a", "b"], g)],
                                dict(max_bond_dim=2, rel_tol=-np.inf, total_tol=-np.inf))
    assert kept == [2]
    u, s, vh = np.linalg.svd(psi.reshape(3, 5))
    best = (u[:, :2] * s[:2]) @ vh[:2]
    np.testing.assert_allclose(OD.state_to_vector(topo, st.A), best.reshape(-1), atol=1e-12)
    # relative tolerance cut
    st = OracleState(topo, dict(A))
    kept = OT.run_one_time_step(st, [(["a", "b"], g)],
                                dict(max_bond_dim=100, rel_tol=s[1] / s[0] * 0.999,
                                     total_tol=-np.inf))
    assert kept == [2]
