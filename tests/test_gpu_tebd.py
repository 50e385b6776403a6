"""TEBD on the device against brute force and the oracle.

The small cases follow the reference's ``tests/time_evolution/test_tebd.py``
(``TestTEBDsmall``: root (5, 6, 2), c1 (5, 3), c2 (6, 4); gates on the root,
on (root, c1) and on (c2, root); no truncation): the device state must equal
the gate product applied to the dense state vector.  Longer runs with SWAPs and
truncation compare the dense vectors and kept bond dimensions with the oracle's
TEBD sweep."""
import numpy as np
import pytest
import torch
from scipy.linalg import expm

import pytreenet_amd as ptn
from oracle import dense as OD
from oracle import tebd as OT
from oracle.tdvp import OracleState
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import PAULI_X, PAULI_Z
from tests.helpers import host_tensors, topology_from_ttn

pytestmark = pytest.mark.gpu
NO_TRUNC = ptn.SVDParameters(float("inf"), float("-inf"), float("-inf"))


def _say(msg):
    print(msg, flush=True)


def _crandn(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def _dense(st):
    return OD.state_to_vector(topology_from_ttn(st), host_tensors(st))


def _small(seed):
    rng = np.random.default_rng(seed)
    ids = ["root", "c1", "c2"]
    parent = {"root": None, "c1": "root", "c2": "root"}
    children = {"root": ["c1", "c2"], "c1": [], "c2": []}
    tensors = {"root": _crandn(rng, 5, 6, 2), "c1": _crandn(rng, 5, 3), "c2": _crandn(rng, 6, 4)}
    st = build_state(ids, parent, children, tensors, device="cuda")
    tps = [ptn.TensorProduct({"root": _crandn(rng, 2, 2)}),
           ptn.TensorProduct({"root": _crandn(rng, 2, 2), "c1": _crandn(rng, 3, 3)}),
           ptn.TensorProduct({"c2": _crandn(rng, 4, 4), "root": _crandn(rng, 2, 2)})]
    meas = [ptn.TensorProduct({"c1": _crandn(rng, 3, 3)})]
    trotter = ptn.TrotterSplitting([ptn.TrotterStep(tp, 0.2) for tp in tps])
    tebd = ptn.TEBD(st, trotter, 0.1, 1.0, meas, svd_parameters=NO_TRUNC)
    return st, tps, tebd


def test_tebd_init_and_exponents(gpu):
    _say("tebd init")
    st, tps, tebd = _small(0)
    assert tebd.trotter_splitting is not None and tebd.svd_parameters is NO_TRUNC
    assert tebd.time_step_size == 0.1 and tebd.final_time == 1.0
    f = -1j * 0.1 * 0.2
    np.testing.assert_allclose(tebd.exponents[0].operator, expm(f * tps[0]["root"]),
                               atol=1e-14)
    np.testing.assert_allclose(tebd.exponents[1].operator.reshape(6, 6),
                               expm(f * np.kron(tps[1]["root"], tps[1]["c1"])), atol=1e-14)
    np.testing.assert_allclose(tebd.exponents[2].operator.reshape(8, 8),
                               expm(f * np.kron(tps[2]["c2"], tps[2]["root"])), atol=1e-14)
    assert tebd.exponents[2].node_identifiers == ["c2", "root"]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_tebd_single_trotter_step_matches_dense(gpu, which):
    """One gate: on the root; on (root, c1) in node order; on (c2, root) against it."""
    _say(f"tebd single step {which}")
    st, tps, tebd = _small(10 + which)
    psi0 = _dense(st)
    tebd._apply_one_trotter_step(tebd.exponents[which])
    g = tebd.exponents[which]
    ref = OT.apply_gates_dense(psi0, ["root", "c1", "c2"], [2, 3, 4],
                               [(g.node_identifiers, g.operator)])
    got = _dense(tebd.state)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # the initial state is untouched
    np.testing.assert_array_equal(_dense(tebd.initial_state), psi0)
    # the reference test's own construction (test_tebd.py:77-192)
    f = -1j * 0.1 * 0.2
    if which == 0:
        U = np.kron(expm(f * tps[0]["root"]), np.eye(12))
        np.testing.assert_allclose(got, U @ psi0, atol=1e-12 * np.abs(ref).max())
    elif which == 1:
        U = np.kron(expm(f * np.kron(tps[1]["root"], tps[1]["c1"])), np.eye(4))
        np.testing.assert_allclose(got, U @ psi0, atol=1e-12 * np.abs(ref).max())
    else:
        U = np.kron(expm(f * np.kron(tps[2]["c2"], tps[2]["root"])), np.eye(3))
        p = psi0.reshape(2, 3, 4).transpose(2, 0, 1).reshape(24)
        want = (U @ p).reshape(4, 2, 3).transpose(1, 2, 0).reshape(24)
        np.testing.assert_allclose(got, want, atol=1e-12 * np.abs(ref).max())
    # no truncation keeps min(rows, cols) of the split: (6*2) x 3 and 4 x (5*2)
    assert tebd.kept_dims == {0: [], 1: [3], 2: [4]}[which]


def test_tebd_run_one_time_step_and_three_site_error(gpu):
    _say("tebd full step")
    st, tps, tebd = _small(20)
    psi0 = _dense(st)
    f = -1j * 0.1 * 0.2
    op1 = np.kron(expm(f * tps[0]["root"]), np.eye(12))
    op2 = np.kron(expm(f * np.kron(tps[1]["root"], tps[1]["c1"])), np.eye(4))
    op3 = expm(f * np.kron(tps[2]["root"], np.kron(np.eye(3), tps[2]["c2"])))
    tebd.run_one_time_step()
    ref = op3 @ op2 @ op1 @ psi0
    np.testing.assert_allclose(_dense(tebd.state), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    op = ptn.NumericOperator(_crandn(np.random.default_rng(0), 24, 24), ["root", "c1", "c2"])
    with pytest.raises(NotImplementedError):
        tebd._apply_one_trotter_step(op)


def _tfim_splitting(st, dt_factor=1.0):
    """Field terms, then every edge's XX term, then SWAPs on two edges and back."""
    tps, split = [], []
    for nid in st.nodes:
        tps.append(ptn.TensorProduct({nid: -0.5 * PAULI_Z}))
    for nid, node in st.nodes.items():
        if node.parent is not None:
            tps.append(ptn.TensorProduct({node.parent: -PAULI_X, nid: PAULI_X}))
    edges = [(n.parent, k) for k, n in st.nodes.items() if n.parent is not None]
    swaps = ptn.SWAPlist(edges[:2])
    back = ptn.SWAPlist(list(reversed(edges[:2])))
    before = [ptn.SWAPlist()] * len(tps)
    after = [ptn.SWAPlist()] * len(tps)
    before[-1], after[-1] = swaps, back
    split = [(i, dt_factor) for i in range(len(tps))]
    return ptn.TrotterSplitting.from_lists(tps, split, before, after)


@pytest.mark.parametrize("params", [None, (6, 1e-10, 1e-10), (3, 1e-12, 1e-12)])
def test_tebd_matches_oracle_on_binary_tree(gpu, params):
    _say(f"tebd binary depth 3 params={params}")
    ids, parent, children = complete_tree(2, 3)
    tensors = random_state_tensors(ids, parent, children, 8, 5)
    st = build_state(ids, parent, children, tensors, device="cuda")
    svd = NO_TRUNC if params is None else ptn.SVDParameters(*params)
    trotter = _tfim_splitting(st)
    zs = {f"Z_{n}": ptn.TensorProduct({n: PAULI_Z}) for n in ("site0", "site3")}
    tebd = ptn.TEBD(st, trotter, 0.05, 0.1, zs, svd_parameters=svd)
    topo = topology_from_ttn(st)
    ost = OracleState(topo, dict(tensors))
    gates = [(g.node_identifiers, g.operator) for g in tebd.exponents]
    oparams = (dict(max_bond_dim=np.inf, rel_tol=-np.inf, total_tol=-np.inf) if params is None
               else dict(max_bond_dim=params[0], rel_tol=params[1], total_tol=params[2]))
    psi = OD.state_to_vector(topo, tensors)
    for step in range(3):
        tebd.run_one_time_step()
        kept = OT.run_one_time_step(ost, gates, oparams)
        assert tebd.kept_dims[-len(kept):] == kept, f"step {step}"
        ref = OD.state_to_vector(topo, ost.A)
        got = _dense(tebd.state)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())
        if params is None:
            psi = OT.apply_gates_dense(psi, topo.node_ids, [2] * len(ids), gates)
            np.testing.assert_allclose(got, psi, rtol=0, atol=1e-10 * np.abs(psi).max())


def test_tebd_run_records_expectation_values(gpu):
    """TEBD.run: <Z> at t = 0, dt, 2 dt from the device state equal the dense ones."""
    _say("tebd run")
    ids, parent, children = complete_tree(2, 2)
    tensors = random_state_tensors(ids, parent, children, 4, 9)
    st = build_state(ids, parent, children, tensors, device="cuda")
    trotter = _tfim_splitting(st)
    zs = {"z1": ptn.TensorProduct({"site1": PAULI_Z})}
    tebd = ptn.TEBD(st, trotter, 0.05, 0.1, zs, svd_parameters=NO_TRUNC)
    tebd.run()
    topo = topology_from_ttn(st)
    gates = [(g.node_identifiers, g.operator) for g in tebd.exponents]
    psi = OD.state_to_vector(topo, tensors)
    pos = topo.node_ids.index("site1")
    Zfull = OD.kron_op(len(ids), {pos: PAULI_Z})
    want = []
    for i in range(3):
        if i:
            psi = OT.apply_gates_dense(psi, topo.node_ids, [2] * len(ids), gates)
        want.append(np.vdot(psi, Zfull @ psi))
    got = tebd.results.operator_result("z1")
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    np.testing.assert_allclose(tebd.results.times(), [0, 0.05, 0.1])
