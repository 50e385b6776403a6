"""Host-side logic of the product (no GPU compute): paths, truncation rules, C ABI exports."""
import ctypes
import json
import os
import re
import warnings

import numpy as np
import pytest

import pytreenet_amd as ptn
from pytreenet_amd.random_states import bond_dims, complete_tree
from pytreenet_amd.time_evolution import (find_orthogonalisation_path,
                                          find_update_path_leaf_to_root)

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cpu_state(root, edges, chi=2):
    st = ptn.TreeTensorNetworkState()
    nchildren = {}
    for c, p in edges:
        nchildren[p] = nchildren.get(p, 0) + 1
    def shape(nid, has_parent):
        return ([chi] if has_parent else []) + [chi] * nchildren.get(nid, 0) + [2]
    import torch
    st.add_root(ptn.Node(identifier=root), torch.zeros(shape(root, False), dtype=torch.complex128))
    for c, p in edges:
        t = torch.zeros(shape(c, True), dtype=torch.complex128)
        st.add_child_to_parent(ptn.Node(identifier=c), t, 0, p, st.nodes[p].nvirt_legs())
    return st


with open(os.path.join(GOLD, "paths.json")) as f:
    CASES = json.load(f)["cases"]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_product_paths_match_reference_kats(case):
    st = _cpu_state(case["root"], case["edges"])
    up = find_update_path_leaf_to_root(st)
    assert up == case["update_path"]
    assert find_orthogonalisation_path(up, st) == case["orth_path"]


def test_add_child_to_parent_leg_order():
    import torch
    st = ptn.TreeTensorNetworkState()
    root = torch.arange(2 * 3 * 4, dtype=torch.float64).reshape(2, 3, 4).to(torch.complex128)
    st.add_root(ptn.Node(identifier="r"), root)
    child = torch.zeros((5, 3), dtype=torch.complex128)       # leg 1 meets the root
    st.add_child_to_parent(ptn.Node(identifier="c"), child, 1, "r", 1)
    assert st.nodes["r"].children == ["c"]
    assert st.tensors["r"].shape == (3, 2, 4)                  # (child, open...)
    assert st.tensors["c"].shape == (3, 5)                     # (parent, open)
    assert torch.equal(st.tensors["r"], root.permute(1, 0, 2))


def test_product_truncation_rules_match_kats():
    with open(os.path.join(GOLD, "truncation.json")) as f:
        g = json.load(f)
    s = np.array(g["s"])
    for c in g["value_truncation"]:
        assert len(ptn.value_truncation(s, c["total_tol"], c["rel_tol"])) == c["kept"]
    for c in g["sum_truncation"]:
        assert len(ptn.sum_truncation(s, c["total_tol"])) == c["kept"]
    for c in g["truncate_singular_values"]:
        params = ptn.SVDParameters(c["max_bond_dim"], c["rel_tol"], c["total_tol"])
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            kept, trunc = ptn.truncate_singular_values(s, params)
        assert len(kept) == c["kept"]
        np.testing.assert_allclose(trunc, c["truncated"])
        assert bool(w) == c.get("warns", False)


def test_svd_parameter_validation():
    with pytest.raises(TypeError):
        ptn.SVDParameters(max_bond_dim=2.5)
    with pytest.raises(ValueError):
        ptn.SVDParameters(rel_tol=-1)
    cfg = ptn.TwoSiteTDVPConfig(max_bond_dim=128, rel_tol=1e-10, total_tol=1e-10)
    assert cfg.max_bond_dim == 128 and cfg.time_evo_mode is ptn.TimeEvoMode.FASTEST


def test_step_count_rule():
    from pytreenet_amd.time_evolution import TimeEvolution
    class T(TimeEvolution):
        pass
    import torch
    st = _cpu_state("r", [("a", "r")])
    assert T(st, 0.01, 0.2, []).num_time_steps == 20
    assert T(st, 0.3, 1.0, []).num_time_steps == 4   # 3.33.. -> frac 0.33 >= 0.1 -> +1


def test_bond_dim_caps_of_benchmark_trees():
    ids, parent, children = complete_tree(3, 3)
    b = bond_dims(ids, parent, children, 256)
    assert len(ids) == 40
    assert sorted(set(b.values())) == [2, 16, 256]
    ids, parent, children = complete_tree(2, 5)
    b = bond_dims(ids, parent, children, 64)
    assert len(ids) == 63 and sorted(set(b.values())) == [2, 8, 64]


def test_network_planner_optimal_root_cost():
    """Flop-optimal order at the ternary root (chi=256, D=5, d=2): 2472.3 GFLOP."""
    from pytreenet_amd.engine import plan_network
    labels = [[("a", i) for i in range(3)] + ["p"],
              [("w", i) for i in range(3)] + ["q", "p"]] + \
             [[("a", i), ("w", i), ("b", i)] for i in range(3)]
    dims = {("a", i): 256 for i in range(3)}
    dims.update({("b", i): 256 for i in range(3)})
    dims.update({("w", i): 5 for i in range(3)})
    dims.update({"p": 2, "q": 2})
    plan = plan_network(labels, dims, [("b", i) for i in range(3)] + ["q"])
    assert abs(plan.flops / 1e9 - 2472.3) < 0.1
    assert len(plan.steps) == 4


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "ptn_hip.h")).read()
    declared = set(re.findall(r"\b(ptn_[a-z0-9_]+)\s*\(", header))
    from pytreenet_amd import _lib
    assert declared == set(_lib.EXPORTED)
    if not _lib.LIB_PATH.exists():
        pytest.skip("libptnhip.so not built")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert _lib.load().ptn_version() >= 100


def test_product_refuses_cpu_tensors():
    import torch
    from pytreenet_amd import engine as E
    a = torch.zeros((2, 2), dtype=torch.complex128)
    with pytest.raises(ptn.PtnBackendError):
        E.contract(a, "ij", a, "jk", "ik")


def test_network_planner_block_sparse_root_cost():
    """With the XYZ TTNO's 12 live bond-state blocks the root apply needs 1312.1 GFLOP."""
    import torch
    from pytreenet_amd.engine import BlockSparse, plan_network
    from pytreenet_amd.random_states import bond_dims, build_state
    ids, parent, children = complete_tree(3, 3)
    b = bond_dims(ids, parent, children, 256)
    tensors = {n: np.zeros(([b[n]] if parent[n] else []) + [b[x] for x in children[n]] + [2])
               for n in ids}
    st = build_state(ids, parent, children, tensors, device="cpu")
    H = ptn.random_xyz_model(st, 1)
    bs = BlockSparse.from_dense(H.tensors["site0"], range(3))
    assert len(bs.tuples) == 12
    labels = [[("a", i) for i in range(3)] + ["p"], [("w", i) for i in range(3)] + ["q", "p"]] + \
             [[("a", i), ("w", i), ("b", i)] for i in range(3)]
    dims = {("a", i): 256 for i in range(3)}
    dims.update({("b", i): 256 for i in range(3)})
    dims.update({("w", i): 5 for i in range(3)})
    dims.update({"p": 2, "q": 2})
    plan = plan_network(labels, dims, [("b", i) for i in range(3)] + ["q"],
                        (1, [("w", i) for i in range(3)], bs.tuples))
    assert abs(plan.flops / 1e9 - 1312.1) < 0.1


def test_ttn_save_load_round_trip(tmp_path):
    """The reference's two interchange files: tensors under node ids in canonical leg
    order (.npz) plus the structure (.json); the round trip keeps everything."""
    from pytreenet_amd.random_states import build_state, random_state_tensors
    ids, parent, children = complete_tree(2, 2)
    tensors = random_state_tensors(ids, parent, children, 4, seed=3)
    st = build_state(ids, parent, children, tensors, device="cpu")
    st.orthogonality_center_id = ids[1]
    path = str(tmp_path / "state")
    st.save(path)
    meta = json.load(open(path + ".json"))
    assert meta["root_id"] == st.root_id and meta["orthogonality_center_id"] == ids[1]
    assert set(meta["nodes"]) == set(ids)
    for nid in ids:
        assert meta["nodes"][nid]["children"] == list(st.nodes[nid].children)
        assert meta["nodes"][nid]["shape"] == list(st.tensors[nid].shape)
    back = ptn.TreeTensorNetworkState.load(path, device="cpu")
    assert isinstance(back, ptn.TreeTensorNetworkState)
    assert back.root_id == st.root_id and back.orthogonality_center_id == ids[1]
    for nid in ids:
        assert back.nodes[nid].parent == st.nodes[nid].parent
        assert back.nodes[nid].children == st.nodes[nid].children
        assert np.array_equal(back.tensors[nid].numpy(), st.tensors[nid].numpy())


def test_ttn_load_rejects_missing_and_inconsistent_files(tmp_path):
    with pytest.raises(FileNotFoundError):
        ptn.TreeTensorNetworkState.load(str(tmp_path / "nothing"), device="cpu")
    np.savez(str(tmp_path / "bad.npz"), r=np.ones((3, 2), complex), c=np.ones((2, 2), complex))
    json.dump({"root_id": "r", "orthogonality_center_id": None,
               "nodes": {"r": {"identifier": "r", "parent": None, "children": ["c"],
                               "shape": [3, 2]},
                         "c": {"identifier": "c", "parent": "r", "children": [],
                               "shape": [2, 2]}}}, open(str(tmp_path / "bad.json"), "w"))
    with pytest.raises(ValueError, match="bond r-c"):
        ptn.TreeTensorNetworkState.load(str(tmp_path / "bad"), device="cpu")


with open(os.path.join(GOLD, "effective_hamiltonian.json")) as f:
    EFF = json.load(f)


def _node(nid, nbs, n_open, has_parent):
    n = ptn.Node(identifier=nid)
    if has_parent:
        n.parent, n.children = nbs[0], list(nbs[1:])
    else:
        n.parent, n.children = None, list(nbs)
    n.shape = (2,) * (len(nbs) + n_open)
    return n


@pytest.mark.parametrize("case", EFF["leg_permutations"], ids=lambda c: c["name"])
def test_product_leg_permutation_kats(case):
    from pytreenet_amd.effective_hamiltonians import find_tensor_leg_permutation
    has_parent = not case["name"].startswith("root")
    s = _node("node", case["state_nbs"], case["n_open"], has_parent)
    h = _node("node", case["ham_nbs"], case["n_open"], has_parent)
    assert find_tensor_leg_permutation(s, h) == tuple(case["perm"])


def test_ttno_from_tensors_own_children_order_and_checks():
    """A TTNO may list its children in its own order (legs permuted to match); bond,
    leg-count and physical-dimension mismatches raise."""
    import torch
    st = _cpu_state("r", [("a", "r"), ("b", "r"), ("c", "a")], chi=3)
    rng = np.random.default_rng(0)
    W = {"r": rng.standard_normal((4, 5, 2, 2)), "a": rng.standard_normal((4, 6, 2, 2)),
         "b": rng.standard_normal((5, 2, 2)), "c": rng.standard_normal((6, 2, 2))}
    op = ptn.TTNO.from_tensors(st, W)
    assert op.nodes["r"].children == ["a", "b"]
    Wrev = dict(W, r=np.transpose(W["r"], (1, 0, 2, 3)))
    op2 = ptn.TTNO.from_tensors(st, Wrev, children={"r": ["b", "a"]})
    assert op2.nodes["r"].children == ["b", "a"] and op2.nodes["r"].neighbour_index("a") == 1
    op3 = op.with_children_order({"r": ["b", "a"]})
    assert torch.equal(op3.tensors["r"], op2.tensors["r"])
    with pytest.raises(ValueError):          # children order given, legs not permuted
        ptn.TTNO.from_tensors(st, W, children={"r": ["b", "a"]})
    with pytest.raises(ValueError):          # not a reordering of the state's children
        ptn.TTNO.from_tensors(st, W, children={"r": ["a", "c"]})
    with pytest.raises(ValueError):          # wrong physical dimension
        ptn.TTNO.from_tensors(st, dict(W, c=rng.standard_normal((6, 3, 3))))
    with pytest.raises(ValueError):          # missing leg
        ptn.TTNO.from_tensors(st, dict(W, b=rng.standard_normal((5, 2))))


def test_krylov_hessenberg_buffer_grows_by_doubling():
    """The Krylov buffers start small and double (no n x n allocation when
    max_krylov_dim is None); the grown buffer keeps the old entries."""
    from pytreenet_amd.krylov import KRYLOV_START_CAP, _grow_hess
    h = np.arange(9, dtype=complex).reshape(3, 3)
    g = _grow_hess(h, 5)
    assert g.shape == (6, 6) and np.array_equal(g[:3, :3], h) and not g[3:].any()
    assert KRYLOV_START_CAP <= 32


def test_interchange_files_between_product_and_oracle(tmp_path):
    """The oracle's reader/writer of the reference's two files agrees with the
    product's save/load, children order included."""
    import torch
    from oracle.io import load_ttn, save_ttn
    from pytreenet_amd.random_states import build_state, random_state_tensors
    from tests.helpers import reversed_children_operator, topology_from_spec
    ids, parent, children = complete_tree(2, 2)
    tensors = random_state_tensors(ids, parent, children, 3, 5)
    st = build_state(ids, parent, children, tensors, device="cpu")
    st.save(str(tmp_path / "s"))
    topo, A, center = load_ttn(str(tmp_path / "s"))
    assert center is None and topo.root == "site0"
    assert all(np.array_equal(A[n], tensors[n]) for n in ids)
    assert all(topo.children[n] == children[n] for n in ids)
    W = {n: np.asarray(t) for n, t in ptn.transverse_field_ising(st).tensors.items()}
    op_topo, Wp = reversed_children_operator(topology_from_spec(ids, parent, children), W)
    save_ttn(str(tmp_path / "w"), op_topo, Wp, center="site1")
    H = ptn.TTNO.load(str(tmp_path / "w"), device="cpu")
    assert H.nodes["site0"].children == ["site2", "site1"]
    assert H.orthogonality_center_id == "site1"
    assert all(torch.equal(H.tensors[n], torch.as_tensor(Wp[n])) for n in ids)


def test_device_built_offset_tables_equal_host_tables():
    """Large offset tables are built with elementwise tensor ops instead of numpy; the
    builder runs on any torch device, here the CPU, and must give the host tables."""
    import torch
    from pytreenet_amd import engine as E
    cases = [((4,), (3,)), ((3, 5), (10, 1)), ((2, 3, 4), (1, 100, 7)),
             ((7, 1, 3), (5, 999, 2)), ((64, 2, 64), (2, 1, 8192))]
    for dims, strides in cases:
        out = torch.empty(int(np.prod(dims)), dtype=torch.int64)
        E._device_offsets(dims, strides, out)
        assert np.array_equal(out.numpy(), E._offsets(dims, strides)), (dims, strides)


def _dense_of(t):
    from pytreenet_amd import engine as E
    if isinstance(t, E.BlockSparse):
        full = np.zeros(t.shape, dtype=complex)
        data = t.data.numpy()
        for r, tup in enumerate(t.tuples):
            idx = [slice(None)] * len(t.shape)
            for a, v in zip(t.block_axes, tup):
                idx[a] = int(v)
            full[tuple(idx)] = data[r]
        return full
    return t.numpy()


def _einsum_network(tensors, labels, out_labels, conj):
    letters = {}
    for lab in list(labels) + [out_labels]:
        for l in lab:
            letters.setdefault(repr(l), "abcdefghijklmnopqrstuvwxyzABCDEFGHIJ"[len(letters)])
    spec = ",".join("".join(letters[repr(l)] for l in lab) for lab in labels) + "->" + \
        "".join(letters[repr(l)] for l in out_labels)
    arrs = [np.conj(_dense_of(t)) if c else _dense_of(t) for t, c in zip(tensors, conj)]
    return np.einsum(spec, *arrs, optimize="optimal")


@pytest.mark.parametrize("model", ["xyz", "field"])
def test_identity_split_sums_to_the_network(model):
    """The parts of a network split at its identity environment blocks, each
    contracted on the host with einsum and added into its slice of the output,
    equal the whole network, for the effective Hamiltonians and every
    environment update of the root, an inner node and a leaf of a ternary tree
    (environments random except for their identity blocks).  The identity
    states are those of the edge-term construction: 0 towards a parent, 1
    towards a child."""
    import torch
    from pytreenet_amd import engine as E
    from pytreenet_amd.contractions import _env_labels, _op_labels, heff_network
    from pytreenet_amd.random_states import build_state, random_state_tensors
    ids, parent, children = complete_tree(3, 2)
    st = build_state(ids, parent, children,
                     random_state_tensors(ids, parent, children, 3, seed=0), device="cpu")
    edges = [(n.parent, nid) for nid, n in st.nodes.items() if n.parent]
    H = ptn.random_xyz_model(st, seed=1) if model == "xyz" else \
        ptn.xyz_model(st, {e: (0.3, -0.7, 1.1) for e in edges}, {n: 0.4 for n in st.nodes})
    rng = np.random.default_rng(0)
    for nid, node in st.nodes.items():
        for nb in node.neighbouring_nodes():
            assert H.identity_states(nid, nb) == ((0,) if node.parent == nb else (1,))
    for nid in ["site0", "site1", "site4"]:
        node, opn = st.nodes[nid], H.nodes[nid]
        W = E.BlockSparse.from_dense(H.tensors[nid], range(opn.nneighbours()))
        nbs = node.neighbouring_nodes()
        ident = {2 + i: H.identity_states(nb, nid) for i, nb in enumerate(nbs)}
        envs = []
        for i, nb in enumerate(nbs):
            chi, D = 2 + i % 2, H.tensors[nid].shape[opn.neighbour_index(nb)]
            e = rng.standard_normal((chi, D, chi)) + 1j * rng.standard_normal((chi, D, chi))
            for s in ident[2 + i]:
                e[:, s, :] = np.eye(chi)
            envs.append(torch.as_tensor(e))
        x = torch.as_tensor(rng.standard_normal(tuple(e.shape[0] for e in envs) + (2,)) + 0j)

        def check(ts, ls, out, cj, idd):
            ref = _einsum_network(ts, ls, out, cj)
            parts = E._split_identity(ts, ls, out, cj, idd)
            assert parts is not None and parts[0].full
            tot = np.zeros(ref.shape, dtype=complex)
            for p in parts:
                idx = [slice(None)] * tot.ndim
                for ax, v in p.fixed:
                    idx[ax] = v
                tot[tuple(idx)] += _einsum_network(p.tensors, p.labels, p.out_labels, p.conj)
            np.testing.assert_allclose(tot, ref, rtol=0, atol=1e-13 * np.abs(ref).max())

        tensors, labels, out = heff_network(node, opn, W, envs)
        tensors[0] = x
        check(tensors, labels, out, [False] * len(tensors), ident)
        k = len(nbs)
        for to in range(k):
            ts = [x, W, x]
            ls = [[("a", i) for i in range(k)] + ["p"], _op_labels(node, opn),
                  [("b", i) for i in range(k)] + ["q"]]
            idd = {}
            for i in range(k):
                if i != to:
                    idd[len(ts)] = ident[2 + i]
                    ts.append(envs[i])
                    ls.append(list(_env_labels(i)))
            check(ts, ls, list(_env_labels(to)), [False, False, True] + [False] * (len(ts) - 3),
                  idd)


def test_two_site_operator_tensor_and_its_split():
    """The merged two-site operator (block sparse over the outer bonds) equals the two
    operator tensors contracted over their bond, and the two-site network split at
    its identity environment blocks sums to the network with both tensors."""
    import torch
    from pytreenet_amd import engine as E
    from pytreenet_amd.contractions import _op_labels
    from pytreenet_amd.random_states import build_state, random_state_tensors
    ids, parent, children = complete_tree(3, 2)
    st = build_state(ids, parent, children,
                     random_state_tensors(ids, parent, children, 3, seed=0), device="cpu")
    edges = [(n.parent, nid) for nid, n in st.nodes.items() if n.parent]
    H = ptn.xyz_model(st, {e: (0.3, -0.7, 1.1) for e in edges}, {n: 0.4 for n in st.nodes})
    rng = np.random.default_rng(3)
    for n1, n2 in [("site0", "site1"), ("site1", "site4"), ("site2", "site0")]:
        a1, a2 = H.nodes[n1], H.nodes[n2]
        nb1 = [n for n in st.nodes[n1].neighbouring_nodes() if n != n2]
        nb2 = [n for n in st.nodes[n2].neighbouring_nodes() if n != n1]
        k1, k2 = len(nb1), len(nb2)
        W12 = H.two_site_block_tensor(n1, n2, nb1, nb2, torch.complex128)
        envs, ident = [], {}
        for i, (nb, me) in enumerate([(nb, n1) for nb in nb1] + [(nb, n2) for nb in nb2]):
            node = H.nodes[me]
            chi, D = 2 + i % 2, H.tensors[me].shape[node.neighbour_index(nb)]
            e = rng.standard_normal((chi, D, chi)) + 1j * rng.standard_normal((chi, D, chi))
            ident[2 + i] = H.identity_states(nb, me)
            for s in ident[2 + i]:
                e[:, s, :] = np.eye(chi)
            envs.append(torch.as_tensor(e))
        x = torch.as_tensor(rng.standard_normal(tuple(e.shape[0] for e in envs) + (2, 2)) + 0j)
        env_l = [[("a", i), ("w", i), ("b", i)] for i in range(k1 + k2)]
        x_l = [("a", i) for i in range(k1 + k2)] + ["p1", "p2"]
        out = [("b", i) for i in range(k1 + k2)] + ["q1", "q2"]

        def op_lab(node, own, offset, other, qo, pi):
            lab = [None] * (node.nneighbours() + 2)
            for i, nb in enumerate(own):
                lab[node.neighbour_index(nb)] = ("w", offset + i)
            lab[node.neighbour_index(other)] = "wl"
            lab[-2], lab[-1] = qo, pi
            return lab
        ref = _einsum_network([x, H.tensors[n1], H.tensors[n2]] + envs,
                              [x_l, op_lab(a1, nb1, 0, n2, "q1", "p1"),
                               op_lab(a2, nb2, k1, n1, "q2", "p2")] + env_l,
                              out, [False] * (3 + k1 + k2))
        split_l = [x_l, [("w", i) for i in range(k1 + k2)] + ["q1", "q2", "p1", "p2"]] + env_l
        whole = _einsum_network([x, W12] + envs, split_l, out, [False] * (2 + k1 + k2))
        np.testing.assert_allclose(whole, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        parts = E._split_identity([x, W12] + envs, split_l, out, [False] * (2 + k1 + k2), ident)
        assert parts is not None
        tot = sum(_einsum_network(p.tensors, p.labels, p.out_labels, p.conj) for p in parts)
        np.testing.assert_allclose(tot, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_krylov_early_queueing_rule():
    """When a long-vector Krylov solve queues the next application before reading a
    step's scalars: always without a tolerance or before the first estimate, after
    it only far above the tolerance, later when the estimate extrapolated two steps
    at the rate of the last two stays above it; never when switched off."""
    from pytreenet_amd import krylov as K
    inf = float("inf")
    tol = 1e-10
    assert K._stop_unlikely(inf, inf, tol)                 # step 0
    assert K._stop_unlikely(0.5, 0.4, 0.0)                 # no tolerance
    assert K._stop_unlikely(1e-3, inf, tol)                # step 1, far above
    assert not K._stop_unlikely(1e-5, inf, tol)            # step 1, within 1e6 of tol
    # rate 1e-2: 1e-4 -> 1e-8 in two steps (above), 1e-6 -> 1e-10 (not above)
    assert K._stop_unlikely(1e-4, 1e-2, tol)
    assert not K._stop_unlikely(1e-6, 1e-4, tol)
    # an estimate that grew is extrapolated flat
    assert K._stop_unlikely(1e-9, 1e-12, tol)
    assert not K._stop_unlikely(1e-11, 1e-12, tol)
    old = K.SPECULATE_STEPS
    try:
        K.SPECULATE_STEPS = 0
        assert not K._stop_unlikely(inf, inf, tol)
    finally:
        K.SPECULATE_STEPS = old
