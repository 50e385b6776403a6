"""Dense effective-Hamiltonian modes (FASTEST / EXPM / CHEBYSHEV / SPARSE).

The reference evolves with the dense H_eff matrix in these modes
(``effective_hamiltonians.py:20-470``, ``std_utils.py:134-175``).  Pins: the dense
matrix reproduces the network's action column by column; the device exponential equals
``scipy.linalg.expm``; TDVP in a dense mode equals TDVP with the Krylov exponential run
to exactness (``krylov_tol=0``, no dimension cap), which equals exp(-iHt) by its own
pins.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

import pytreenet_amd as ptn
from pytreenet_amd import engine as E
from pytreenet_amd.contractions import heff_network, operator_tensor
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import pauli_matrices, random_xyz_model

pytestmark = pytest.mark.gpu


def _state(depth, chi, seed):
    ids, parent, children = complete_tree(2, depth)
    st = build_state(ids, parent, children,
                     random_state_tensors(ids, parent, children, chi, seed=seed), device="cuda")
    return st


def test_dense_operator_matches_action(gpu):
    print("[dense] operator vs action", flush=True)
    st = _state(2, 4, 3)
    H = random_xyz_model(st, seed=4)
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.05, [], config=cfg)
    for nid in list(alg.state.nodes)[:3]:
        node, t = alg.state[nid]
        envs = [alg.partial_tree_cache.get_entry(nb, nid) for nb in node.neighbouring_nodes()
                if alg.partial_tree_cache.contains(nb, nid)]
        if len(envs) != node.nneighbours():
            continue
        tensors, labels, out = heff_network(node, H.nodes[nid],
                                            operator_tensor(H, nid, "fp64"), envs)
        tensors[0] = t
        net = E.CompiledNetwork(tensors, labels, out)
        M = E.dense_operator(net)
        rng = np.random.default_rng(1)
        x = torch.as_tensor(rng.standard_normal(t.shape) + 1j * rng.standard_normal(t.shape)).to(
            "cuda")
        y = net(x).reshape(-1)
        np.testing.assert_allclose((M @ x.reshape(-1)).cpu().numpy(), y.cpu().numpy(),
                                   atol=1e-12 * float(y.abs().max()))


@pytest.mark.parametrize("mode", ["EXPM", "FASTEST", "CHEBYSHEV", "SPARSE", "KRYLOV"])
def test_time_evolve_dense_vs_expm(gpu, mode):
    print(f"[dense] time_evolve {mode}", flush=True)
    rng = np.random.default_rng(24)
    A = rng.standard_normal((24, 24)) + 1j * rng.standard_normal((24, 24))
    Hm = A + A.conj().T
    psi = rng.standard_normal(24) + 1j * rng.standard_normal(24)
    m = getattr(ptn.TimeEvoMode, mode)
    opts = {"krylov_tol": 0.0} if mode == "KRYLOV" else {}
    for forward in (True, False):
        got = m.time_evolve(torch.as_tensor(psi).to("cuda"), torch.as_tensor(Hm).to("cuda"),
                            0.07, forward=forward, **opts).cpu().numpy()
        ref = scipy.linalg.expm((-1j if forward else 1j) * 0.07 * Hm) @ psi
        np.testing.assert_allclose(got, ref, atol=1e-12 * np.abs(ref).max())


def _obs(alg):
    st = alg.state
    _, _, Z = pauli_matrices()
    return np.array([ptn.expectation_value(st, alg.hamiltonian).real, st.norm()] +
                    [st.single_site_operator_expectation_value(n, Z).real for n in st.nodes])


@pytest.mark.parametrize("mode,two_site", [("EXPM", False), ("FASTEST", False),
                                           ("FASTEST", True)])
def test_tdvp_dense_modes_match_exact_krylov(gpu, mode, two_site):
    print(f"[dense] tdvp {mode} two_site={two_site}", flush=True)
    runs = []
    for m in (getattr(ptn.TimeEvoMode, mode), ptn.TimeEvoMode.KRYLOV):
        st = _state(2, 4, 7)
        H = random_xyz_model(st, seed=8)
        if two_site:
            cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=m, max_bond_dim=8, rel_tol=1e-12,
                                        total_tol=1e-12)
            cls = ptn.SecondOrderTwoSiteTDVP
        else:
            cfg = ptn.TDVPConfig(time_evo_mode=m)
            cls = ptn.SecondOrderOneSiteTDVP
        alg = cls(st, H, 0.05, 0.1, [], config=cfg,
                  solver_options={"krylov_tol": 0.0} if m is ptn.TimeEvoMode.KRYLOV else {})
        for _ in range(2):
            alg.run_one_time_step()
        runs.append(alg)
    a, b = _obs(runs[0]), _obs(runs[1])
    np.testing.assert_allclose(a, b, atol=1e-10)
    if two_site:
        assert runs[0].kept_dims == runs[1].kept_dims
