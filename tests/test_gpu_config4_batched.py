"""The batching configuration at full size: 512 independent random-field Heisenberg
instances (h_i ~ U[-2, 2], seed per instance) on the 31-node binary tree,
second-order two-site TDVP with truncation at 1e-10 and chi_max = 128, from bond
dimension 8, one step of all 512 in lockstep (``BatchedTDVP``).

Checked: the bonds grow unevenly towards 128 across instances and bonds; seeded
sample instances run alone give the same observables (1e-12) and the same kept
bond dimensions; and one instance at chi_max = 16 matches the oracle (energy
within 1e-12 max(|E|, 1), identical kept dimensions).  The reference runs such instances
one after another through ``tdvp_algorithms/twositetdvp.py:89-125``."""
import time

import numpy as np
import pytest
import torch

import pytreenet_amd as ptn
from oracle.tdvp import OracleState, OracleTDVP
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import pauli_matrices, random_field_heisenberg
from tests.helpers import host_tensors, topology_from_ttn

pytestmark = pytest.mark.gpu
COUNT = 512
SEED0 = 4000
SAMPLE = (0, 257, 511)


def _instance(b, chi_max=128, chi0=8):
    ids, parent, children = complete_tree(2, 4)
    tensors = random_state_tensors(ids, parent, children, chi0, seed=SEED0 + b)
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = random_field_heisenberg(st, 2.0, seed=SEED0 + 100000 + b)
    cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, max_bond_dim=chi_max,
                                rel_tol=1e-10, total_tol=1e-10)
    a = ptn.SecondOrderTwoSiteTDVP(st, H, 0.05, 0.05, [], config=cfg,
                                   solver_options={"max_krylov_dim": 10, "krylov_tol": 1e-10})
    c = a.state.orthogonality_center_id
    a.state.tensors[c] = a.state.tensors[c] / a.state.norm()
    return a, tensors


def _observables(a):
    st = a.state
    _, _, Z = pauli_matrices()
    zs = [st.single_site_operator_expectation_value(n, Z).real for n in list(st.nodes)[:4]]
    return np.array([ptn.expectation_value(st, a.hamiltonian).real, st.norm()] + zs)


def test_config4_512_instances_one_batched_step(gpu):
    print(f"config C4: {COUNT} instances, one batched two-site step", flush=True)
    t0 = time.time()
    algs = [_instance(b)[0] for b in range(COUNT)]
    torch.cuda.synchronize()
    print(f"  built {COUNT} instances in {time.time() - t0:.1f} s", flush=True)
    t0 = time.time()
    bt = ptn.BatchedTDVP(algs)
    bt.progress = lambda done, kind: (print(f"  {done} lockstep points ({kind}), "
                                            f"{time.time() - t0:.0f} s", flush=True)
                                      if done % 25 == 0 else None)
    bt.run(1)
    torch.cuda.synchronize()
    dt = time.time() - t0
    print(f"  batched step: {dt:.1f} s = {COUNT / dt:.2f} instance-steps/s", flush=True)
    sizes = np.array([sum(t.numel() for t in a.state.tensors.values()) for a in algs])
    bonds = np.array([sorted(a.state.bond_dims().values()) for a in algs])
    assert bonds.max() == 128 and bonds.min() <= 8
    assert len(np.unique(sizes)) > 1, "the instances' bond dimensions should differ"
    for a in algs:
        assert abs(a.state.norm() - 1.0) < 1e-10
    sample = {b: algs[b] for b in SAMPLE}
    del algs
    torch.cuda.empty_cache()
    for b, a_batch in sample.items():
        a_single = _instance(b)[0]
        a_single.run_one_time_step()
        ob, os_ = _observables(a_batch), _observables(a_single)
        np.testing.assert_allclose(ob, os_, rtol=0, atol=1e-12 * max(1.0, np.abs(os_).max()))
        assert a_batch.kept_dims == a_single.kept_dims
        print(f"  instance {b}: batched = single (E={ob[0]:.12f})", flush=True)


def test_config4_instance_matches_oracle_at_reduced_bond_dim(gpu):
    print("config C4 instance at chi_max 16 vs oracle", flush=True)
    a, tensors = _instance(SAMPLE[1], chi_max=16, chi0=8)
    W = host_tensors(a.hamiltonian)
    otd = OracleTDVP(OracleState(topology_from_ttn(a.state), tensors), W, 0.05, kind="second2",
                     krylov_tol=1e-10, max_krylov_dim=10,
                     svd_params=dict(max_bond_dim=16, rel_tol=1e-10, total_tol=1e-10))
    otd.s.normalise()
    a.run_one_time_step()
    otd.run_one_time_step()
    e, e_ref = ptn.expectation_value(a.state, a.hamiltonian), otd.energy()
    # relative to max(|E|, 1): this instance's energy (0.08) is a near-cancelling
    # sum of O(1) terms
    assert abs(e - e_ref) < 1e-12 * max(abs(e_ref), 1.0)
    dims = {k: tuple(v.shape) for k, v in a.state.tensors.items()}
    assert dims == {k: tuple(v.shape) for k, v in otd.s.A.items()}
    assert max(max(s) for s in dims.values()) == 16
