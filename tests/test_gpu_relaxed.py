"""Relaxed-precision Krylov (TDVPConfig(precision="relaxed")) against the oracle.

The relaxed form keeps every environment, split and Krylov vector in FP64 and
runs an application of H_eff in 3M FP32 only when its weight in the local
result (tau * coefficient estimate * ||T|| * low_eps) is below relax_tol.
The small trees here set relax_min_size = 0 (every local space may drop).
Bars: with relax_tol = 0 no application drops precision and the run is the FP64
path (1e-12 relative to the oracle); with relax_tol = 1e-12 some applications
do drop, and energies and <Z_i> stay within 1e-9 relative of the oracle, while
the plain FP32 form on the same inputs is off by more than that.
"""
import numpy as np
import pytest
import torch

import pytreenet_amd as ptn
from oracle.tdvp import OracleState, OracleTDVP
from pytreenet_amd import krylov as K
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import PAULI_Z
from tests.helpers import host_tensors, topology_from_ttn

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pair(kind, precision, relax_tol, chi=16, seed=5, steps=3, svd=None):
    ids, parent, children = complete_tree(2, 3)
    tensors = random_state_tensors(ids, parent, children, chi, seed)
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = ptn.random_xyz_model(st, seed)
    ops = {nid: ptn.TensorProduct({nid: PAULI_Z}) for nid in list(st.nodes)[:4]}
    opts = {"krylov_tol": 1e-10, "max_krylov_dim": 10}
    if kind == "second2":
        cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV,
                                    precision=precision, relax_tol=relax_tol,
                                    relax_min_size=0, **(svd or {}))
        alg = ptn.SecondOrderTwoSiteTDVP(st, H, 0.05, 0.05 * steps, ops, config=cfg,
                                         solver_options=opts)
    else:
        cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision=precision,
                             relax_tol=relax_tol, relax_min_size=0)
        alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.05 * steps, ops, config=cfg,
                                         solver_options=opts)
    otd = OracleTDVP(OracleState(topology_from_ttn(st), tensors), host_tensors(H), 0.05,
                     kind=kind, krylov_tol=1e-10, max_krylov_dim=10, svd_params=svd)
    return alg, otd, ops


def _run(alg, otd, steps):
    errs = []
    for _ in range(steps):
        alg.run_one_time_step()
        otd.run_one_time_step()
        e = ptn.expectation_value(alg.state, alg.hamiltonian)
        e_ref = otd.energy()
        errs.append(abs(e - e_ref) / abs(e_ref))
    return errs


def _z_err(alg, otd, ops):
    """Largest |<Z_i> - <Z_i>_oracle| relative to <psi|psi> (the states are not
    normalised)."""
    worst = 0.0
    for nid in ops:
        z = alg.state.single_site_operator_expectation_value(nid, PAULI_Z)
        z_ref = otd.expect_site(nid, PAULI_Z)
        worst = max(worst, abs(z - z_ref))
    return worst / otd.norm_sq()


def test_relaxed_zero_tol_is_fp64(gpu):
    _say("relaxed precision, relax_tol=0: FP64 parity 1e-12")
    alg, otd, _ = _pair("second1", "relaxed", 0.0)
    low0 = K.STATS.low_applies
    errs = _run(alg, otd, 2)
    assert K.STATS.low_applies == low0
    assert max(errs) < 1e-12, errs


def test_relaxed_one_site_error_bound(gpu):
    _say("relaxed precision one-site, relax_tol=1e-12 vs oracle and vs plain fp32")
    alg, otd, ops = _pair("second1", "relaxed", 1e-12)
    low0, app0 = K.STATS.low_applies, K.STATS.applies
    errs = _run(alg, otd, 3)
    low = K.STATS.low_applies - low0
    _say(f"  low applications {low} of {K.STATS.applies - app0}; rel dE {max(errs):.2e}")
    assert low > 0
    assert max(errs) < 1e-9, errs
    assert _z_err(alg, otd, ops) < 1e-9
    # the plain FP32 form on the same inputs is far less accurate
    a32, o32, _ = _pair("second1", "fp32", 0.0)
    errs32 = _run(a32, o32, 3)
    assert max(errs32) > 10 * max(errs), (errs32, errs)


def test_relaxed_more_low_with_looser_tol(gpu):
    _say("relaxed precision: looser relax_tol drops more applications")
    counts = []
    for tol in (1e-13, 1e-8):
        alg, otd, _ = _pair("second1", "relaxed", tol, steps=1)
        low0 = K.STATS.low_applies
        errs = _run(alg, otd, 1)
        counts.append(K.STATS.low_applies - low0)
        assert max(errs) < 1e-6, (tol, errs)
    assert counts[1] > counts[0], counts


def test_relaxed_two_site_kept_dims(gpu):
    _say("relaxed precision two-site with truncation: kept dims and energy")
    svd = {"max_bond_dim": 12, "rel_tol": 1e-10, "total_tol": 1e-10}
    alg, otd, _ = _pair("second2", "relaxed", 1e-12, chi=8, steps=2, svd=svd)
    low0 = K.STATS.low_applies
    errs = _run(alg, otd, 2)
    assert K.STATS.low_applies > low0
    for nid, t in alg.state.tensors.items():
        assert tuple(t.shape) == tuple(otd.s.A[nid].shape), nid
    assert max(errs) < 1e-9, errs


def test_relaxed_instances_batched(gpu):
    """Relaxed instances step in lockstep: local spaces from SEQUENTIAL_N on run the
    relaxed single-instance solve (applications drop to FP32), smaller ones the
    batched FP64 solve; each instance stays within 1e-9 of its oracle run."""
    _say("relaxed instances in BatchedTDVP")
    pairs = [_pair("second1", "relaxed", 1e-12, seed=s, steps=2) for s in (5, 6)]
    bt = ptn.BatchedTDVP([a for a, _, _ in pairs])
    bt.SEQUENTIAL_N = 1000
    low0 = K.STATS.low_applies
    for _ in range(2):
        bt.run_one_time_step()
        for _, otd, _ in pairs:
            otd.run_one_time_step()
    assert K.STATS.low_applies > low0
    for alg, otd, _ in pairs:
        e, e_ref = ptn.expectation_value(alg.state, alg.hamiltonian), otd.energy()
        assert abs(e - e_ref) < 1e-9 * abs(e_ref)
    with pytest.raises(ValueError):
        ptn.BatchedTDVP([pairs[0][0], _pair("second1", "fp64", 0.0, steps=1)[0]])


def test_relaxed_bf16x6_level_matches_oracle(gpu):
    """relax_low="bf16x6": the FP32-accurate level runs as six products of
    three-piece bf16 splits; with relax_tol 1e-12 it stays within 1e-9 of the
    oracle, as the FP32 level does."""
    _say("relaxed precision, bf16x6 level")
    ids, parent, children = complete_tree(2, 3)
    tensors = random_state_tensors(ids, parent, children, 16, 5)
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = ptn.random_xyz_model(st, 5)
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision="relaxed",
                         relax_tol=1e-12, relax_min_size=0, relax_low="bf16x6")
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.15, [], config=cfg,
                                     solver_options={"krylov_tol": 1e-10, "max_krylov_dim": 10})
    otd = OracleTDVP(OracleState(topology_from_ttn(st), tensors), host_tensors(H), 0.05,
                     kind="second1", krylov_tol=1e-10, max_krylov_dim=10)
    low0 = K.STATS.low_applies
    errs = _run(alg, otd, 3)
    assert K.STATS.low_applies > low0
    assert max(errs) < 1e-9, errs


def test_config3_relaxed_20_steps_energy_error(gpu):
    """The chi=256 configuration (ternary tree N=40, random XYZ, dt=0.05) over its
    20 steps: the relaxed path at the benchmark's relax_tol, with and without the
    bf16 level, stays within |dE| <= 1e-8 of the FP64 path, and so does <Z_i> on
    every one of the 40 sites for the FP32 level (2e-7 with the bf16 level), while
    part of the applications ran in reduced precision."""
    _say("config C3 full size, 20 steps, fp64 vs relaxed")
    import bench
    from pytreenet_amd import engine as E
    res = {}
    for run in ("fp64", "relaxed", "relaxed_bf16"):
        prec = "fp64" if run == "fp64" else "relaxed"
        st, H, _ = bench.make_problem(256, torch.device("cuda"))
        cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision=prec,
                             relax_tol=bench.RELAX_TOL, relax_bf16=run == "relaxed_bf16")
        alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 1.0, [], config=cfg,
                                         solver_options={"max_krylov_dim": 10,
                                                         "krylov_tol": 1e-10})
        c = alg.state.orthogonality_center_id
        alg.state.tensors[c] = alg.state.tensors[c] / alg.state.norm()
        assert alg.num_time_steps == 20
        low0 = K.STATS.low_applies
        for step in range(20):
            alg.run_one_time_step()
        from pytreenet_amd.contractions import multi_single_site_expectation_value
        zs = multi_single_site_expectation_value(alg.state, {n: PAULI_Z for n in alg.state.nodes})
        res[run] = (ptn.expectation_value(alg.state, H).real,
                    [zs[n][0].real for n in sorted(zs)], K.STATS.low_applies - low0)
        _say(f"  {run}: E={res[run][0]:.15f} reduced-precision applications {res[run][2]}")
        del alg, st, H
        E.clear_caches()
        torch.cuda.empty_cache()
    e64, z64, _ = res["fp64"]
    # the bf16 level keeps the energy within the bar; its <Z_i> drift over all 40
    # sites reaches 1.06e-7 after the 20 steps (measured), so its bar there is 2e-7:
    # it misses the 1e-8 bar the FP32 level meets, and the bench does not take it
    for run, zbar in (("relaxed", 1e-8), ("relaxed_bf16", 2e-7)):
        er, zr, low = res[run]
        assert low > 0, run
        assert abs(er - e64) <= 1e-8, (run, er, e64)
        assert max(abs(a - b) for a, b in zip(zr, z64)) <= zbar, (run, zr, z64)


def test_relaxed_min_size_keeps_small_spaces_fp64(gpu):
    """Below relax_min_size every application runs in FP64: the run is the FP64
    path (1e-12 against the oracle) and no application dropped."""
    _say("relaxed precision with the default relax_min_size on a small tree")
    ids, parent, children = complete_tree(2, 3)
    tensors = random_state_tensors(ids, parent, children, 16, 5)
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = ptn.random_xyz_model(st, 5)
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision="relaxed",
                         relax_tol=1e-6)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                     solver_options={"krylov_tol": 1e-10, "max_krylov_dim": 10})
    otd = OracleTDVP(OracleState(topology_from_ttn(st), tensors), host_tensors(H), 0.05,
                     kind="second1", krylov_tol=1e-10, max_krylov_dim=10)
    low0 = K.STATS.low_applies
    errs = _run(alg, otd, 2)
    assert K.STATS.low_applies == low0
    assert max(errs) < 1e-12, errs


def test_relaxed_bf16_level(gpu):
    """With relax_bf16 the applications of the smallest weights run as bf16
    products; the run stays within 1e-9 of the oracle (relax_tol 1e-10)."""
    _say("relaxed precision with the bf16 level")
    ids, parent, children = complete_tree(2, 3)
    tensors = random_state_tensors(ids, parent, children, 16, 5)
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = ptn.random_xyz_model(st, 5)
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision="relaxed",
                         relax_tol=1e-10, relax_min_size=0, relax_bf16=True)
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.15, [], config=cfg,
                                     solver_options={"krylov_tol": 1e-10, "max_krylov_dim": 10})
    otd = OracleTDVP(OracleState(topology_from_ttn(st), tensors), host_tensors(H), 0.05,
                     kind="second1", krylov_tol=1e-10, max_krylov_dim=10)
    b0, l0 = K.STATS.lower_applies, K.STATS.low_applies
    errs = _run(alg, otd, 3)
    _say(f"  bf16 applications {K.STATS.lower_applies - b0}, reduced {K.STATS.low_applies - l0};"
         f" rel dE {max(errs):.2e}")
    assert K.STATS.lower_applies > b0
    assert max(errs) < 1e-9, errs
