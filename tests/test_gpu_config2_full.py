"""The 63-node configuration at full size against the oracle's committed series.

Binary tree of depth 5 (63 nodes), bond dimension 64, XXZ (Jxy = 1, Jz = 0.8),
second-order one-site TDVP with dt = 0.02 and Krylov dimension <= 10, from the
seeded random state.  ``golden/config2_chi64.json`` holds the oracle's energy,
overlap <psi_0|psi_t> and <Z_i> of every site after each of its steps, at
krylov_tol 0 (fixed Krylov dimension) and 1e-10 (``golden/make_config2_fixture.py``
wrote it on the CPU).  Bar: 1e-12 relative on the energy and the overlap, 1e-12
absolute on <Z_i> (the state is normalised); the Krylov dimension of every solve
equals the oracle's."""
import json
import os

import numpy as np
import pytest

import pytreenet_amd as ptn
from pytreenet_amd.contractions import multi_single_site_expectation_value
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import PAULI_Z

pytestmark = pytest.mark.gpu
FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "config2_chi64.json")
REL = 1e-12


def _load():
    if not os.path.exists(FIXTURE):
        pytest.skip("config2_chi64.json not generated")
    with open(FIXTURE) as f:
        return json.load(f)


@pytest.mark.parametrize("tol", ["0.0", "1e-10"])
def test_config2_full_size_matches_oracle_series(gpu, tol):
    data = _load()
    cfg, ref = data["config"], data["series"][tol]
    print(f"config C2 chi 64, krylov_tol {tol}, {cfg['steps']} steps vs oracle", flush=True)
    ids, parent, children = complete_tree(2, 5)
    tensors = random_state_tensors(ids, parent, children, cfg["chi"], cfg["seed"])
    st = build_state(ids, parent, children, tensors, device="cuda")
    H = ptn.xxz_model(st, 1.0, 0.8)
    alg = ptn.SecondOrderOneSiteTDVP(
        st, H, cfg["dt"], cfg["dt"] * cfg["steps"], [],
        config=ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV),
        solver_options={"max_krylov_dim": 10, "krylov_tol": float(tol)})
    c = alg.state.orthogonality_center_id
    alg.state.tensors[c] = alg.state.tensors[c] / alg.state.norm()
    psi0 = alg.state.copy()
    first_miss = None
    worst = 0.0
    dims_differ = []
    for step in range(cfg["steps"] + 1):
        if step:
            nlog = len(alg.apply_log)
            alg.run_one_time_step()
            # Krylov dimension of every site and bond solve, in sweep order
            dims = [a for _, _, a in alg.apply_log[nlog:]]
            if dims != ref["krylov_dims"][step - 1]:
                dims_differ.append(step)
        e = ptn.expectation_value(alg.state, H)
        ov = ptn.contract_two_ttns(alg.state, psi0, conj_second=True)    # <psi_0|psi_t>
        z = multi_single_site_expectation_value(alg.state, {n: PAULI_Z for n in ids})
        e_ref = complex(*ref["energy"][step])
        ov_ref = complex(*ref["overlap"][step])
        dz = max(abs(z[n][0].real - ref["z"][step][i]) for i, n in enumerate(ids))
        de = abs(e - e_ref) / abs(e_ref)
        dov = abs(ov - ov_ref) / abs(ov_ref)
        worst = max(worst, de, dov, dz)
        print(f"  step {step}: dE/E={de:.1e} dov={dov:.1e} max dZ={dz:.1e}", flush=True)
        if first_miss is None and max(de, dov, dz) >= REL:
            first_miss = step
    assert first_miss is None, f"first step beyond 1e-12: {first_miss} (worst {worst:.1e})"
    assert not dims_differ, f"Krylov dimensions differ from the oracle's at steps {dims_differ}"
