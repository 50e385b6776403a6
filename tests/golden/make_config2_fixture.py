"""Write the full-size reference series of the 63-node configuration.

Runs the oracle (NumPy restatement of the reference TDVP, reference contraction
order) on the 63-node binary tree at bond dimension 64 -- XXZ with Jxy = 1,
Jz = 0.8, second-order one-site TDVP, dt = 0.02, Krylov dimension <= 10 -- for
STEPS steps at krylov_tol 0 and 1e-10, and stores per step the energy, the
overlap <psi_0|psi_t> and <Z_i> for every site in ``config2_chi64.json``.

The state is ``random_state_tensors(complete_tree(2, 5), 64, seed=63)``; the
operator is built by the edge-term construction that the oracle pins check
against the dense Hamiltonian.  Run on the CPU (about 15 minutes on 8 cores):

    python tests/golden/make_config2_fixture.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import contract as OC                      # noqa: E402
from oracle.tdvp import OracleState, OracleTDVP        # noqa: E402
from tests.helpers import topology_from_spec           # noqa: E402

STEPS = 10
SEED = 63
CHI = 64
DT = 0.02
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


def problem():
    from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
    from pytreenet_amd.ttno import build_edge_ttno_tensors
    ids, parent, children = complete_tree(2, 5)
    tensors = random_state_tensors(ids, parent, children, CHI, SEED)
    topo = topology_from_spec(ids, parent, children)
    edges = []
    for n in ids:
        p = parent[n]
        if p is not None:
            edges += [(p, n, -1.0, X, X), (p, n, -1.0, Y, Y), (p, n, -0.8, Z, Z)]
    st = build_state(ids, parent, children, tensors, device="cpu")
    W = build_edge_ttno_tensors(st, {}, edges)
    return ids, topo, tensors, W


def series(topo, tensors, W, ids, tol):
    otd = OracleTDVP(OracleState(topo, tensors), W, DT, kind="second1", krylov_tol=tol,
                     max_krylov_dim=10)
    otd.s.normalise()
    psi0 = {k: v.copy() for k, v in otd.s.A.items()}
    out = {"energy": [], "overlap": [], "z": [], "krylov_dims": []}

    def record():
        e = otd.energy()
        ov = OC.overlap(topo, psi0, otd.s.A)              # <psi_0|psi_t>
        z = [otd.expect_site(n, Z).real for n in ids]
        out["energy"].append([e.real, e.imag])
        out["overlap"].append([ov.real, ov.imag])
        out["z"].append(z)

    record()
    for step in range(STEPS):
        t0 = time.time()
        n0 = len(otd.krylov_dims)
        otd.run_one_time_step()
        out["krylov_dims"].append(otd.krylov_dims[n0:])
        record()
        print(f"tol={tol:g} step {step + 1}: E={out['energy'][-1][0]:.15f} "
              f"({time.time() - t0:.1f} s)", flush=True)
    return out


def main():
    ids, topo, tensors, W = problem()
    data = {"config": {"tree": "binary depth 5 (63 nodes)", "chi": CHI, "seed": SEED,
                       "model": "XXZ Jxy=1 Jz=0.8, H = -sum(XX + YY + 0.8 ZZ)",
                       "algorithm": "second-order one-site TDVP", "dt": DT,
                       "max_krylov_dim": 10, "steps": STEPS, "sites": ids,
                       "normalised": True, "overlap": "<psi_0|psi_t>"},
            "series": {}}
    for tol in (0.0, 1e-10):
        data["series"][repr(tol)] = series(topo, tensors, W, ids, tol)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_chi64.json")
    with open(path, "w") as f:
        json.dump(data, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
