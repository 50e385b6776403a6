"""The oracle's DMRG sweeps, pinned by the exact ground state at full bond
dimensions and by the variational (non-increasing) energies of one-site sweeps."""
import numpy as np
import pytest

from oracle import dense as OD
from oracle.dmrg import OracleDMRG
from tests.helpers import oracle_problem


@pytest.mark.parametrize("kind", ["one", "two"])
def test_oracle_dmrg_reaches_exact_ground_state_at_full_bond_dims(kind):
    topo, st, W, H = oracle_problem(2, 2, 64, 9)
    e_exact, v_exact = np.linalg.eigh(H)
    d = OracleDMRG(st, W, dict(max_bond_dim=1000, rel_tol=1e-14, total_tol=1e-14))
    for _ in range(3):
        es = d.sweep_one_site() if kind == "one" else d.sweep_two_site()
    assert abs(es[-1] - e_exact[0]) < 1e-10
    psi = OD.state_to_vector(topo, d.s.A)
    assert abs(abs(np.vdot(v_exact[:, 0], psi)) - 1) < 1e-9


def test_oracle_one_site_dmrg_is_variational():
    """Below full rank every exact local solve lowers (or keeps) the energy, which
    stays above the exact ground energy."""
    topo, st, W, H = oracle_problem(2, 2, 2, 4)
    e0 = np.linalg.eigvalsh(H)[0]
    d = OracleDMRG(st, W)
    es = d.sweep_one_site() + d.sweep_one_site()
    assert all(b <= a + 1e-12 for a, b in zip(es, es[1:]))
    assert es[-1] >= e0 - 1e-12
    psi = OD.state_to_vector(topo, d.s.A)
    assert abs(np.vdot(psi, H @ psi).real - es[-1]) < 1e-11
