"""CPU oracle for the PyTreeNet time-evolution hot path.

TEST INFRASTRUCTURE ONLY.  This package is a from-scratch NumPy restatement of
the reference algorithms (PyTreeNet, pure Python/NumPy) that the GPU library is
checked against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline -- never as a product code path.

Modules and the reference behaviour each one restates:

* ``tree``       -- tree topology helpers and the TDVP update/orthogonalisation
                    paths (``core/tree_structure.py``,
                    ``time_evolution/time_evo_util/update_path.py``).
* ``contract``   -- sandwich environments, H_eff / bond actions, expectation
                    values and overlaps in the reference contraction order
                    (``contractions/state_operator_contraction.py``,
                    ``contractions/contraction_util.py``,
                    ``contractions/state_state_contraction.py``).
* ``krylov``     -- the MGS2 Arnoldi exponential with the Saad stop
                    (``time_evolution/time_evolution.py:696-821``).
* ``splitting``  -- QR (KEEP/REDUCED) and truncated SVD with the reference
                    truncation rules (``util/tensor_splitting.py``).
* ``tdvp``       -- first/second-order one-site and second-order two-site TDVP
                    sweeps (``time_evolution/tdvp_algorithms/*.py``).
* ``dense``      -- brute-force state vectors and Hamiltonians for pinning.

Pinning (what makes each function trustworthy, checked in ``tests/``):

* tree paths: the hard-coded update/orth paths of the reference tests
  (``tests/golden/paths.json``).
* truncation: the deterministic singular-value KATs of the reference tests
  (``tests/golden/truncation.json``).
* krylov: ``scipy.linalg.expm`` on random Hermitian matrices; exactness at
  breakdown.
* contractions: brute-force dense vectors / Hamiltonians on small trees.
* tdvp: one-site TDVP at full (uncapped) bond dimensions equals exact
  state-vector evolution; energy and norm conservation; two-site TDVP without
  truncation equals exact evolution as well.
"""
