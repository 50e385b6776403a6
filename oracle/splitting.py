"""QR / SVD splitting and truncation rules (oracle, test infrastructure only).

Restates ``util/tensor_splitting.py``:

* ``qr_split``              -- ``tensor_qr_decomposition`` (``:60-107``): reduced QR of
  the matricised tensor; KEEP zero-pads Q's new columns and R's new rows up to
  the original bond dimension (``:98-106``).
* ``value_truncation``      -- ``:397-418``: keep s > max(rel_tol * s_0, total_tol).
* ``sum_truncation_index``  -- ``:354-395``.
* ``truncate_singular_values`` -- ``:420-486``: cap at max_bond_dim; keep s_0 when all
  are cut; optional 2-norm renormalisation (``:330-352``).
* ``svd_split``             -- ``truncated_tensor_svd`` + VCONTR
  (``:488-572``): singular values are absorbed into V.
"""
from __future__ import annotations

import warnings
from typing import Sequence, Tuple

import numpy as np


def matricise(t: np.ndarray, row_legs: Sequence[int], col_legs: Sequence[int]) -> np.ndarray:
    perm = list(row_legs) + list(col_legs)
    tp = np.transpose(t, perm)
    rows = int(np.prod([t.shape[i] for i in row_legs]))
    return tp.reshape(rows, -1)


def qr_split(t: np.ndarray, q_legs: Sequence[int], r_legs: Sequence[int],
             keep: bool = True) -> Tuple[np.ndarray, np.ndarray]:
    """Q legs (q_legs..., new), R legs (new, r_legs...)."""
    mat = matricise(t, q_legs, r_legs)
    q, r = np.linalg.qr(mat, mode="reduced")
    qshape = [t.shape[i] for i in q_legs]
    rshape = [t.shape[i] for i in r_legs]
    if keep:
        orig = int(np.prod(rshape))
        diff = orig - q.shape[1]
        if diff > 0:
            q = np.pad(q, [(0, 0), (0, diff)])
            r = np.pad(r, [(0, diff), (0, 0)])
    return q.reshape(qshape + [q.shape[1]]), r.reshape([r.shape[0]] + rshape)


def value_truncation(s: np.ndarray, total_tol: float, rel_tol: float) -> np.ndarray:
    cut = max(rel_tol * s[0], total_tol)
    return s[s > cut]


def sum_truncation_index(s: np.ndarray, total_tol: float, norming: bool = True) -> int:
    normsq = float(np.linalg.norm(s)) ** 2
    if normsq == 0:
        return 0
    thresh = total_tol ** 2
    acc = 0.0
    for i, sv in enumerate(reversed(s)):
        acc += sv ** 2
        comp = acc / normsq if norming else acc
        if comp > thresh:
            return len(s) - i
    return 0


def truncate_singular_values(s: np.ndarray, max_bond_dim=100, rel_tol=1e-15,
                             total_tol=1e-15, renorm=False, sum_trunc=False,
                             sum_renorm=False) -> Tuple[np.ndarray, np.ndarray]:
    if len(s) == 0:
        raise ValueError("No singular values to truncate!")
    if sum_trunc:
        kept = s[:sum_truncation_index(s, total_tol, sum_renorm)]
    else:
        kept = value_truncation(s, total_tol, rel_tol)
    if len(kept) > max_bond_dim:
        new_s, s_trunc = kept[:max_bond_dim], s[max_bond_dim:]
    elif len(kept) == 0:
        warnings.warn("All singular values were truncated. Returning only the "
                      "largest singular value.")
        new_s, s_trunc = np.array([s[0]]), s[1:]
    else:
        new_s, s_trunc = kept, s[len(kept):]
    if renorm:
        new_s = new_s * np.linalg.norm(s) / np.linalg.norm(new_s)
    return new_s, s_trunc


def svd_split(t: np.ndarray, u_legs: Sequence[int], v_legs: Sequence[int],
              **params) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Returns (U (u_legs..., new), S V (new, v_legs...), s_kept)."""
    mat = matricise(t, u_legs, v_legs)
    u, s, vh = np.linalg.svd(mat, full_matrices=False)
    s_kept, _ = truncate_singular_values(s, **params)
    k = len(s_kept)
    u = u[:, :k]
    svh = s_kept[:, None] * vh[:k, :]
    ushape = [t.shape[i] for i in u_legs]
    vshape = [t.shape[i] for i in v_legs]
    return u.reshape(ushape + [k]), svh.reshape([k] + vshape), s_kept
