"""Matrix-free Krylov exponential (oracle, test infrastructure only).

Restates ``krylov_exp_action`` and ``_krylov_small_propagator``
(``time_evolution/time_evolution.py:696-821``):

* Arnoldi with full re-orthogonalisation by two passes of modified
  Gram-Schmidt, both passes accumulated into the same Hessenberg column
  (``:791-797``);
* happy breakdown when beta <= max(eps*n, 1e-12*||A q_j||) (``:778-809``);
* Saad a-posteriori stop ||psi|| * beta * |[exp(f dt T_m) e_1]_{m-1}| < krylov_tol
  (``:810-814``), disabled for krylov_tol <= 0;
* small propagator from eigh of (T + T^H)/2 when hermitian, else expm(f dt T)
  (``:696-718``).
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np
from scipy.linalg import expm


def small_propagator(t_mat: np.ndarray, dt: float, exp_factor: complex,
                     hermitian: bool = True) -> np.ndarray:
    if hermitian:
        herm = 0.5 * (t_mat + t_mat.conj().T)
        evals, vecs = np.linalg.eigh(herm)
        return vecs @ (np.exp(exp_factor * dt * evals) * vecs[0].conj())
    return expm(exp_factor * dt * t_mat)[:, 0]


def krylov_exp_action(psi: np.ndarray, action: Callable[[np.ndarray], np.ndarray],
                      dt: float, exp_factor: complex, krylov_tol: float = 1e-10,
                      max_krylov_dim: Optional[int] = None, hermitian: bool = True,
                      stats: Optional[dict] = None) -> np.ndarray:
    shape = psi.shape
    v0 = np.asarray(psi, dtype=complex).reshape(-1)
    n = v0.size
    beta0 = float(np.linalg.norm(v0))
    if beta0 == 0.0:
        return v0.copy().reshape(shape)
    max_steps = n if max_krylov_dim is None else min(n, int(max_krylov_dim))
    breakdown = np.finfo(float).eps * max(n, 1)
    q = [v0 / beta0]
    hess = np.zeros((max_steps + 1, max_steps + 1), dtype=complex)
    m = 0
    for j in range(max_steps):
        w = np.asarray(action(q[j].reshape(shape)), dtype=complex).reshape(-1)
        w0_norm = float(np.linalg.norm(w))
        for _ in range(2):
            for i in range(j + 1):
                c = np.vdot(q[i], w)
                hess[i, j] += c
                w = w - c * q[i]
        beta = float(np.linalg.norm(w))
        hess[j + 1, j] = beta
        m = j + 1
        if beta <= max(breakdown, 1e-12 * w0_norm):
            break
        if krylov_tol > 0.0:
            col = small_propagator(hess[:m, :m], dt, exp_factor, hermitian)
            if beta0 * beta * abs(col[m - 1]) < krylov_tol:
                break
        q.append(w / beta)
    coeffs = beta0 * small_propagator(hess[:m, :m], dt, exp_factor, hermitian)
    out = np.zeros(n, dtype=complex)
    for idx in range(m):
        out += coeffs[idx] * q[idx]
    if stats is not None:
        stats["m"] = m
    return out.reshape(shape)
