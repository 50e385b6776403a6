"""QR / SVD node splits with the reference truncation rules.

Drop-in counterparts of ``util/tensor_splitting.py``: ``SplitMode`` (``:23-58``),
``tensor_qr_decomposition`` (``:60-107``), ``tensor_svd`` (``:122-177``),
``SVDParameters`` (``:253-328``), ``renormalise_singular_values`` (``:330-352``),
``sum_truncation`` (``:354-395``), ``value_truncation`` (``:397-418``),
``truncate_singular_values`` (``:420-486``), ``truncated_tensor_svd``
(``:488-527``) and ``contr_truncated_svd_splitting`` (``:541-572``).

Device side: the matricisation is one strided copy into a column-major panel
(``ptn_permute``), the factorisation runs in FP64 on the device
(``ptn_qr`` / ``ptn_svd``), and Q, U, S*Vh are written straight into their
final leg layout by strided copies or one gather-GEMM.  Only the singular
values travel to the host, where the truncation rules (tiny vectors, exact
reference semantics) pick the kept dimension.
"""
from __future__ import annotations

from contextlib import contextmanager

import ctypes
import warnings
from dataclasses import dataclass
from enum import Enum
from numbers import Integral
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import engine as E


class SplitMode(Enum):
    FULL = "full"
    REDUCED = "reduced"
    KEEP = "keep"


class ContractionMode(Enum):
    UCONTR = "ucontr"
    VCONTR = "vcontr"
    EQUAL = "equal"


@dataclass
class SVDParameters:
    max_bond_dim: int = 100
    rel_tol: float = 1e-15
    total_tol: float = 1e-15
    renorm: bool = False
    sum_trunc: bool = False
    sum_renorm: bool = False
    random: bool = False

    def __post_init__(self):
        self.check_truncation_parameters()

    def check_truncation_parameters(self):
        mbd = self.max_bond_dim
        if not isinstance(mbd, Integral) and mbd != float("inf"):
            raise TypeError(f"'max_bond_dim' has to be an integer or inf, not {type(mbd)}!")
        if mbd <= 0:
            raise ValueError("'max_bond_dim' has to be positive.")
        if self.rel_tol < 0 and self.rel_tol != float("-inf"):
            raise ValueError("'rel_tol' has to be positive or -inf.")
        if self.total_tol < 0 and self.total_tol != float("-inf"):
            raise ValueError("'total_tol' has to be positive or -inf.")
        if self.random:
            raise ValueError("randomised SVD is not part of the device path")


# ---------------------------------------------------------------------------------------
# truncation rules (host, on the singular-value vector)
# ---------------------------------------------------------------------------------------
def renormalise_singular_values(s: np.ndarray, new_s: np.ndarray) -> np.ndarray:
    return new_s * np.linalg.norm(s) / np.linalg.norm(new_s)


def _sum_truncation_index(s: np.ndarray, total_tol: float, norming: bool = True) -> int:
    normsq = float(np.linalg.norm(s)) ** 2
    if normsq == 0:
        return 0
    thresh = total_tol ** 2
    acc = 0.0
    for i, sv in enumerate(reversed(s)):
        acc += float(sv) ** 2
        if (acc / normsq if norming else acc) > thresh:
            return len(s) - i
    return 0


def sum_truncation(s: np.ndarray, total_tol: float, norming: bool = True) -> np.ndarray:
    return s[: _sum_truncation_index(s, total_tol, norming=norming)]


def value_truncation(s: np.ndarray, total_tol: float, rel_tol: float) -> np.ndarray:
    return s[s > max(rel_tol * s[0], total_tol)]


def truncate_singular_values(s: np.ndarray, svd_params) -> Tuple[np.ndarray, np.ndarray]:
    if len(s) == 0:
        raise ValueError("No singular values to truncate!")
    if getattr(svd_params, "sum_trunc", False):
        kept = sum_truncation(s, svd_params.total_tol, getattr(svd_params, "sum_renorm", False))
    else:
        kept = value_truncation(s, svd_params.total_tol, svd_params.rel_tol)
    mbd = svd_params.max_bond_dim
    if len(kept) > mbd:
        new_s, s_trunc = kept[: int(mbd)], s[int(mbd):]
    elif len(kept) == 0:
        warnings.warn("All singular values were truncated. Returning only the largest "
                      "singular value.")
        new_s, s_trunc = np.array([s[0]]), s[1:]
    else:
        new_s, s_trunc = kept, s[len(kept):]
    if getattr(svd_params, "renorm", False):
        new_s = renormalise_singular_values(s, new_s)
    return new_s, s_trunc


# ---------------------------------------------------------------------------------------
# device factorisations
# ---------------------------------------------------------------------------------------
def _cstrides(dims: Sequence[int]) -> List[int]:
    out, acc = [], 1
    for d in reversed(list(dims)):
        out.append(acc)
        acc *= d
    return list(reversed(out))


def _colmajor_panel(t: torch.Tensor, row_legs: Sequence[int], col_legs: Sequence[int]):
    """Column-major (m x n) copy: row-major (n, m) = t permuted to (col_legs, row_legs)."""
    return E.permute_copy(t, list(col_legs) + list(row_legs), dtype=torch.complex128)


def qr_colmajor(panel: torch.Tensor, m: int, n: int) -> Tuple[torch.Tensor, torch.Tensor, int]:
    """In-place reduced QR of a column-major m x n panel. Returns (panel, R_cm, k)."""
    k = min(m, n)
    R = torch.empty(k * n, dtype=torch.complex128, device=panel.device)
    tau = torch.empty(k, dtype=torch.complex128, device=panel.device)
    _lib.check(_lib.load().ptn_qr(ctypes.c_void_p(E.stream_handle()), m, n,
                                  ctypes.c_void_p(panel.data_ptr()),
                                  ctypes.c_void_p(R.data_ptr()),
                                  ctypes.c_void_p(tau.data_ptr())), "ptn_qr")
    return panel, R, k


def tensor_qr_decomposition(tensor: torch.Tensor, q_legs: Sequence[int], r_legs: Sequence[int],
                            mode: SplitMode = SplitMode.REDUCED
                            ) -> Tuple[torch.Tensor, torch.Tensor]:
    """Q legs (q_legs..., new), R legs (new, r_legs...); KEEP zero-pads to the R dims."""
    E.require_device(tensor)
    if mode is SplitMode.FULL:
        raise NotImplementedError("SplitMode.FULL is not part of the device path")
    q_legs, r_legs = list(q_legs), list(r_legs)
    qdims = [tensor.shape[i] for i in q_legs]
    rdims = [tensor.shape[i] for i in r_legs]
    m, n = int(np.prod(qdims)), int(np.prod(rdims))
    panel = _colmajor_panel(tensor, q_legs, r_legs)
    panel, R, k = qr_colmajor(panel, m, n)
    kk = n if (mode is SplitMode.KEEP and k < n) else k
    Q = torch.empty(qdims + [kk], dtype=torch.complex128, device=tensor.device)
    if kk > k:
        E.fill_(Q, 0.0)
    # panel rows 0..k-1 of the row-major (n, m) buffer are Q's columns
    E.strided_copy([k] + qdims, panel, [m] + _cstrides(qdims), Q,
                   [Q.stride(-1)] + list(Q.stride()[:-1]))
    Rt = torch.empty([kk] + rdims, dtype=torch.complex128, device=tensor.device)
    if kk > k:
        E.fill_(Rt, 0.0)
    # R_cm element (i, j) at j*k + i
    E.strided_copy([k] + rdims, R, [1] + [k * s for s in _cstrides(rdims)], Rt,
                   list(Rt.stride()))
    return Q, Rt


_U = 2.0 ** -53


def _potrf(G: torch.Tensor, shift_factor: float, min_pivot: float, status: torch.Tensor):
    n = G.shape[0]
    L = torch.empty((n, n), dtype=torch.complex128, device=G.device)
    _lib.check(_lib.load().ptn_potrf_lower(ctypes.c_void_p(E.stream_handle()), n,
                                           ctypes.c_void_p(G.data_ptr()), G.stride(0),
                                           ctypes.c_void_p(L.data_ptr()), shift_factor,
                                           min_pivot, ctypes.c_void_p(status.data_ptr())),
               "ptn_potrf_lower")
    X = torch.empty((n, n), dtype=torch.complex128, device=G.device)
    _lib.check(_lib.load().ptn_trtri_lower(ctypes.c_void_p(E.stream_handle()), n,
                                           ctypes.c_void_p(L.data_ptr()),
                                           ctypes.c_void_p(X.data_ptr())), "ptn_trtri_lower")
    return L, X


def cholesky_qr3_leg(t: torch.Tensor, leg: int):
    """Shifted CholeskyQR3 of ``t`` matricised as (all other legs) x (leg ``leg``).

    Returns (Q with the new bond at position ``leg``, R (new, old), ok).  Each pass
    is a Gram gather-GEMM G = A^H A, a device Cholesky L L^H = G (+ shift in the
    first pass, s = 11 (mn + n(n+1)) u tr(G)), and the update A <- A L^{-H} as a
    gather-GEMM straight in the tensor layout.  ``ok`` is False when a pivot
    collapsed (numerically rank-deficient panel); the caller then falls back to
    Householder QR.
    """
    n = t.shape[leg]
    m = t.numel() // n
    lab = list(range(t.dim()))
    la = list(lab)
    la[leg] = "c"
    lb = list(lab)
    lb[leg] = "c2"
    status = torch.zeros(3, dtype=torch.int32, device=t.device)
    q = t
    Ls = []
    for p in range(3):
        # (the Cholesky reads the lower triangle only)
        G = E.contract(q, la, q, lb, ["c", "c2"], conj_a=True, lower=True)
        shift = 11.0 * (m * n + n * (n + 1)) * _U if p == 0 else 0.0
        L, X = _potrf(G, shift, 0.0 if p == 0 else 1e-14, status[p:p + 1])
        Ls.append(L)
        lq = list(lab)
        lq[leg] = "k"
        q = E.contract(q, lq, X, ["c", "k"], la, conj_b=True)
    P = E.contract(Ls[0], ["a", "b"], Ls[1], ["b", "c"], ["a", "c"])
    P = E.contract(P, ["a", "c"], Ls[2], ["c", "d"], ["a", "d"])
    R = E.permute_copy(P, [1, 0], conj=True)
    if _DEFER.active:
        # checked once at the end of the time step (deferred_qr_checks)
        _DEFER.flags.append(status)
        return q, R, True
    ok = int(status.sum().cpu().item()) == 0
    return q, R, ok


class _DeferredChecks:
    """Pivot statuses of the CholeskyQR3 splits of one time step, read back once."""

    def __init__(self):
        self.active = False
        self.flags: List[torch.Tensor] = []

    def failed(self) -> bool:
        if not self.flags:
            return False
        return bool(torch.cat(self.flags).sum().item() != 0)


_DEFER = _DeferredChecks()


@contextmanager
def deferred_qr_checks():
    """Within the block, a CholeskyQR3 split returns its Q and R without reading
    its pivot status back (which would stall the device queue); the statuses
    are collected and ``failed()`` reads them all with one copy.  The caller
    must redo the work of the block with the checks on when a split failed."""
    prev = (_DEFER.active, _DEFER.flags)
    _DEFER.active, _DEFER.flags = True, []
    try:
        yield _DEFER
    finally:
        _DEFER.active, _DEFER.flags = prev


QR_STATS = {"cholqr3": 0, "householder": 0, "small": 0}
# tensors up to this many elements are split by the one-launch Householder kernel
QR_SMALL_MAX = 3072


def qr_split_small(t: torch.Tensor, leg: int) -> Tuple[torch.Tensor, torch.Tensor]:
    """QR(KEEP) of ``t`` matricised as (other legs) x (leg ``leg``) in one launch
    (``ptn_qr_split_small``): Q with t's shape (new bond at ``leg``), R (new, old);
    diag(R) real and non-negative."""
    E.require_device(t)
    t = t if t.dtype == torch.complex128 else t.to(torch.complex128)
    n = t.shape[leg]
    q = torch.empty(t.shape, dtype=torch.complex128, device=t.device)
    r = torch.empty((n, n), dtype=torch.complex128, device=t.device)
    arr = ctypes.c_int64 * t.dim()
    _lib.check(_lib.load().ptn_qr_split_small(
        ctypes.c_void_p(E.stream_handle()), t.dim(), arr(*t.shape), arr(*t.stride()), leg,
        ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(q.data_ptr()),
        ctypes.c_void_p(r.data_ptr())), "ptn_qr_split_small")
    return q, r


def qr_split_to_neighbour(ttn, node_id: str, neighbour_id: str, mode=SplitMode.KEEP):
    """QR of a node off its leg to ``neighbour_id``.

    Returns (Q in the node's canonical layout with the new bond at the old
    leg's position, R with legs (new, old)).  Small tensors run the one-launch
    Householder kernel; larger panels with rows >= columns use shifted
    CholeskyQR3, wide or rank-deficient ones Householder QR.  The modes differ
    only for a bond wider than the product m of the node's other dimensions:
    KEEP (also for ``mode=None``) zero-pads Q and R to the old bond, REDUCED
    shrinks the bond to m.
    """
    t = ttn.tensors[node_id]
    leg = ttn.nodes[node_id].neighbour_index(neighbour_id)
    n = t.shape[leg]
    m = t.numel() // n
    if mode is SplitMode.REDUCED and m < n:
        others = [i for i in range(t.dim()) if i != leg]
        q, r = tensor_qr_decomposition(t, others, [leg], mode=SplitMode.REDUCED)
        perm = list(range(q.dim() - 1))
        perm.insert(leg, q.dim() - 1)
        return E.permute_copy(q, perm), r
    if t.numel() <= QR_SMALL_MAX and n <= 256 and t.dim() <= 8:
        QR_STATS["small"] += 1
        return qr_split_small(t, leg)
    if m >= n and n <= 1024:
        q, r, ok = cholesky_qr3_leg(t, leg)
        if ok:
            QR_STATS["cholqr3"] += 1
            return q, r
    QR_STATS["householder"] += 1
    others = [i for i in range(t.dim()) if i != leg]
    q, r = tensor_qr_decomposition(t, others, [leg], mode=SplitMode.KEEP)
    perm = list(range(q.dim() - 1))
    perm.insert(leg, q.dim() - 1)
    return E.permute_copy(q, perm), r


def absorb_into_neighbour(ttn, node_id: str, from_id: str, r: torch.Tensor):
    """node[.., new, ..] = sum_old r[new, old] node[.., old, ..] on the leg to ``from_id``."""
    t = ttn.tensors[node_id]
    leg = ttn.nodes[node_id].neighbour_index(from_id)
    lt = list(range(t.dim()))
    lt[leg] = "old"
    lo = list(range(t.dim()))
    lo[leg] = "new"
    ttn.tensors[node_id] = E.contract(r, ["new", "old"], t, lt, lo)


SVD_STATS = {"direct": 0, "qr_tall": 0, "qr_wide": 0, "qr_fallback": 0, "jacobi": 0,
             "jacobi_fallback": 0}
# aspect ratio from which a panel is reduced by QR before the SVD
SVD_QR_RATIO = 4


class SvdRequest:
    """One pending thin SVD of a column-major panel (the small dense step of a
    truncated split): answered by ``answer_svd`` one at a time, or together with
    other instances' requests by ``answer_svd_batched``."""
    __slots__ = ("panel", "m", "n", "on_rows", "result")

    def __init__(self, panel: torch.Tensor, m: int, n: int, on_rows: Optional[int] = None):
        self.panel, self.m, self.n, self.on_rows = panel, m, n, on_rows
        self.result = None


def answer_svd(req: SvdRequest):
    req.result = _svd_direct(req.panel, req.m, req.n, req.on_rows)


def drive_svd(gen):
    """Run a generator of SvdRequests, answering each one at a time."""
    try:
        req = next(gen)
        while True:
            answer_svd(req)
            req = gen.send(None)
    except StopIteration as stop:
        return stop.value


def svd_colmajor(panel: torch.Tensor, m: int, n: int):
    """Thin SVD of a column-major m x n panel (destroyed): (s host, U_cm, Vh_cm, k).

    Strongly rectangular panels (long side >= SVD_QR_RATIO x short side, short
    side 8..1024) are first reduced by a CholeskyQR3 of the long side, A = Q R
    (tall) or A^T = Q R (wide), and only the small triangle goes through the
    SVD: A = (Q U_R) S V_R^H, or A = U_B S (W^H Q^T) with R^T = U_B S W^H.
    Both steps are backward stable, so the singular values equal those of the
    direct SVD to O(u ||A||).  A collapsed CholeskyQR pivot (numerically rank
    deficient panel) falls back to the direct SVD.
    """
    return drive_svd(svd_colmajor_gen(panel, m, n))


def svd_colmajor_gen(panel: torch.Tensor, m: int, n: int):
    """``svd_colmajor`` as a generator: the dense SVD step is handed out as an
    ``SvdRequest`` (``result = yield request``)."""
    short, long_ = min(m, n), max(m, n)
    if 8 <= short <= 1024 and long_ >= SVD_QR_RATIO * short:
        hit = yield from _svd_via_qr_gen(panel, m, n)
        if hit is not None:
            return hit
        SVD_STATS["qr_fallback"] += 1
    SVD_STATS["direct"] += 1
    req = SvdRequest(panel, m, n)
    yield req
    return req.result


def _svd_via_qr_gen(panel: torch.Tensor, m: int, n: int):
    dev = panel.device
    T = panel.view(n, m)                 # T[j][i] = A(i, j)
    if m >= n:
        # A = Q R: matricise T as (i) x (j), the QR leg is j (leg 0)
        q, R, ok = cholesky_qr3_leg(T, 0)  # q[r][i] = Q(i, r), R[r][j]
        if not ok:
            return None
        B = E.permute_copy(R, [1, 0])      # B col-major n x n = R
        # Jacobi on the rows of the triangle (columns of R^H): QR-preconditioned
        req = SvdRequest(B.reshape(-1), n, n, on_rows=1)
        yield req
        s, Ub, Vhb, k = req.result
        U = torch.empty(m * k, dtype=torch.complex128, device=dev)
        E.contract(q, ["r", "i"], torch.as_strided(Ub, (n, k), (1, n)), ["r", "c"], ["i", "c"],
                   out=torch.as_strided(U, (m, k), (1, m)))
        SVD_STATS["qr_tall"] += 1
        return s, U, Vhb, k
    # A^T = Q R: matricise T as (j) x (i), the QR leg is i (leg 1)
    q, R, ok = cholesky_qr3_leg(T, 1)      # q[j][r] = Q(j, r), R[r][i]
    if not ok:
        return None
    B = R.contiguous().clone()             # col-major m x m of R^T: B(i, r) = R[r][i]
    # Jacobi on the rows of R^T (columns of conj(R)): QR-preconditioned
    req = SvdRequest(B.reshape(-1), m, m, on_rows=1)
    yield req
    s, Ub, Whb, k = req.result
    Vh = torch.empty(k * n, dtype=torch.complex128, device=dev)
    E.contract(torch.as_strided(Whb, (k, m), (1, k)), ["c", "r"], q, ["j", "r"], ["c", "j"],
               out=torch.as_strided(Vh, (k, n), (1, k)))
    SVD_STATS["qr_wide"] += 1
    return s, Ub, Vh, k


# panels whose short side is at most SVD_JACOBI_MAX_SHORT (and long side at most
# SVD_JACOBI_MAX_LONG) go to the one-workgroup Jacobi SVD (ptn_svd_jacobi); larger
# ones, and the rare panel on which Jacobi does not converge in 64 sweeps, to
# rocSOLVER's bidiagonal SVD (ptn_svd)
SVD_JACOBI_MAX_SHORT = 256
SVD_JACOBI_MAX_LONG = 4096


def _svd_jacobi(panel: torch.Tensor, m: int, n: int, on_rows: int):
    k = min(m, n)
    dev = panel.device
    lib = _lib.load()
    S = torch.empty(k, dtype=torch.float64, device=dev)
    U = torch.empty(m * k, dtype=torch.complex128, device=dev)
    Vh = torch.empty(k * n, dtype=torch.complex128, device=dev)
    ws = torch.empty(int(lib.ptn_svd_jacobi_work_bytes(m, n)) // 16, dtype=torch.complex128,
                     device=dev)
    info = torch.zeros(2, dtype=torch.float64, device=dev)   # int32 info in the first word
    _lib.check(lib.ptn_svd_jacobi(ctypes.c_void_p(E.stream_handle()), m, n,
                                  ctypes.c_void_p(panel.data_ptr()), m, on_rows,
                                  ctypes.c_void_p(S.data_ptr()), ctypes.c_void_p(U.data_ptr()),
                                  ctypes.c_void_p(Vh.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                                  ctypes.c_void_p(info.data_ptr())), "ptn_svd_jacobi")
    host = torch.cat([S, info]).cpu().numpy()
    if int(np.frombuffer(host[k:k + 1].tobytes(), dtype=np.int32)[0]) != 0:
        return None
    SVD_STATS["jacobi"] += 1
    return host[:k].copy(), U, Vh, k


def _jacobi_ok(m: int, n: int) -> bool:
    return min(m, n) <= SVD_JACOBI_MAX_SHORT and max(m, n) <= SVD_JACOBI_MAX_LONG


def answer_svd_batched(reqs: Sequence[SvdRequest]):
    """Answer many independent SVD requests: those within the Jacobi limits in one
    ``ptn_svd_jacobi_batched`` launch (one workgroup per matrix) and one host
    read-back; the others, and any that did not converge, one at a time."""
    todo = [r for r in reqs if _jacobi_ok(r.m, r.n)]
    for r in reqs:
        if not _jacobi_ok(r.m, r.n):
            answer_svd(r)
    if not todo:
        return
    if len(todo) == 1:
        answer_svd(todo[0])
        return
    lib = _lib.load()
    dev = todo[0].panel.device
    ks = [min(r.m, r.n) for r in todo]
    outs = []
    ptrs = {k: [] for k in ("A", "S", "U", "Vh", "ws")}
    meta = {"m": [], "n": [], "lda": [], "rows": []}
    kt = sum(ks)
    S_all = torch.empty(kt + 2 * len(todo), dtype=torch.float64, device=dev)
    info = S_all[kt:].view(torch.int32)           # one int32 per request, after the S blocks
    off = 0
    for r, k in zip(todo, ks):
        U = torch.empty(r.m * k, dtype=torch.complex128, device=dev)
        Vh = torch.empty(k * r.n, dtype=torch.complex128, device=dev)
        ws = torch.empty(int(lib.ptn_svd_jacobi_work_bytes(r.m, r.n)) // 16,
                         dtype=torch.complex128, device=dev)
        outs.append((U, Vh, ws, off))
        ptrs["A"].append(r.panel.data_ptr())
        ptrs["S"].append(S_all[off:].data_ptr())
        ptrs["U"].append(U.data_ptr())
        ptrs["Vh"].append(Vh.data_ptr())
        ptrs["ws"].append(ws.data_ptr())
        meta["m"].append(r.m)
        meta["n"].append(r.n)
        meta["lda"].append(r.m)
        meta["rows"].append(int(r.m < r.n) if r.on_rows is None else int(r.on_rows))
        off += k
    info_p = [info[i:].data_ptr() for i in range(len(todo))]
    arr = {k: np.ascontiguousarray(np.array(v, dtype=np.int64)) for k, v in ptrs.items()}
    m_a = np.ascontiguousarray(np.array(meta["m"], dtype=np.int64))
    n_a = np.ascontiguousarray(np.array(meta["n"], dtype=np.int64))
    ld_a = np.ascontiguousarray(np.array(meta["lda"], dtype=np.int64))
    rw_a = np.ascontiguousarray(np.array(meta["rows"], dtype=np.int32))
    in_a = np.ascontiguousarray(np.array(info_p, dtype=np.int64))
    work = torch.empty(int(lib.ptn_svd_jacobi_batched_work_bytes(len(todo))), dtype=torch.uint8,
                       device=dev)
    cp = lambda a: ctypes.c_void_p(a.ctypes.data)
    _lib.check(lib.ptn_svd_jacobi_batched(ctypes.c_void_p(E.stream_handle()), len(todo),
                                          cp(m_a), cp(n_a), cp(arr["A"]), cp(ld_a), cp(rw_a),
                                          cp(arr["S"]), cp(arr["U"]), cp(arr["Vh"]),
                                          cp(arr["ws"]), cp(in_a),
                                          ctypes.c_void_p(work.data_ptr())),
               "ptn_svd_jacobi_batched")
    host = S_all.cpu().numpy()
    infos = np.frombuffer(host[kt:].tobytes(), dtype=np.int32)
    for i, (r, k, (U, Vh, _ws, o)) in enumerate(zip(todo, ks, outs)):
        if infos[i] != 0:
            SVD_STATS["jacobi_fallback"] += 1
            answer_svd(r)
            continue
        SVD_STATS["jacobi"] += 1
        r.result = (host[o:o + k].copy(), U, Vh, k)


def _svd_direct(panel: torch.Tensor, m: int, n: int, on_rows: Optional[int] = None):
    """Thin SVD of the column-major panel: one-sided Jacobi (on the columns of the
    panel when m >= n, else of its conjugate transpose, or as ``on_rows`` says)
    within the Jacobi limits, rocSOLVER beyond them."""
    k = min(m, n)
    if k <= SVD_JACOBI_MAX_SHORT and max(m, n) <= SVD_JACOBI_MAX_LONG:
        rows = int(m < n) if on_rows is None else on_rows
        hit = _svd_jacobi(panel, m, n, rows)
        if hit is not None:
            return hit
        SVD_STATS["jacobi_fallback"] += 1
    dev = panel.device
    S = torch.empty(k, dtype=torch.float64, device=dev)
    U = torch.empty(m * k, dtype=torch.complex128, device=dev)
    Vh = torch.empty(k * n, dtype=torch.complex128, device=dev)
    Ew = torch.empty(k + 2, dtype=torch.float64, device=dev)
    _lib.check(_lib.load().ptn_svd(ctypes.c_void_p(E.stream_handle()), m, n,
                                   ctypes.c_void_p(panel.data_ptr()),
                                   ctypes.c_void_p(S.data_ptr()),
                                   ctypes.c_void_p(U.data_ptr()),
                                   ctypes.c_void_p(Vh.data_ptr()),
                                   ctypes.c_void_p(Ew.data_ptr())), "ptn_svd")
    tail = Ew[k:].cpu()
    info = int(np.frombuffer(tail.numpy().tobytes(), dtype=np.int32)[0])
    if info != 0:
        raise np.linalg.LinAlgError(f"device SVD did not converge (info={info})")
    return S.cpu().numpy(), U, Vh, k


def tensor_svd(tensor: torch.Tensor, u_legs: Sequence[int], v_legs: Sequence[int]):
    """(U (u_legs..., k), s (host), Vh (k, v_legs...)) -- thin SVD."""
    E.require_device(tensor)
    u_legs, v_legs = list(u_legs), list(v_legs)
    udims = [tensor.shape[i] for i in u_legs]
    vdims = [tensor.shape[i] for i in v_legs]
    m, n = int(np.prod(udims)), int(np.prod(vdims))
    panel = _colmajor_panel(tensor, u_legs, v_legs)
    s, U, Vh, k = svd_colmajor(panel, m, n)
    Ut = torch.empty(udims + [k], dtype=torch.complex128, device=tensor.device)
    E.strided_copy([k] + udims, U, [m] + _cstrides(udims), Ut,
                   [Ut.stride(-1)] + list(Ut.stride()[:-1]))
    Vt = torch.empty([k] + vdims, dtype=torch.complex128, device=tensor.device)
    E.strided_copy([k] + vdims, Vh, [1] + [k * s_ for s_ in _cstrides(vdims)], Vt,
                   list(Vt.stride()))
    return Ut, s, Vt


def truncated_tensor_svd(tensor: torch.Tensor, u_legs: Sequence[int], v_legs: Sequence[int],
                         svd_params: SVDParameters):
    u, s, vh = tensor_svd(tensor, u_legs, v_legs)
    s_new, _ = truncate_singular_values(s, svd_params)
    r = len(s_new)
    return u[..., :r].contiguous(), np.asarray(s_new), vh[:r].contiguous()


def contr_truncated_svd_splitting(tensor: torch.Tensor, u_legs: Sequence[int],
                                  v_legs: Sequence[int],
                                  contr_mode: ContractionMode = ContractionMode.VCONTR,
                                  svd_params: SVDParameters = SVDParameters()):
    u, s, vh = truncated_tensor_svd(tensor, u_legs, v_legs, svd_params)
    dev = tensor.device
    if contr_mode == ContractionMode.VCONTR:
        return u, scale_leg(vh, 0, s)
    if contr_mode == ContractionMode.UCONTR:
        return scale_leg(u, u.dim() - 1, s), vh
    root = np.sqrt(s)
    return scale_leg(u, u.dim() - 1, root), scale_leg(vh, 0, root)


def scale_leg(t: torch.Tensor, leg: int, s: np.ndarray) -> torch.Tensor:
    """t with leg ``leg`` scaled by the vector s (diag(s) contracted on that leg)."""
    d = torch.as_tensor(np.diag(np.asarray(s, dtype=np.complex128))).to(t.device)
    lt = list(range(t.dim()))
    lt[leg] = "old"
    lo = list(range(t.dim()))
    lo[leg] = "new"
    return E.contract(d, ["new", "old"], t, lt, lo)


# ---------------------------------------------------------------------------------------
# two-site merge and truncated split (two-site TDVP, TEBD)
# ---------------------------------------------------------------------------------------
def merge_two_sites(ttn, n1: str, n2: str) -> torch.Tensor:
    """The tensors of neighbours n1 and n2 contracted over their shared bond.

    Leg order of the result: n1's other virtual legs in n1's order, n2's other
    virtual legs in n2's order, then n1's and n2's physical legs.  This is the
    merged node of ``TTN.contract_nodes`` (``core/ttn.py:1072-1115``) with the
    open legs last in (n1, n2) order, as ``legs_before_combination``
    (``core/ttn.py:1163-1210``) records them; the virtual legs keep each node's
    own order so that ``split_two_sites`` can put every leg back in place."""
    A1, A2 = ttn.tensors[n1], ttn.tensors[n2]
    i12 = ttn.nodes[n1].neighbour_index(n2)
    i21 = ttn.nodes[n2].neighbour_index(n1)
    l1 = [("o1", i) for i in range(A1.dim() - 1)] + ["p1"]
    l1[i12] = "bond"
    l2 = [("o2", i) for i in range(A2.dim() - 1)] + ["p2"]
    l2[i21] = "bond"
    outer1 = [x for x in l1[:-1] if x != "bond"]
    outer2 = [x for x in l2[:-1] if x != "bond"]
    return E.contract(A1, l1, A2, l2, outer1 + outer2 + ["p1", "p2"])


def split_two_sites(ttn, x: torch.Tensor, n1: str, n2: str, svd_params,
                    contr_mode: "ContractionMode" = None) -> np.ndarray:
    """``split_two_sites_gen`` with its SVD answered at once."""
    return drive_svd(split_two_sites_gen(ttn, x, n1, n2, svd_params, contr_mode))


def split_two_sites_gen(ttn, x: torch.Tensor, n1: str, n2: str, svd_params,
                        contr_mode: "ContractionMode" = None):
    """Split a merged two-site tensor (legs as ``merge_two_sites`` returns them)
    back into n1 and n2 by a truncated SVD, and write both node tensors.

    U goes to n1 and S Vh to n2 (``ContractionMode.VCONTR``, the default of
    ``contr_truncated_svd_splitting``, ``util/tensor_splitting.py:541-572``, as
    used by ``TTN.split_node_svd``, ``core/ttn.py:1476-1497``); UCONTR puts S on
    U, EQUAL splits sqrt(S) over both.  The new bond takes the place of the old
    one in each node.  Returns the kept singular values (host).  A generator: the
    dense SVD is handed out as an ``SvdRequest``."""
    if contr_mode is None:
        contr_mode = ContractionMode.VCONTR
    A1, A2 = ttn.tensors[n1], ttn.tensors[n2]
    i12 = ttn.nodes[n1].neighbour_index(n2)
    i21 = ttn.nodes[n2].neighbour_index(n1)
    k1, k2 = A1.dim() - 2, A2.dim() - 2
    dims = list(x.shape)
    udims = dims[:k1] + [dims[k1 + k2]]
    vdims = dims[k1:k1 + k2] + [dims[k1 + k2 + 1]]
    u_legs = list(range(k1)) + [k1 + k2]
    v_legs = list(range(k1, k1 + k2)) + [k1 + k2 + 1]
    m, n = int(np.prod(udims)), int(np.prod(vdims))
    panel = _colmajor_panel(x, u_legs, v_legs)
    svals, U, Vh, kfull = yield from svd_colmajor_gen(panel, m, n)
    s_new, _ = truncate_singular_values(svals, svd_params)
    r = len(s_new)
    dev = x.device
    s_u = s_v = None
    if contr_mode == ContractionMode.VCONTR:
        s_v = s_new
    elif contr_mode == ContractionMode.UCONTR:
        s_u = s_new
    else:
        s_u = s_v = np.sqrt(s_new)
    # U (m x kfull, column-major) -> n1's layout with the new bond at i12
    shape1 = [A1.shape[i] for i in range(A1.dim())]
    shape1[i12] = r
    new1 = torch.empty(shape1, dtype=torch.complex128, device=dev)
    st1 = list(new1.stride())
    dst_u = [st1[i12]] + [st1[j] for j in range(A1.dim()) if j != i12]
    E.strided_copy([r] + udims, U, [m] + _cstrides(udims), new1, dst_u)
    if s_u is not None:
        new1 = scale_leg(new1, i12, s_u)
    # Vh (kfull x n, column-major) -> n2's layout with the new bond at i21
    vh_view = torch.as_strided(Vh, [r] + vdims, [1] + [kfull * c for c in _cstrides(vdims)])
    lv = ["old"] + [("v", i) for i in range(len(vdims))]
    lo = [("v", i) for i in range(len(vdims) - 1)]
    lo.insert(i21, "new")
    lo.append(("v", len(vdims) - 1))
    scale = s_v if s_v is not None else np.ones(r)
    sdiag = torch.as_tensor(np.diag(np.asarray(scale, dtype=np.complex128))).to(dev)
    new2 = E.contract(sdiag, ["new", "old"], vh_view, lv, lo)
    ttn.tensors[n1] = new1
    ttn.tensors[n2] = new2
    return np.asarray(s_new)
