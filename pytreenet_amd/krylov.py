"""Matrix-free Krylov exponential on the device, and the time-evolution enums.

``krylov_exp_action`` keeps the reference algorithm and stop rules
(``time_evolution/time_evolution.py:696-821``): Arnoldi with two
re-orthogonalisation passes accumulated into one Hessenberg column, happy
breakdown at beta <= max(eps*n, 1e-12*||A q_j||), the Saad estimate
||psi|| * beta * |[exp(f dt T_m) e_1]_{m-1}| < krylov_tol, and the small
propagator from eigh((T+T^H)/2) (hermitian) or expm(T).

MI355X form: the basis lives in one device buffer.  q_0 = psi / ||psi|| is
formed on the device (``ptn_normalize``).  Each step's two projection passes
run as block classical Gram-Schmidt (CGS2 -- two passes of Q^H w / w -= Q h,
each streaming the basis once) inside ``ptn_krylov_orthonormalize``, which
also returns ||w||^2 before and after and writes q_{j+1} = w / ||w||, so the
next H_eff application does not wait for the host.  One asynchronous
device->host copy of the (j+3) scalars per step feeds the host-side m x m
propagator (m <= max dim, FP64), as in the reference.  For short vectors
(n <= LOOKAHEAD_N) the next application is queued before that copy is read:
if the stop rule then ends the recurrence, that one application is discarded
-- the accepted q_0..q_{m-1} and T_m are the same either way.  Longer vectors
queue it only when the last two Saad estimates make a stop at this step
unlikely (``_stop_unlikely``), since a discarded application costs more there
than the device idling while the host reads the scalars.  The two-pass CGS
and the reference's two-pass MGS differ only in rounding.
"""
from __future__ import annotations

import ctypes
import math
from enum import Enum
from contextlib import contextmanager
from typing import Any, Callable, Dict, List, Optional, Union

import numpy as np
import torch
from scipy.linalg import expm

from . import _lib
from . import engine as E


class EvoDirection(Enum):
    """FORWARD evolves with exp(-i dt H), BACKWARD with exp(+i dt H)."""
    FORWARD = -1
    BACKWARD = 1

    @staticmethod
    def from_bool(forward: bool) -> "EvoDirection":
        return EvoDirection.FORWARD if forward else EvoDirection.BACKWARD

    def exp_sign(self) -> int:
        return self.value

    def exp_factor(self) -> complex:
        return 1.0j * self.exp_sign()


class TimeEvoMode(Enum):
    """Solver selection (``time_evolution.py:431-655``).  KRYLOV is the device path."""
    FASTEST = "fastest"
    EXPM = "expm"
    CHEBYSHEV = "chebyshev"
    SPARSE = "sparse"
    RK45 = "RK45"
    RK23 = "RK23"
    DOP853 = "DOP853"
    BDF = "BDF"
    KRYLOV = "krylov"

    def __str__(self):
        return self.value

    @staticmethod
    def default_solver_options(mode: "TimeEvoMode") -> Dict[str, Any]:
        if mode is TimeEvoMode.KRYLOV:
            return {"krylov_tol": 1e-10}
        return {}

    @staticmethod
    def valid_solver_options(mode: "TimeEvoMode") -> set:
        if mode is TimeEvoMode.KRYLOV:
            return {"krylov_tol", "max_krylov_dim", "hermitian", "exp_factor"}
        return {"exp_factor"}

    def action_evolvable(self) -> bool:
        return self is TimeEvoMode.KRYLOV

    def is_dense(self) -> bool:
        """Modes that exponentiate a dense matrix (``util/std_utils.py:134-175``)."""
        return self in (TimeEvoMode.FASTEST, TimeEvoMode.EXPM, TimeEvoMode.CHEBYSHEV,
                        TimeEvoMode.SPARSE)

    def time_evolve(self, psi: torch.Tensor, hamiltonian: torch.Tensor,
                    time_difference: float,
                    forward: Union[EvoDirection, bool] = EvoDirection.FORWARD,
                    **options) -> torch.Tensor:
        """exp(f dt H) psi for a dense device matrix H (``time_evolution.py:617-655``).

        KRYLOV runs the Krylov recurrence on H @ y.  The dense modes (FASTEST, EXPM,
        CHEBYSHEV, SPARSE) all compute the exact exponential action, here as
        ``torch.linalg.matrix_exp`` of the (n x n) exponent on the device.
        """
        if isinstance(forward, bool):
            forward = EvoDirection.from_bool(forward)
        if self.action_evolvable():
            if psi.numel() <= SMALL_N:
                y = small_time_evolve(psi, hamiltonian, time_difference, forward, options)
                if y is not None:
                    return y
            return self.time_evolve_action(psi, lambda t, y: (hamiltonian @ y.reshape(-1)),
                                           time_difference, forward=forward, **options)
        if not self.is_dense():
            raise NotImplementedError(f"mode {self.value}: no device implementation")
        E.require_device(psi, hamiltonian)
        override = options.pop("exp_factor", None)
        exp_factor = (forward.exp_factor() if override is None
                      else (override if forward is EvoDirection.FORWARD else -override))
        return dense_exp_action(hamiltonian, psi, exp_factor * time_difference)

    def time_evolve_action(self, psi: torch.Tensor, time_evo_action: Callable,
                           time_difference: float,
                           forward: Union[EvoDirection, bool] = EvoDirection.FORWARD,
                           **options) -> torch.Tensor:
        if isinstance(forward, bool):
            forward = EvoDirection.from_bool(forward)
        override = options.pop("exp_factor", None)
        exp_factor = (forward.exp_factor() if override is None
                      else (override if forward is EvoDirection.FORWARD else -override))
        hermitian = options.pop("hermitian", True)
        if self is not TimeEvoMode.KRYLOV:
            raise NotImplementedError(
                f"mode {self.value}: the device path evolves matrix-free with "
                "TimeEvoMode.KRYLOV only")
        relaxed = options.pop("relaxed", None)
        return krylov_exp_action(psi, lambda y: time_evo_action(0.0, y), time_difference,
                                 exp_factor, krylov_tol=options.pop("krylov_tol", 1e-10),
                                 max_krylov_dim=options.pop("max_krylov_dim", None),
                                 hermitian=hermitian, relaxed=relaxed)


def dense_exp_action(hamiltonian: torch.Tensor, psi: torch.Tensor, factor: complex
                     ) -> torch.Tensor:
    """exp(factor * H) psi with H dense (n x n) on the device, complex128."""
    H = hamiltonian.to(torch.complex128)
    n = psi.numel()
    if H.shape != (n, n):
        raise ValueError(f"dense operator {tuple(H.shape)} does not act on a vector of {n}")
    U = torch.linalg.matrix_exp(H * complex(factor))
    y = U @ psi.to(torch.complex128).reshape(-1)
    return y.reshape(psi.shape)


# ---------------------------------------------------------------------------------------
# device level-1 helpers
# ---------------------------------------------------------------------------------------
_WORK: Dict[torch.device, torch.Tensor] = {}


def _work(device) -> torch.Tensor:
    w = _WORK.get(device)
    if w is None:
        nbytes = int(_lib.load().ptn_krylov_work_bytes(0, 1))
        w = torch.empty(nbytes // 16 + 1, dtype=torch.complex128, device=device)
        _WORK[device] = w
    return w


def _vp(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _flat(t: torch.Tensor) -> torch.Tensor:
    if t.dtype != torch.complex128:
        raise TypeError("Krylov vectors are complex128")
    return t.reshape(-1) if t.is_contiguous() else t.contiguous().reshape(-1)


def dotc(x: torch.Tensor, y: torch.Tensor) -> complex:
    """sum conj(x) * y (complex128 device vectors)."""
    E.require_device(x, y)
    xf, yf = _flat(x), _flat(y)
    out = torch.empty(1, dtype=torch.complex128, device=x.device)
    _lib.check(_lib.load().ptn_dotc(ctypes.c_void_p(E.stream_handle()), xf.numel(), _vp(xf),
                                    _vp(yf), _vp(out), _vp(_work(x.device))), "ptn_dotc")
    return complex(out.cpu().item())


def norm_sq(x: torch.Tensor) -> float:
    E.require_device(x)
    xf = _flat(x)
    out = torch.empty(1, dtype=torch.complex128, device=x.device)
    _lib.check(_lib.load().ptn_norm_sq(ctypes.c_void_p(E.stream_handle()), xf.numel(),
                                       _vp(xf), _vp(out), _vp(_work(x.device))),
               "ptn_norm_sq")
    return float(out.cpu().item().real)


def combine(Q: torch.Tensor, coeffs, out: torch.Tensor, nvec: Optional[int] = None):
    """out = sum_i coeffs[i] Q[i] over the first nvec rows of Q."""
    nvec = len(coeffs) if nvec is None else nvec
    c = torch.as_tensor(np.asarray(coeffs, dtype=np.complex128)).pin_memory().to(
        Q.device, non_blocking=True)
    lib = _lib.load()
    stream = ctypes.c_void_p(E.stream_handle())
    if nvec <= 128:
        _lib.check(lib.ptn_krylov_combine(stream, Q.shape[1], _vp(Q), Q.stride(0), nvec,
                                          _vp(c), _vp(out)), "ptn_krylov_combine")
        return out
    part = torch.empty_like(out)
    for c0 in range(0, nvec, 128):
        cnt = min(128, nvec - c0)
        dst = out if c0 == 0 else part
        _lib.check(lib.ptn_krylov_combine(stream, Q.shape[1], _vp(Q[c0]), Q.stride(0), cnt,
                                          _vp(c[c0:]), _vp(dst)), "ptn_krylov_combine")
        if c0:
            out.add_(part)
    return out


def lincomb(tensors, coeffs, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """out = sum_i coeffs[i] * tensors[i] (same shape, complex128), on the device."""
    t0 = tensors[0]
    n = t0.numel()
    buf = torch.empty((len(tensors), n), dtype=torch.complex128, device=t0.device)
    for i, t in enumerate(tensors):
        E.strided_copy([n], _flat(t), [1], buf[i], [1])
    if out is None:
        out = torch.empty_like(t0, dtype=torch.complex128)
    combine(buf, coeffs, out.reshape(-1))
    return out


def small_propagator(t_mat: np.ndarray, dt: float, exp_factor: complex,
                     hermitian: bool = True) -> np.ndarray:
    """First column of exp(exp_factor * dt * T) for the m x m projection T (host FP64)."""
    if hermitian:
        herm = 0.5 * (t_mat + t_mat.conj().T)
        evals, vecs = np.linalg.eigh(herm)
        return vecs @ (np.exp(exp_factor * dt * evals) * vecs[0].conj())
    return expm(exp_factor * dt * t_mat)[:, 0]


class KrylovStats:
    def __init__(self):
        self.calls = 0
        self.applies = 0
        self.low_applies = 0     # applications run in reduced precision (relaxed form)
        self.lower_applies = 0   # of them, in bf16
        self.dims = []

    def record(self, m: int):
        self.calls += 1
        self.applies += m
        self.dims.append(m)


STATS = KrylovStats()


class PendingApplies:
    """Krylov dimension of a one-launch solve whose info has not been read back
    yet (``deferred_small_info``); ``resolve`` turns it into an int."""
    __slots__ = ("index",)

    def __init__(self, index: int):
        self.index = index


class _SmallDefer:
    def __init__(self):
        self.active = False
        self.infos: List[torch.Tensor] = []


_SMALL_DEFER = _SmallDefer()


@contextmanager
def deferred_small_info():
    """Within the block, a one-launch Krylov solve whose depth cannot run out
    (the allowed depth fits the kernel) returns its result without reading its
    (m, zero, stopped) info back, which would stall the device queue once per
    solve.  The solve's Krylov dimension is handed out as a ``PendingApplies``
    (``last_pending``); ``resolve_pending`` reads every info of the block with
    one copy, records the dimensions in STATS and maps each PendingApplies to
    its int."""
    prev = (_SMALL_DEFER.active, _SMALL_DEFER.infos)
    _SMALL_DEFER.active, _SMALL_DEFER.infos = True, []
    try:
        yield _SMALL_DEFER
    finally:
        _SMALL_DEFER.active, _SMALL_DEFER.infos = prev


def pending_count() -> int:
    """Number of deferred one-launch solves so far in the current block."""
    return len(_SMALL_DEFER.infos) if _SMALL_DEFER.active else 0


def resolve_pending(defer: _SmallDefer) -> List[int]:
    """Krylov dimensions of the block's deferred solves (one device -> host copy),
    recorded in STATS in solve order."""
    if not defer.infos:
        return []
    ms = [int(v) for v in torch.stack(defer.infos)[:, 0].cpu().tolist()]
    for m in ms:
        STATS.record(m)
    defer.infos = []
    return ms


LOOKAHEAD_N = 1 << 16
MAXV_PER_CALL = 128      # basis vectors per ptn_krylov_orthonormalize call
# longer vectors: the next application is queued before the scalars of a step are
# read when the estimate, extrapolated SPECULATE_STEPS steps at the rate of the
# last two, stays above krylov_tol (0 switches this off)
SPECULATE_STEPS = 2
# a speculative reduced-precision application must satisfy the relaxed rule with
# this much to spare (its weight uses the previous step's ||T||)
SPECULATE_MARGIN = 4.0


# with fewer than two estimates, the first one must exceed krylov_tol this much
SPECULATE_FIRST = 1e6


def _stop_unlikely(est: float, est_prev: float, krylov_tol: float) -> bool:
    """Whether the Saad estimate of the step being finished, predicted from the
    previous two, stays above krylov_tol (the recurrence then continues).  Before
    the first estimate (step 0) only a breakdown could stop the recurrence; after
    it (step 1) the first estimate has to lie SPECULATE_FIRST above krylov_tol."""
    if SPECULATE_STEPS <= 0:
        return False
    if krylov_tol <= 0.0 or est == math.inf:
        return True
    if not math.isfinite(est_prev):
        return math.isfinite(est) and est > krylov_tol * SPECULATE_FIRST
    if est_prev <= 0.0:
        return False
    rate = min(est / est_prev, 1.0)
    return est * rate ** SPECULATE_STEPS > krylov_tol


def _orthonormalize_chunked(lib, stream, n: int, Q: torch.Tensor, nvec: int, w: torch.Tensor,
                            packed: torch.Tensor, work: torch.Tensor):
    """The orthonormalisation step for a basis longer than one call takes: the
    basis is projected out chunk by chunk (two passes per chunk), the chunks'
    coefficients land in packed[1 : nvec+1] and (|w_in|^2, |w_out|^2) in
    packed[nvec+1], as the one-call form writes them."""
    tmp = torch.empty(MAXV_PER_CALL + 1, dtype=torch.complex128, device=Q.device)
    first = None
    for c0 in range(0, nvec, MAXV_PER_CALL):
        cnt = min(MAXV_PER_CALL, nvec - c0)
        last = c0 + cnt == nvec
        _lib.check(lib.ptn_krylov_orthonormalize(
            stream, n, _vp(Q[c0]), Q.stride(0), cnt, _vp(w),
            _vp(Q[nvec]) if last else ctypes.c_void_p(0), _vp(tmp), _vp(work)),
            "ptn_krylov_orthonormalize")
        packed[1 + c0: 1 + c0 + cnt].copy_(tmp[:cnt])
        if first is None:
            first = tmp[cnt].real.clone()
        if last:
            packed[nvec + 1] = torch.complex(first, tmp[cnt].imag)
_PINNED: List[Optional[torch.Tensor]] = [None]


def _pinned(count: int) -> torch.Tensor:
    """Pinned host staging of at least ``count`` complex128 scalars.  One buffer,
    grown by doubling and reused by every solve (each solve has read it back
    before it returns), so its size follows the longest recurrence run."""
    buf = _PINNED[0]
    if buf is None or buf.numel() < count:
        size = max(count, 64 if buf is None else 2 * buf.numel())
        buf = torch.empty(size, dtype=torch.complex128, pin_memory=True)
        _PINNED[0] = buf
    return buf[:count]


def _grow_hess(hess: np.ndarray, cap: int) -> np.ndarray:
    """The (cap+1) x (cap+1) Hessenberg buffer holding ``hess`` in its top-left corner."""
    out = np.zeros((cap + 1, cap + 1), dtype=complex)
    out[: hess.shape[0], : hess.shape[1]] = hess
    return out


# columns of T held at the start of a recurrence; buffers double as it runs longer
KRYLOV_START_CAP = 16


def krylov_exp_action(psi: torch.Tensor, action: Callable[[torch.Tensor], torch.Tensor],
                      time_difference: float, exp_factor: complex,
                      krylov_tol: float = 1e-10, max_krylov_dim: Optional[int] = None,
                      hermitian: bool = True, relaxed=None) -> torch.Tensor:
    """exp(exp_factor * dt * A) psi with A given by ``action`` (device tensors).

    ``relaxed`` (an ``engine.RelaxedNetwork`` of the same A): application j runs
    its low-precision form when tau * c_j * low_eps * ||T|| <= relax_tol * ||psi||
    (relax_tol bounds the change of the result relative to ||psi||), with
    tau = |exp_factor * dt|, ||T|| the largest column sum of |T| so far (it bounds
    ||A q_j||, hence the error of one application over low_eps) and c_j the
    latest Saad estimate beta_0 * beta * |[exp(f dt T) e_1]_{m-1}|.  In the
    perturbed Arnoldi relation the error of application j enters the result
    weighted by at most tau times the coefficient of q_j, which the estimate of
    step j-1 approximates; the estimates decrease along the recurrence, so an
    older one (the look-ahead uses the estimate of step j-2) overstates the
    weight.  The first two applications always run in full precision.  When the
    relaxed network also holds a bf16 form, an application whose weight allows
    lower_eps instead of low_eps runs in bf16."""
    E.require_device(psi)
    shape = psi.shape
    dev = psi.device
    v0 = _flat(psi.to(torch.complex128))
    n = v0.numel()
    max_steps = n if max_krylov_dim is None else min(n, int(max_krylov_dim))
    lib = _lib.load()
    work = _work(dev)
    stream = ctypes.c_void_p(E.stream_handle())
    # Q, T and the scalar staging grow by doubling (the reference grows its
    # Hessenberg buffer the same way, time_evolution.py:782-787): with no
    # max_krylov_dim the recurrence may run up to n steps, but it stops long
    # before that, and buffers sized for n would need n^2 elements.
    cap = min(max_steps, KRYLOV_START_CAP)
    Q = torch.empty((cap + 1, n), dtype=torch.complex128, device=dev)
    # packed[0] = (||psi||^2, 0); packed[1..j+2) = T[:j+1, j]; packed[j+2] = (w0sq, beta^2)
    packed = torch.empty(cap + 3, dtype=torch.complex128, device=dev)
    _lib.check(lib.ptn_normalize(stream, n, _vp(v0), _vp(Q[0]), _vp(packed), _vp(work)),
               "ptn_normalize")
    hess = np.zeros((cap + 1, cap + 1), dtype=complex)
    breakdown = np.finfo(float).eps * max(n, 1)
    done = torch.cuda.Event()
    lookahead = n <= LOOKAHEAD_N
    beta0 = 0.0
    w_next = None
    m = 0
    low = None
    if relaxed is not None and relaxed.relax_tol > 0.0:
        low = relaxed.low
        tau = abs(complex(exp_factor) * time_difference)
    est = math.inf          # Saad estimate of the latest step
    est_prev = math.inf     # and of the step before
    t_norm = 0.0            # largest column sum of |T| so far

    def apply(k, speculative=False):
        """H_eff q_k; ``speculative``: queued before the scalars of step k-1 are
        read -- a relaxed application then runs in reduced precision only if the
        rule holds with SPECULATE_MARGIN to spare, and None is returned when it
        would have to run in full precision (the caller then waits and decides
        on the newer estimate)."""
        if low is not None and k >= 2:
            weight = tau * est * t_norm * (SPECULATE_MARGIN if speculative else 1.0)
            net = None
            if relaxed.lower is not None and weight * relaxed.lower_eps <= \
                    relaxed.relax_tol * beta0:
                net = relaxed.lower
                STATS.lower_applies += 1
            elif weight * relaxed.low_eps <= relaxed.relax_tol * beta0:
                net = low
            if net is not None:
                STATS.low_applies += 1
                # complex64 input: every product of the reduced-precision network
                # then has complex64 operands
                return net(Q[k].view(shape).to(torch.complex64))
            if speculative:
                return None
        return action(Q[k].view(shape))

    for j in range(max_steps):
        if j >= cap:
            # column j of T and basis vector j+1 do not fit: double every buffer
            # (the look-ahead application only reads Q[j], which is copied)
            cap = min(2 * cap, max_steps)
            grown = torch.empty((cap + 1, n), dtype=torch.complex128, device=dev)
            grown[: Q.shape[0]].copy_(Q)
            Q = grown
            grown_p = torch.empty(cap + 3, dtype=torch.complex128, device=dev)
            grown_p[: packed.numel()].copy_(packed)
            packed = grown_p
            hess = _grow_hess(hess, cap)
        host_buf = _pinned(j + 3)
        w = w_next if w_next is not None else apply(j)
        w_next = None
        w = _flat(w) if w.dtype == torch.complex128 else _flat(w.to(torch.complex128))
        if w.data_ptr() == Q[j].data_ptr():
            w = w.clone()
        if j + 1 <= MAXV_PER_CALL:
            _lib.check(lib.ptn_krylov_orthonormalize(stream, n, _vp(Q), Q.stride(0), j + 1,
                                                     _vp(w), _vp(Q[j + 1]), _vp(packed[1:]),
                                                     _vp(work)), "ptn_krylov_orthonormalize")
        else:
            _orthonormalize_chunked(lib, stream, n, Q, j + 1, w, packed, work)
        host_buf.copy_(packed[: j + 3], non_blocking=True)
        done.record()
        if j + 1 < max_steps:
            if lookahead:
                w_next = apply(j + 1)
            elif _stop_unlikely(est, est_prev, krylov_tol):
                w_next = apply(j + 1, speculative=True)
        done.synchronize()
        host = host_buf.numpy()
        if j == 0:
            beta0 = math.sqrt(max(host[0].real, 0.0))
            if beta0 == 0.0:
                return psi.clone()
        hess[: j + 1, j] = host[1: j + 2]
        w0_norm = math.sqrt(max(host[j + 2].real, 0.0))
        beta = math.sqrt(max(host[j + 2].imag, 0.0))
        hess[j + 1, j] = beta
        m = j + 1
        if beta <= max(breakdown, 1e-12 * w0_norm):
            break
        if krylov_tol > 0.0 or low is not None:
            col = small_propagator(hess[:m, :m], time_difference, exp_factor, hermitian)
            est_prev, est = est, beta0 * beta * abs(col[m - 1])
            t_norm = max(t_norm, float(np.abs(hess[: j + 2, j]).sum()))
            if krylov_tol > 0.0 and est < krylov_tol:
                break
    coeffs = beta0 * small_propagator(hess[:m, :m], time_difference, exp_factor, hermitian)
    out = torch.empty(n, dtype=torch.complex128, device=dev)
    combine(Q, coeffs, out, nvec=m)
    STATS.record(m)
    return out.view(shape)


# Local spaces up to SMALL_N run the whole recurrence in one launch
# (ptn_krylov_expm_small) on the dense effective Hamiltonian, for at most
# SMALL_MAX_STEPS steps.
SMALL_N = 256
SMALL_MAX_STEPS = 24
_PINNED_INFO: Dict[int, torch.Tensor] = {}


def krylov_exp_small(psi: torch.Tensor, hamiltonian: torch.Tensor, time_difference: float,
                     exp_factor: complex, krylov_tol: float = 1e-10,
                     max_krylov_dim: Optional[int] = None, hermitian: bool = True,
                     transposed: bool = False) -> Optional[torch.Tensor]:
    """exp(exp_factor * dt * H) psi for a dense device matrix H (n <= 1024) with the
    recurrence of krylov_exp_action run inside one kernel.  ``transposed``: the
    argument holds H^T (H in column-major order, as the kernel reads it).
    Returns None when the stop rules have not ended the recurrence within
    SMALL_MAX_STEPS steps and the allowed depth is larger (the caller then runs
    the general form)."""
    E.require_device(psi, hamiltonian)
    n = psi.numel()
    dev = psi.device
    H = hamiltonian.to(torch.complex128)
    if H.shape != (n, n):
        raise ValueError(f"dense operator {tuple(H.shape)} does not act on a vector of {n}")
    H = H.contiguous() if transposed else H.t().contiguous()
    x = _flat(psi.to(torch.complex128))
    max_steps = n if max_krylov_dim is None else min(n, int(max_krylov_dim))
    steps = min(max_steps, SMALL_MAX_STEPS)
    lib = _lib.load()
    work = torch.empty(n * (steps + 1), dtype=torch.complex128, device=dev)
    info = torch.empty(3, dtype=torch.int32, device=dev)
    y = torch.empty(n, dtype=torch.complex128, device=dev)
    stream = ctypes.c_void_p(E.stream_handle())
    f = complex(exp_factor)
    _lib.check(lib.ptn_krylov_expm_small(stream, n, _vp(H), _vp(x), _vp(y), steps,
                                         float(time_difference), f.real, f.imag,
                                         float(krylov_tol), 1 if hermitian else 0,
                                         _vp(work), _vp(info)), "ptn_krylov_expm_small")
    if _SMALL_DEFER.active and steps == max_steps:
        # the recurrence cannot need the general form (its whole allowed depth
        # runs in the kernel), and a zero input leaves y = x: y is final, only
        # the Krylov dimension is read later
        _SMALL_DEFER.infos.append(info)
        return y.view(psi.shape)
    dev_idx = dev.index or 0
    host = _PINNED_INFO.get(dev_idx)
    if host is None:
        host = torch.empty(3, dtype=torch.int32, pin_memory=True)
        _PINNED_INFO[dev_idx] = host
    host.copy_(info, non_blocking=True)
    done = torch.cuda.Event()
    done.record()
    done.synchronize()
    m, zero, stopped = (int(v) for v in host.tolist())
    if zero:
        return psi.clone()
    if not stopped and steps < max_steps:
        return None
    STATS.record(m)
    return y.view(psi.shape)


def small_time_evolve(psi: torch.Tensor, hamiltonian: torch.Tensor, time_difference: float,
                      forward: "EvoDirection", options: Dict[str, Any],
                      transposed: bool = False) -> Optional[torch.Tensor]:
    """TimeEvoMode.KRYLOV's exponential of a dense H through krylov_exp_small, with
    the solver options resolved as time_evolve_action resolves them (None: the
    recurrence needs more than SMALL_MAX_STEPS steps)."""
    override = options.get("exp_factor")
    exp_factor = (forward.exp_factor() if override is None
                  else (override if forward is EvoDirection.FORWARD else -override))
    return krylov_exp_small(psi, hamiltonian, time_difference, exp_factor,
                            krylov_tol=options.get("krylov_tol", 1e-10),
                            max_krylov_dim=options.get("max_krylov_dim"),
                            hermitian=options.get("hermitian", True), transposed=transposed)


# ---------------------------------------------------------------------------------------
# batched form
# ---------------------------------------------------------------------------------------
_BWORK: Dict[tuple, torch.Tensor] = {}


def _bwork(dev, count: int) -> torch.Tensor:
    key = (dev, count)
    w = _BWORK.get(key)
    if w is None:
        nb = int(_lib.load().ptn_krylov_batched_work_bytes(count))
        w = _BWORK[key] = torch.empty(nb, dtype=torch.uint8, device=dev)
    return w


def _ptrs(ts) -> np.ndarray:
    return np.array([t.data_ptr() for t in ts], dtype=np.int64)


def _cp(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def krylov_exp_action_batched(psis, apply: Callable, time_difference: float,
                              exp_factor: complex, krylov_tol: float = 1e-10,
                              max_krylov_dim: Optional[int] = None, hermitian: bool = True,
                              stats: Optional[list] = None):
    """``krylov_exp_action`` for many independent instances at once.

    ``apply(idxs, xs)`` returns H_b x for the active instances ``idxs``.  Each
    instance follows the single-instance recurrence and stop rules exactly
    (its own beta0, Hessenberg matrix, breakdown and Saad test); the vector
    work of all instances with the same length n runs as one batched launch
    per step, and one device->host copy per Krylov step carries the scalars of
    every active instance.
    """
    count = len(psis)
    dev = psis[0].device
    shapes = [p.shape for p in psis]
    v0 = [_flat(p.to(torch.complex128)) for p in psis]
    ns = [v.numel() for v in v0]
    steps = [n if max_krylov_dim is None else min(n, int(max_krylov_dim)) for n in ns]
    mmax = max(steps)
    lib = _lib.load()
    stream = ctypes.c_void_p(E.stream_handle())
    by_n: Dict[int, List[int]] = {}
    for b, n in enumerate(ns):
        by_n.setdefault(n, []).append(b)
    # every buffer starts at KRYLOV_START_CAP columns and doubles as the
    # recurrence runs longer (see krylov_exp_action)
    caps = [min(st, KRYLOV_START_CAP) for st in steps]
    Q = [torch.empty((caps[b] + 1, ns[b]), dtype=torch.complex128, device=dev)
         for b in range(count)]
    pcap = min(mmax, KRYLOV_START_CAP)
    packed = torch.empty((count, pcap + 3), dtype=torch.complex128, device=dev)
    pk = packed.data_ptr()
    prow = packed.stride(0) * 16
    for n, mem in by_n.items():
        work = _bwork(dev, len(mem))
        xs = _ptrs([v0[b] for b in mem])
        ys = _ptrs([Q[b][0] for b in mem])
        outs = np.array([pk + prow * b for b in mem], dtype=np.int64)
        _lib.check(lib.ptn_normalize_batched(stream, n, len(mem), _cp(xs), _cp(ys), _cp(outs),
                                             ctypes.c_void_p(work.data_ptr())),
                   "ptn_normalize_batched")
    host_buf = torch.empty((count, pcap + 3), dtype=torch.complex128, pin_memory=True)
    done_ev = torch.cuda.Event()
    hess = [np.zeros((caps[b] + 1, caps[b] + 1), dtype=complex) for b in range(count)]
    beta0 = [0.0] * count
    m = [0] * count
    active = list(range(count))
    zero = [False] * count
    for j in range(mmax):
        active = [b for b in active if j < steps[b]]
        if not active:
            break
        if j >= pcap:
            pcap = min(2 * pcap, mmax)
            grown = torch.empty((count, pcap + 3), dtype=torch.complex128, device=dev)
            grown[:, : packed.shape[1]].copy_(packed)
            packed = grown
            pk = packed.data_ptr()
            prow = packed.stride(0) * 16
            host_buf = torch.empty((count, pcap + 3), dtype=torch.complex128,
                                   pin_memory=True)
        for b in active:
            if j >= caps[b]:
                caps[b] = min(2 * caps[b], steps[b])
                grown = torch.empty((caps[b] + 1, ns[b]), dtype=torch.complex128, device=dev)
                grown[: Q[b].shape[0]].copy_(Q[b])
                Q[b] = grown
                hess[b] = _grow_hess(hess[b], caps[b])
        ws = apply(active, [Q[b][j].view(shapes[b]) for b in active])
        ws = [_flat(w) if w.dtype == torch.complex128 else _flat(w.to(torch.complex128))
              for w in ws]
        wmap = dict(zip(active, ws))
        groups: Dict[int, List[int]] = {}
        for b in active:
            groups.setdefault(ns[b], []).append(b)
        for n, mem in groups.items():
            if j + 1 > MAXV_PER_CALL:
                # bases longer than one call: the single-instance chunked form
                for b in mem:
                    _orthonormalize_chunked(lib, stream, n, Q[b], j + 1, wmap[b], packed[b],
                                            _work(dev))
                continue
            work = _bwork(dev, len(mem))
            qs = _ptrs([Q[b] for b in mem])
            ld = np.array([Q[b].stride(0) for b in mem], dtype=np.int64)
            wp = _ptrs([wmap[b] for b in mem])
            qn = _ptrs([Q[b][j + 1] for b in mem])
            outs = np.array([pk + prow * b + 16 for b in mem], dtype=np.int64)
            _lib.check(lib.ptn_krylov_orthonormalize_batched(
                stream, n, j + 1, len(mem), _cp(qs), _cp(ld), _cp(wp), _cp(qn), _cp(outs),
                ctypes.c_void_p(work.data_ptr())), "ptn_krylov_orthonormalize_batched")
        host_buf[:, : j + 3].copy_(packed[:, : j + 3], non_blocking=True)
        done_ev.record()
        done_ev.synchronize()
        host = host_buf[:, : j + 3].numpy()
        still = []
        for b in active:
            h = host[b]
            if j == 0:
                beta0[b] = math.sqrt(max(h[0].real, 0.0))
                if beta0[b] == 0.0:
                    zero[b] = True
                    continue
            hess[b][: j + 1, j] = h[1: j + 2]
            w0_norm = math.sqrt(max(h[j + 2].real, 0.0))
            beta = math.sqrt(max(h[j + 2].imag, 0.0))
            hess[b][j + 1, j] = beta
            m[b] = j + 1
            if beta <= max(np.finfo(float).eps * max(ns[b], 1), 1e-12 * w0_norm):
                continue
            if krylov_tol > 0.0:
                col = small_propagator(hess[b][: m[b], : m[b]], time_difference, exp_factor,
                                       hermitian)
                if beta0[b] * beta * abs(col[m[b] - 1]) < krylov_tol:
                    continue
            still.append(b)
        active = still
    results: List[Optional[torch.Tensor]] = [None] * count
    for b in range(count):
        if zero[b]:
            results[b] = psis[b].clone()
    live = [b for b in range(count) if not zero[b]]
    by_nm: Dict[tuple, List[int]] = {}
    for b in live:
        by_nm.setdefault((ns[b], m[b]), []).append(b)
    for (n, mb), mem in by_nm.items():
        coeffs = np.stack([beta0[b] * small_propagator(hess[b][:mb, :mb], time_difference,
                                                       exp_factor, hermitian)
                           for b in mem]).astype(np.complex128)
        outs = [torch.empty(n, dtype=torch.complex128, device=dev) for _ in mem]
        work = _bwork(dev, len(mem))
        qs = _ptrs([Q[b] for b in mem])
        ld = np.array([Q[b].stride(0) for b in mem], dtype=np.int64)
        op = _ptrs(outs)
        _lib.check(lib.ptn_krylov_combine_batched(stream, n, mb, len(mem), _cp(qs), _cp(ld),
                                                  ctypes.c_void_p(coeffs.ctypes.data), _cp(op),
                                                  ctypes.c_void_p(work.data_ptr())),
                   "ptn_krylov_combine_batched")
        for b, o in zip(mem, outs):
            results[b] = o.view(shapes[b])
    for b in range(count):
        STATS.record(m[b])
        if stats is not None:
            stats.append(m[b])
    return results
