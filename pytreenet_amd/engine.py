"""Tensor-network contraction engine over the gather-GEMM kernel.

A pairwise contraction of labelled tensors becomes one ``ptn_gemm_gather``
launch: the labels of A that are not in B form M, the shared labels form K,
the labels of B that are not in A form N.  The flattened (M, K), (K, N),
(M, N) index spaces are described by per-mode element-offset tables built once
per shape/stride signature and cached in device memory, so permuted views and
arbitrary output layouts cost nothing extra at run time.

A network of several tensors (H_eff applies, environment updates, bond
actions) is contracted in the pairwise order that minimises the flop count
(exhaustive dynamic programming over subsets, cached per signature).  The
reference contracts the same networks with a fixed ket-first ``np.tensordot``
sequence (``contractions/contraction_util.py:176-290``,
``contractions/state_operator_contraction.py:227-273``); the order is free
because only the summation order differs.
"""
from __future__ import annotations

import ctypes
import math
from collections import OrderedDict
from dataclasses import dataclass
from typing import Dict, Hashable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import GemmDesc, PtnBackendError

Label = Hashable

FP64 = "fp64"
FP32 = "fp32"
_COMPUTE = {FP64: _lib.PTN_COMPUTE_F64_4M, FP32: _lib.PTN_COMPUTE_F32_3M}
_DTYPE_CODE = {torch.complex128: _lib.PTN_C128, torch.complex64: _lib.PTN_C64}


def storage_dtype(precision: str) -> torch.dtype:
    return torch.complex128 if precision == FP64 else torch.complex64


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def require_device(*ts: torch.Tensor):
    for t in ts:
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise PtnBackendError(
                "pytreenet_amd computes on the ROCm device only; got a tensor on "
                f"{getattr(t, 'device', type(t))} (there is no CPU fallback)")
        if t.dtype not in _DTYPE_CODE:
            raise PtnBackendError(f"unsupported dtype {t.dtype}; use complex128/complex64")


# --------------------------------------------------------------------------------------
# offset tables
# --------------------------------------------------------------------------------------
def _offsets(dims: Sequence[int], strides: Sequence[int]) -> np.ndarray:
    off = np.zeros(1, dtype=np.int64)
    for d, s in zip(dims, strides):
        off = (off[:, None] + np.arange(d, dtype=np.int64)[None, :] * s).reshape(-1)
    return off


class _OffsetCache:
    """Device-resident offset tables keyed by contraction signature (LRU)."""

    def __init__(self, capacity: int = 4096):
        self.cap = capacity
        self.d: "OrderedDict[tuple, tuple]" = OrderedDict()

    def get(self, key, builder, device):
        hit = self.d.get(key)
        if hit is not None:
            self.d.move_to_end(key)
            return hit
        tables = builder()
        sizes = [len(t) for t in tables]
        host = np.concatenate([t if len(t) else np.zeros(1, np.int64) for t in tables])
        dev = torch.from_numpy(host).to(device)
        ptrs, pos = [], 0
        for t, n in zip(tables, sizes):
            ptrs.append(dev.data_ptr() + pos * 8)
            pos += max(n, 1)
        val = (dev, ptrs)
        self.d[key] = val
        if len(self.d) > self.cap:
            self.d.popitem(last=False)
        return val


_CACHE = _OffsetCache()


class LaunchProfiler:
    """Executed-flop counters and HIP-event timings of gather-GEMM launches.

    Events are recorded on the stream the kernels are launched on (torch's
    current stream) around every launch with at least ``min_flops``; the
    per-shape averages give the duration of the dominant kernel.
    """

    def __init__(self, min_flops: float = 1e9):
        self.min_flops = min_flops
        self.active = False
        self.reset()

    def reset(self):
        self.flops = 0.0
        self.launches = 0
        self.events: Dict[tuple, list] = {}

    def start(self):
        self.reset()
        self.active = True

    def stop(self):
        self.active = False

    def summary(self) -> List[dict]:
        torch.cuda.synchronize()
        rows = []
        for key, evs in self.events.items():
            ms = [a.elapsed_time(b) for a, b in evs]
            rows.append({"M": key[0], "N": key[1], "K": key[2], "compute": key[3],
                         "flops_per_launch": 8.0 * key[0] * key[1] * key[2],
                         "launches": len(ms), "total_ms": float(sum(ms)),
                         "avg_ms": float(sum(ms) / len(ms))})
        rows.sort(key=lambda r: -r["total_ms"])
        return rows


PROFILER = LaunchProfiler()


def clear_caches():
    _CACHE.d.clear()
    _PLANS.clear()


_SPLITK_WS: Dict[torch.device, torch.Tensor] = {}
SPLITK_WS_BYTES = 256 << 20


def _splitk_workspace(device) -> torch.Tensor:
    """Per-device scratch for split-K partial sums (kernels on one stream share it)."""
    ws = _SPLITK_WS.get(device)
    if ws is None:
        ws = torch.empty(SPLITK_WS_BYTES // 16, dtype=torch.complex128, device=device)
        _SPLITK_WS[device] = ws
    return ws


def _inner_is_unit(dims, strides) -> bool:
    for d, s in zip(reversed(list(dims)), reversed(list(strides))):
        if d > 1:
            return s == 1
    return False


# --------------------------------------------------------------------------------------
# pairwise contraction
# --------------------------------------------------------------------------------------
def contract(A: torch.Tensor, la: Sequence[Label], B: torch.Tensor, lb: Sequence[Label],
             out_labels: Sequence[Label], *, conj_a: bool = False, conj_b: bool = False,
             precision: str = FP64, out: Optional[torch.Tensor] = None,
             out_dtype: Optional[torch.dtype] = None, accumulate: bool = False,
             alpha: complex = 1.0) -> torch.Tensor:
    """sum over shared labels of opA(A) * opB(B), laid out as ``out_labels``."""
    require_device(A, B)
    la, lb, lo = tuple(la), tuple(lb), tuple(out_labels)
    if len(la) != A.dim() or len(lb) != B.dim():
        raise ValueError("label count does not match tensor rank")
    if len(set(la)) != len(la) or len(set(lb)) != len(lb):
        raise ValueError("repeated label within one operand")
    dim = {}
    for l, d in list(zip(la, A.shape)) + list(zip(lb, B.shape)):
        if dim.setdefault(l, d) != d:
            raise ValueError(f"dimension mismatch on label {l!r}")
    k_l = [l for l in la if l in lb]
    m_l = [l for l in la if l not in lb]
    n_l = [l for l in lb if l not in la]
    if sorted(map(repr, m_l + n_l)) != sorted(map(repr, lo)) or len(lo) != len(m_l) + len(n_l):
        raise ValueError(f"output labels {lo} must be the open labels {m_l + n_l}")
    oshape = tuple(dim[l] for l in lo)
    if out is None:
        odt = out_dtype or (A.dtype if A.dtype == B.dtype else torch.complex128)
        if precision == FP64:
            odt = torch.complex128
        out = torch.empty(oshape, dtype=odt, device=A.device)
    else:
        require_device(out)
        if tuple(out.shape) != oshape:
            raise ValueError("out has the wrong shape")
    if out.numel() == 0:
        return out
    if precision == FP64 and (A.dtype != torch.complex128 or B.dtype != torch.complex128
                              or out.dtype != torch.complex128):
        raise ValueError("fp64 contraction requires complex128 operands")
    sa = dict(zip(la, A.stride()))
    sb = dict(zip(lb, B.stride()))
    so = dict(zip(lo, out.stride()))
    key = (tuple(dim[l] for l in m_l), tuple(dim[l] for l in k_l), tuple(dim[l] for l in n_l),
           tuple(sa[l] for l in m_l), tuple(sa[l] for l in k_l), tuple(sb[l] for l in k_l),
           tuple(sb[l] for l in n_l), tuple(so[l] for l in m_l), tuple(so[l] for l in n_l))
    mdims, kdims, ndims = key[0], key[1], key[2]

    def build():
        return [_offsets(mdims, key[3]), _offsets(kdims, key[4]), _offsets(kdims, key[5]),
                _offsets(ndims, key[6]), _offsets(mdims, key[7]), _offsets(ndims, key[8])]

    _, ptrs = _CACHE.get(key, build, A.device)
    M = math.prod(mdims)
    N = math.prod(ndims)
    K = math.prod(kdims) if kdims else 1
    if not kdims:
        # outer product: a single K index with zero offsets
        pass
    d = GemmDesc()
    d.M, d.N, d.K = M, N, K
    d.A = A.data_ptr()
    d.a_off_m, d.a_off_k = ptrs[0], ptrs[1]
    d.a_dtype = _DTYPE_CODE[A.dtype]
    d.a_conj = int(conj_a)
    d.a_kmajor = int(_inner_is_unit(kdims, key[4]) and not _inner_is_unit(mdims, key[3]))
    d.B = B.data_ptr()
    d.b_off_k, d.b_off_n = ptrs[2], ptrs[3]
    d.b_dtype = _DTYPE_CODE[B.dtype]
    d.b_conj = int(conj_b)
    d.b_kmajor = int(_inner_is_unit(kdims, key[5]) and not _inner_is_unit(ndims, key[6]))
    d.C = out.data_ptr()
    d.c_off_m, d.c_off_n = ptrs[4], ptrs[5]
    d.c_dtype = _DTYPE_CODE[out.dtype]
    d.beta_one = int(accumulate)
    alpha = complex(alpha)
    d.alpha_re, d.alpha_im = alpha.real, alpha.imag
    d.compute = _COMPUTE[precision]
    ws = _splitk_workspace(A.device)
    d.ws, d.ws_bytes = ws.data_ptr(), ws.numel() * ws.element_size()
    lib = _lib.load()
    prof = PROFILER
    if prof.active:
        fl = 8.0 * M * N * K
        prof.flops += fl
        prof.launches += 1
        if fl >= prof.min_flops:
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            _lib.check(lib.ptn_gemm_gather(ctypes.c_void_p(stream_handle()), ctypes.byref(d)),
                       "ptn_gemm_gather")
            ev1.record()
            prof.events.setdefault((M, N, K, precision), []).append((ev0, ev1))
            return out
    _lib.check(lib.ptn_gemm_gather(ctypes.c_void_p(stream_handle()), ctypes.byref(d)),
               "ptn_gemm_gather")
    return out


def permute_copy(src: torch.Tensor, perm: Sequence[int], out: Optional[torch.Tensor] = None,
                 conj: bool = False, dtype: Optional[torch.dtype] = None) -> torch.Tensor:
    """out = op(src).permute(perm) materialised through the ptn_permute kernel."""
    require_device(src)
    perm = list(perm)
    dims = [src.shape[p] for p in perm]
    if out is None:
        out = torch.empty(dims, dtype=dtype or src.dtype, device=src.device)
    require_device(out)
    if list(out.shape) != dims:
        raise ValueError("permute_copy: out has the wrong shape")
    return strided_copy(dims, src, [src.stride(p) for p in perm], out, list(out.stride()),
                        conj)


def strided_copy(dims, src, src_strides, dst, dst_strides, conj=False):
    rank = len(dims)
    if rank > 8:
        raise ValueError("rank > 8 not supported")
    arr = (ctypes.c_int64 * max(rank, 1))
    lib = _lib.load()
    _lib.check(lib.ptn_permute(ctypes.c_void_p(stream_handle()), rank, arr(*dims),
                               ctypes.c_void_p(src.data_ptr()), arr(*src_strides),
                               _DTYPE_CODE[src.dtype], ctypes.c_void_p(dst.data_ptr()),
                               arr(*dst_strides), _DTYPE_CODE[dst.dtype], int(conj)),
               "ptn_permute")
    return dst


def fill_(t: torch.Tensor, value: complex = 0.0) -> torch.Tensor:
    require_device(t)
    if not t.is_contiguous():
        raise ValueError("fill_ needs a contiguous tensor")
    v = complex(value)
    _lib.check(_lib.load().ptn_fill(ctypes.c_void_p(stream_handle()),
                                    ctypes.c_void_p(t.data_ptr()), t.numel(),
                                    _DTYPE_CODE[t.dtype], v.real, v.imag), "ptn_fill")
    return t


def zeros(shape, dtype=torch.complex128, device=None) -> torch.Tensor:
    t = torch.empty(tuple(shape), dtype=dtype, device=device)
    if t.numel():
        fill_(t, 0.0)
    return t


# --------------------------------------------------------------------------------------
# network contraction
# --------------------------------------------------------------------------------------
@dataclass
class NetPlan:
    steps: List[Tuple[int, int, Tuple[Label, ...]]]   # (slot_a, slot_b, result labels)
    flops: float                                       # 8 * MACs (4M counting)
    peak: int                                          # largest intermediate (elements)


_PLANS: Dict[tuple, NetPlan] = {}


def _merge_cost(l1: frozenset, l2: frozenset, dims: Dict[Label, int]) -> int:
    return math.prod(dims[l] for l in (l1 | l2))


def plan_network(labels: Sequence[Sequence[Label]], dims: Dict[Label, int],
                 out_labels: Sequence[Label]) -> NetPlan:
    """Optimal pairwise order (min total flops) by DP over subsets."""
    n = len(labels)
    key = (tuple(tuple(l) for l in labels), tuple(sorted((repr(k), v) for k, v in dims.items())),
           tuple(out_labels))
    hit = _PLANS.get(key)
    if hit is not None:
        return hit
    sets = [frozenset(l) for l in labels]
    out_set = frozenset(out_labels)
    count: Dict[Label, int] = {}
    for s in sets:
        for l in s:
            count[l] = count.get(l, 0) + 1

    def legs(mask: int) -> frozenset:
        inside: Dict[Label, int] = {}
        for i in range(n):
            if mask >> i & 1:
                for l in sets[i]:
                    inside[l] = inside.get(l, 0) + 1
        return frozenset(l for l, c in inside.items() if c < count[l] or l in out_set)

    full = (1 << n) - 1
    leg_of = {m: legs(m) for m in range(1, full + 1)}
    best: Dict[int, Tuple[float, int, Optional[tuple]]] = {}
    for i in range(n):
        best[1 << i] = (0.0, 0, None)
    for mask in range(1, full + 1):
        if mask & (mask - 1) == 0:
            continue
        cand = None
        sub = (mask - 1) & mask
        while sub:
            other = mask ^ sub
            if sub < other and sub in best and other in best:
                shared = leg_of[sub] & leg_of[other]
                if shared or mask == full:
                    c = 8.0 * _merge_cost(leg_of[sub], leg_of[other], dims)
                    tot = best[sub][0] + best[other][0] + c
                    size = math.prod(dims[l] for l in leg_of[mask])
                    peak = max(best[sub][1], best[other][1], size)
                    if cand is None or (tot, peak) < (cand[0], cand[1]):
                        cand = (tot, peak, (sub, other))
            sub = (sub - 1) & mask
        if cand is not None:
            best[mask] = cand
    if full not in best:
        raise ValueError("network is not connected")
    steps: List[Tuple[int, int, Tuple[Label, ...]]] = []
    slot_of: Dict[int, int] = {1 << i: i for i in range(n)}
    next_slot = [n]

    def emit(mask: int) -> int:
        if mask in slot_of:
            return slot_of[mask]
        a, b = best[mask][2]
        sa, sb = emit(a), emit(b)
        slot = next_slot[0]
        next_slot[0] += 1
        res = tuple(out_labels) if mask == full else None
        steps.append((sa, sb, res))
        slot_of[mask] = slot
        return slot

    emit(full)
    plan = NetPlan(steps, best[full][0], best[full][1])
    _PLANS[key] = plan
    return plan


def contract_network(tensors: Sequence[torch.Tensor], labels: Sequence[Sequence[Label]],
                     out_labels: Sequence[Label], *, conj: Optional[Sequence[bool]] = None,
                     precision: str = FP64, out: Optional[torch.Tensor] = None,
                     out_dtype: torch.dtype = torch.complex128) -> torch.Tensor:
    """Contract a whole network; ``conj[i]`` conjugates tensor i on load."""
    n = len(tensors)
    conj = list(conj) if conj is not None else [False] * n
    dims: Dict[Label, int] = {}
    for t, l in zip(tensors, labels):
        if len(l) != t.dim():
            raise ValueError("label count does not match tensor rank")
        for x, d in zip(l, t.shape):
            if dims.setdefault(x, d) != d:
                raise ValueError(f"dimension mismatch on label {x!r}")
    if n == 1:
        perm = [list(labels[0]).index(l) for l in out_labels]
        return permute_copy(tensors[0], perm, out=out, conj=conj[0], dtype=out_dtype)
    plan = plan_network(labels, dims, out_labels)
    slots: List[Optional[Tuple[torch.Tensor, Tuple[Label, ...], bool]]] = [
        (t, tuple(l), c) for t, l, c in zip(tensors, labels, conj)]
    inter = storage_dtype(precision)
    for idx, (sa, sb, res) in enumerate(plan.steps):
        ta, la_, ca = slots[sa]
        tb, lb_, cb = slots[sb]
        last = idx == len(plan.steps) - 1
        if res is None:
            res = tuple([l for l in la_ if l not in lb_] + [l for l in lb_ if l not in la_])
        # a label shared by A and B but also needed outside is not contracted:
        # networks here never have such hyper-edges.
        r = contract(ta, la_, tb, lb_, res, conj_a=ca, conj_b=cb, precision=precision,
                     out=out if last else None,
                     out_dtype=out_dtype if last else inter)
        slots.append((r, res, False))
        slots[sa] = None
        slots[sb] = None
    return slots[-1][0]
