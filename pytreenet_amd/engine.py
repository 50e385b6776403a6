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
from ._lib import GemmDesc, PtnBackendError, StridedDesc

Label = Hashable

FP64 = "fp64"
FP32 = "fp32"
# 3M bf16 products with FP32 accumulation (the lowest level of RelaxedNetwork)
BF16 = "bf16"
# 3M products of three-piece bf16 splits (six piece products, FP32 accumulation):
# FP32 accuracy at the bf16 MFMA rate (a low level of RelaxedNetwork)
BF16X6 = "bf16x6"
# FP64 environments and Krylov applications, each application run in 3M FP32 when
# its contribution to the Krylov result is small enough (RelaxedNetwork)
RELAXED = "relaxed"
_COMPUTE = {FP64: _lib.PTN_COMPUTE_F64_4M, FP32: _lib.PTN_COMPUTE_F32_3M,
            BF16: _lib.PTN_COMPUTE_BF16_3M, BF16X6: _lib.PTN_COMPUTE_BF16X6_3M}
_DTYPE_CODE = {torch.complex128: _lib.PTN_C128, torch.complex64: _lib.PTN_C64}


def storage_dtype(precision: str) -> torch.dtype:
    return torch.complex128 if precision == FP64 else torch.complex64


_raw_stream = torch._C._cuda_getCurrentRawStream
_cur_device = torch._C._cuda_getDevice


def stream_handle() -> int:
    """HIP stream the library launches on: torch's current stream on the current
    device, read at every call (two C calls, no Python stream object), so work
    issued inside a ``torch.cuda.stream(...)`` block is ordered with the
    allocations torch makes on that stream."""
    return _raw_stream(_cur_device())


def refresh_stream() -> int:
    """Kept for callers that refreshed a cached handle; the handle is never cached."""
    return stream_handle()


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


class _TableSpec:
    """An offset table given by its modes, built where it is cheapest."""
    __slots__ = ("dims", "strides", "n")

    def __init__(self, dims: Sequence[int], strides: Sequence[int]):
        self.dims, self.strides = tuple(dims), tuple(strides)
        self.n = math.prod(self.dims)


def _device_offsets(dims: Sequence[int], strides: Sequence[int], out: torch.Tensor) -> None:
    """The offset table of the modes written into the int64 device vector ``out``."""
    dev = out.device
    off = torch.zeros(1, dtype=torch.int64, device=dev)
    for i, (d, s) in enumerate(zip(dims, strides)):
        step = torch.arange(d, dtype=torch.int64, device=dev) * s
        if i == len(dims) - 1:
            torch.add(off[:, None], step[None, :], out=out.view(-1, d))
            return
        off = (off[:, None] + step[None, :]).reshape(-1)
    out.copy_(off)


class _OffsetCache:
    """Device-resident offset tables keyed by contraction signature (LRU).

    A builder returns host tables (numpy) or ``_TableSpec``s.  Specs of more than
    DEVICE_BUILD_MIN entries are built on the device (a few elementwise kernels)
    instead of on the host and copied: the tables of the large two-site networks
    run to millions of entries, and their signatures change whenever truncation
    changes a bond dimension."""

    DEVICE_BUILD_MIN = 1 << 16

    def __init__(self, capacity: int = 4096):
        self.cap = capacity
        self.d: "OrderedDict[tuple, tuple]" = OrderedDict()

    def get(self, key, builder, device):
        hit = self.d.get(key)
        if hit is not None:
            self.d.move_to_end(key)
            return hit
        tables = builder()
        on_dev = [isinstance(t, _TableSpec) and t.n > self.DEVICE_BUILD_MIN for t in tables]
        host = [t if not isinstance(t, _TableSpec) else
                (None if d else _offsets(t.dims, t.strides)) for t, d in zip(tables, on_dev)]
        sizes = [t.n if d else len(h) for t, d, h in zip(tables, on_dev, host)]
        # host tables first (one copy), device-built ones after them
        order = [i for i in range(len(tables)) if not on_dev[i]] + \
                [i for i in range(len(tables)) if on_dev[i]]
        pos, start = 0, [0] * len(tables)
        for i in order:
            start[i] = pos
            pos += max(sizes[i], 1)
        dev = torch.empty(pos, dtype=torch.int64, device=device)
        nhost = sum(max(sizes[i], 1) for i in order if not on_dev[i])
        if nhost:
            cat = np.concatenate([host[i] if sizes[i] else np.zeros(1, np.int64)
                                  for i in order if not on_dev[i]])
            dev[:nhost].copy_(torch.from_numpy(cat))
        for i in order:
            if on_dev[i]:
                _device_offsets(tables[i].dims, tables[i].strides,
                                dev[start[i]:start[i] + sizes[i]])
        ptrs = [dev.data_ptr() + start[i] * 8 for i in range(len(tables))]
        val = (dev, ptrs)
        self.d[key] = val
        if len(self.d) > self.cap:
            self.d.popitem(last=False)
        return val


_CACHE = _OffsetCache()


class LaunchProfiler:
    """Executed-flop counters and HIP-event timings of contraction launches.

    Events are recorded on the stream the kernels are launched on (torch's
    current stream) around every launch that runs the gather-GEMM's LARGE
    instantiation (``gemm_kernel_class`` "large": a tile grid >= LARGE_TILES,
    never split-K), which profilers report as a kernel of its own, and around
    every strided launch with such a grid (a rocBLAS GEMM): the dominant H_eff /
    environment GEMMs.
    """

    LARGE_TILES = 2048   # tile-grid size from which the library uses its LARGE instantiation

    def __init__(self):
        self.active = False
        self.phase = "default"
        self.reset()

    def reset(self):
        self.flops = 0.0
        self.launches = 0
        self.events: Dict[tuple, list] = {}
        # every LARGE gather launch in issue order: (M, N, K, precision, beta_one,
        # c_mmajor, phase), to match per-dispatch profiler counters to shapes
        self.order: List[tuple] = []

    def start(self, phase: str = "default", reset: bool = True):
        if reset:
            self.reset()
        self.phase = phase
        self.active = True

    def stop(self):
        self.active = False

    def summary(self, phase: Optional[str] = None, compute: Optional[str] = None,
                backend: Optional[str] = None) -> List[dict]:
        """Per-shape rows for LARGE launches, optionally of one phase / precision /
        backend ("gather": ptn_gemm_gather, "blas": ptn_gemm_strided)."""
        torch.cuda.synchronize()
        agg: Dict[tuple, list] = {}
        for key, evs in self.events.items():
            if (phase is not None and key[4] != phase) or (compute is not None and
                                                           key[3] != compute) or \
                    (backend is not None and key[5] != backend):
                continue
            agg.setdefault(key[:4] + (key[5],), []).extend(a.elapsed_time(b) for a, b in evs)
        rows = []
        for key, ms in agg.items():
            rows.append({"M": key[0], "N": key[1], "K": key[2], "compute": key[3],
                         "backend": key[4],
                         "flops_per_launch": 8.0 * key[0] * key[1] * key[2],
                         "launches": len(ms), "total_ms": float(sum(ms)),
                         "avg_ms": float(sum(ms) / len(ms))})
        rows.sort(key=lambda r: -r["total_ms"])
        return rows


PROFILER = LaunchProfiler()


def clear_caches():
    _FAST.clear()
    _CACHE.d.clear()
    _PLANS.clear()
    _AFFINE.clear()
    _NET_CACHE.clear()
    _NET_CACHE_BYTES[0] = 0
    _NET_TOO_BIG.clear()


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


def _affine_modes(dims, strides) -> Optional[int]:
    """Stride of the flattened index space if its modes are nested (one stride), else None."""
    nz = [(d, s) for d, s in zip(dims, strides) if d > 1]
    if not nz:
        return 1
    for (_, s0), (d1, s1) in zip(nz[:-1], nz[1:]):
        if s0 != s1 * d1:
            return None
    return nz[-1][1] if nz[-1][1] > 0 else None


def _affine_table(t: np.ndarray) -> Optional[Tuple[int, int]]:
    """(base, stride) if the offset table is an arithmetic progression."""
    t = np.asarray(t, dtype=np.int64)
    if len(t) == 1:
        return int(t[0]), 1
    st = int(t[1] - t[0])
    if st <= 0 or not np.array_equal(t, t[0] + st * np.arange(len(t), dtype=np.int64)):
        return None
    return int(t[0]), st


# --------------------------------------------------------------------------------------
# pairwise contraction
# --------------------------------------------------------------------------------------
def contract(A: torch.Tensor, la: Sequence[Label], B: torch.Tensor, lb: Sequence[Label],
             out_labels: Sequence[Label], *, conj_a: bool = False, conj_b: bool = False,
             precision: str = FP64, out: Optional[torch.Tensor] = None,
             out_dtype: Optional[torch.dtype] = None, accumulate: bool = False,
             alpha: complex = 1.0, lower: bool = False) -> torch.Tensor:
    """sum over shared labels of opA(A) * opB(B), laid out as ``out_labels``.

    ``lower``: only the part on and below the diagonal of a square result (first
    output legs x last, equal extents) will be read, e.g. a Gram matrix: the
    64 x 64 tiles strictly above it may be left unwritten."""
    require_device(A, B)
    la, lb, lo = tuple(la), tuple(lb), tuple(out_labels)
    fk = None
    if out is None and not accumulate:
        # a call that allocates its output: the launch of an earlier call with the
        # same labels, layouts, dtypes and options is replayed from its descriptor
        fk = (la, lb, lo, A.shape, A.stride(), A.dtype, B.shape, B.stride(), B.dtype,
              conj_a, conj_b, precision, out_dtype, complex(alpha), A.device, lower,
              USE_BLAS, BLAS_FP64, _BLAS_OFF)
        hit = _FAST.get(fk)
        if hit is not None:
            return _fast_launch(hit, A, B, precision)
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
        return [_TableSpec(mdims, key[3]), _TableSpec(kdims, key[4]),
                _TableSpec(kdims, key[5]), _TableSpec(ndims, key[6]),
                _TableSpec(mdims, key[7]), _TableSpec(ndims, key[8])]

    tabs, ptrs = _CACHE.get(key, build, A.device)
    M = math.prod(mdims)
    N = math.prod(ndims)
    K = math.prod(kdims) if kdims else 1
    ka = _inner_is_unit(kdims, key[4]) and not _inner_is_unit(mdims, key[3])
    kb = _inner_is_unit(kdims, key[5]) and not _inner_is_unit(ndims, key[6])
    cm, cn = _inner_is_unit(mdims, key[7]), _inner_is_unit(ndims, key[8])
    aff = _AFFINE.get(key)
    if aff is None:
        if len(_AFFINE) > 65536:
            _AFFINE.clear()
        st = [_affine_modes(d, sx) for d, sx in ((mdims, key[3]), (kdims, key[4]),
                                                (kdims, key[5]), (ndims, key[6]),
                                                (mdims, key[7]), (ndims, key[8]))]
        aff = _AFFINE[key] = (False if any(x is None for x in st)
                              else tuple((0, x) for x in st))
    used = _launch(A, B, out, ptrs, M, N, K, ka, kb, conj_a, conj_b, precision, accumulate,
                   alpha, aff or None, tabs, (cm and not cn, cn and not cm),
                   lower and M == N and list(m_l) == list(lo[:len(m_l)]))
    if fk is not None and used is not None:
        if len(_FAST) >= FAST_CAP:
            _FAST.clear()
        d, swapped, Ml, Nl = used
        _FAST[fk] = (oshape, out.dtype, GemmDesc.from_buffer_copy(d), swapped, Ml, Nl, K, tabs)
    return out


# launch descriptors of earlier allocating contract() calls, by labels, layouts,
# dtypes and options (their offset tables are held here too)
_FAST: Dict[tuple, tuple] = {}
FAST_CAP = 4096


def _fast_launch(hit, A: torch.Tensor, B: torch.Tensor, precision: str) -> torch.Tensor:
    oshape, odt, proto, swapped, M, N, K, tabs = hit
    out = torch.empty(oshape, dtype=odt, device=A.device)
    X, Y = (B, A) if swapped else (A, B)
    d = GemmDesc.from_buffer_copy(proto)
    d.A, d.B, d.C = X.data_ptr(), Y.data_ptr(), out.data_ptr()
    if _RECORDER is not None:
        _RECORDER.ops.append(("gemm", d, X, Y, out, M, N, K, precision, tabs))
    _run_gemm(d, M, N, K, precision)
    return out


_AFFINE: Dict[tuple, object] = {}


class _Recorder:
    """Collects the launches of one network contraction for later replay."""

    def __init__(self):
        self.ops: list = []


_RECORDER: Optional[_Recorder] = None


# strided GEMMs go to rocBLAS when the tile grid fills the chip; the gather-GEMM
# keeps the gathered, short and split-K (long K, few tiles) cases.  FP64 stays on
# the gather-GEMM unless BLAS_FP64 is set: its three-product kernel outruns
# rocBLAS zgemm on the strided shapes of the sweep (78.5 vs 68.8 TFLOP/s
# algorithmic at M=131072, N=1280, K=256)
BLAS_MIN_TILES = 256
BLAS_MIN_DIM = 64
BLAS_MAX_K_PER_TILE = 1 << 14
USE_BLAS = True
BLAS_FP64 = False


_BLAS_OFF = 0   # > 0 while recording networks that are replayed as grouped launches


def _blas_desc(A, B, C, M, N, K, conj_a, conj_b, precision, accumulate, alpha, affine):
    if not USE_BLAS or _BLAS_OFF or K < 1 or min(M, N) < BLAS_MIN_DIM:
        return None
    if precision == FP64 and not BLAS_FP64:
        return None
    tiles = -(-M // 64) * -(-N // 64)
    if tiles < BLAS_MIN_TILES or K > BLAS_MAX_K_PER_TILE:
        return None
    dt = A.dtype
    if B.dtype != dt or C.dtype != dt:
        return None
    if not ((precision == FP64 and dt == torch.complex128) or
            (precision == FP32 and dt == torch.complex64)):
        return None
    (am0, asm), (ak0, ask), (bk0, bsk), (bn0, bsn), (cm0, csm), (cn0, csn) = affine
    sd = StridedDesc()
    sd.M, sd.N, sd.K = M, N, K
    es = A.element_size()
    offs = ((am0 + ak0) * es, (bk0 + bn0) * es, (cm0 + cn0) * es)
    sd.A, sd.a_sm, sd.a_sk, sd.a_conj = A.data_ptr() + offs[0], asm, ask, int(conj_a)
    sd.B, sd.b_sk, sd.b_sn, sd.b_conj = B.data_ptr() + offs[1], bsk, bsn, int(conj_b)
    sd.C, sd.c_sm, sd.c_sn = C.data_ptr() + offs[2], csm, csn
    sd.dtype = _DTYPE_CODE[dt]
    sd.beta_one = int(accumulate)
    alpha = complex(alpha)
    sd.alpha_re, sd.alpha_im = alpha.real, alpha.imag
    if not _lib.load().ptn_gemm_strided_ok(ctypes.byref(sd)):
        return None
    return sd, offs


def _run_blas(sd: StridedDesc, M: int, N: int, K: int, precision: str):
    lib = _lib.load()
    prof = PROFILER
    if prof.active:
        prof.flops += 8.0 * M * N * K
        prof.launches += 1
        if -(-M // 64) * -(-N // 64) >= prof.LARGE_TILES:
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            _lib.check(lib.ptn_gemm_strided(ctypes.c_void_p(stream_handle()), ctypes.byref(sd)),
                       "ptn_gemm_strided")
            ev1.record()
            prof.events.setdefault((M, N, K, precision, prof.phase, "blas"), []).append(
                (ev0, ev1))
            return
    _lib.check(lib.ptn_gemm_strided(ctypes.c_void_p(stream_handle()), ctypes.byref(sd)),
               "ptn_gemm_strided")


def _launch(A, B, C, ptrs, M, N, K, kmaj_a, kmaj_b, conj_a, conj_b, precision, accumulate,
            alpha, affine=None, tabs=None, c_major=(False, False), lower=False):
    """One contraction launch from device offset tables
    ptrs = (a_off_m, a_off_k, b_off_k, b_off_n, c_off_m, c_off_n).

    ``tabs`` is the device tensor holding those tables: a recorded launch keeps
    a reference to it, so the tables outlive the offset-table cache's LRU
    eviction for as long as the recording can be replayed.

    ``affine`` holds (base, stride) of each table when all six are arithmetic
    progressions; such a launch with a full tile grid runs as one strided
    rocBLAS GEMM.  Otherwise ptn_gemm_gather runs, with the operand of the
    longer free dimension on the M side (64-row tiles waste less on a short N
    than on a short M).  ``c_major`` tells whether consecutive m (n) are
    contiguous in C: a layout hint for the store order of small products.
    ``lower``: only C's tiles with m >= n are needed (a square Gram product).
    Returns (descriptor, swapped, M, N) of a gather launch (``swapped``: A and B
    exchanged their roles), None for a strided library GEMM."""
    if affine is not None and not lower:
        hit = _blas_desc(A, B, C, M, N, K, conj_a, conj_b, precision, accumulate, alpha, affine)
        if hit is not None:
            sd, offs = hit
            if _RECORDER is not None:
                _RECORDER.ops.append(("blas", sd, A, B, C, M, N, K, precision, offs))
            _run_blas(sd, M, N, K, precision)
            return None
    swapped = False
    if M < N and N >= 64:
        swapped = True
        A, B = B, A
        M, N = N, M
        kmaj_a, kmaj_b = kmaj_b, kmaj_a
        conj_a, conj_b = conj_b, conj_a
        ptrs = (ptrs[3], ptrs[2], ptrs[1], ptrs[0], ptrs[5], ptrs[4])
        c_major = (c_major[1], c_major[0])
    d = GemmDesc()
    d.c_mmajor = int(c_major[0])
    d.M, d.N, d.K = M, N, K
    d.A = A.data_ptr()
    d.a_off_m, d.a_off_k = ptrs[0], ptrs[1]
    d.a_dtype = _DTYPE_CODE[A.dtype]
    d.a_conj = int(conj_a)
    d.a_kmajor = int(kmaj_a)
    d.B = B.data_ptr()
    d.b_off_k, d.b_off_n = ptrs[2], ptrs[3]
    d.b_dtype = _DTYPE_CODE[B.dtype]
    d.b_conj = int(conj_b)
    d.b_kmajor = int(kmaj_b)
    d.C = C.data_ptr()
    d.c_off_m, d.c_off_n = ptrs[4], ptrs[5]
    d.c_dtype = _DTYPE_CODE[C.dtype]
    d.beta_one = int(accumulate)
    alpha = complex(alpha)
    d.alpha_re, d.alpha_im = alpha.real, alpha.imag
    d.compute = _COMPUTE[precision]
    d.c_lower = int(lower)
    ws = _splitk_workspace(A.device)
    d.ws, d.ws_bytes = ws.data_ptr(), ws.numel() * ws.element_size()
    if _RECORDER is not None:
        _RECORDER.ops.append(("gemm", d, A, B, C, M, N, K, precision, tabs))
    _run_gemm(d, M, N, K, precision)
    return d, swapped, M, N


# the row-streaming kernel's limits and the rule's threshold (``ptn_gemm.hip``)
SKINNY_MAX_K, SKINNY_MAX_N, SKINNY_MIN_M = 16, 64, 8192


def gemm_kernel_class(M: int, N: int, K: int, precision: str, a_kmajor: bool = True,
                      c_mmajor: bool = False) -> str:
    """The kernel ``ptn_gemm_gather`` runs a launch of this shape on (its dispatch
    rule, given the descriptor's layout hints ``a_kmajor`` and ``c_mmajor``): "skinny"
    (K <= 16, N <= 64, M >= 8192, A and C m-contiguous: rows streamed with B in LDS),
    "wave" (M or N within one 16-block, or both within two, or, for
    K <= 256, N within three or five 16-blocks below a longer M, or M within three
    below a longer N; either compute mode), "large" (a tile grid of >= 2048 64 x 64 tiles: the LARGE instantiation,
    never split-K) or "small" (the same kernel on a smaller grid, split-K allowed)."""
    if M >= SKINNY_MIN_M and 1 <= K <= SKINNY_MAX_K and 1 <= N <= SKINNY_MAX_N and \
            not a_kmajor and c_mmajor:
        return "skinny"
    bm, bn = -(-M // 16), -(-N // 16)
    if (bm <= 2 and bn <= 2) or bm == 1 or bn == 1:
        return "wave"
    if K <= 256 and ((bn in (3, 5) and bm > bn) or (bm == 3 and bn > bm)):
        return "wave"
    return "large" if -(-M // 64) * -(-N // 64) >= PROFILER.LARGE_TILES else "small"


# ---- launch-configuration autotuner -------------------------------------------------
# With AUTOTUNE set, the first launch of every gather-GEMM signature below the LARGE
# grid (shape, compute mode, dtypes, operand majorness, store-order hint) times the
# candidate configurations of ``ptn_gemm_gather_config`` -- the library's own rule,
# the wave kernel with each block shape, the 64 x 64 kernel without and with K
# splits -- and keeps the fastest for every later launch of that signature
# (recorded network replays included).  Each candidate computes the same product,
# so the timing runs write the caller's output; the chosen configuration runs last.
# Accumulating launches (beta = 1) keep the library's rule.  LARGE launches run the
# direct kernel, which reaches the MFMA peak, and are not tuned.
AUTOTUNE = False
TUNE_REPS = 3
_WAVE_BLOCKS = ((1, 1), (2, 1), (1, 2), (2, 2), (4, 1), (1, 4), (1, 3), (3, 1), (1, 5))
_WAVE_DTYPES = {0, 7, 6, 3, 5}     # a*4 + b*2 + c over (complex128 = 0, complex64 = 1)
_TUNED: Dict[tuple, Tuple[int, int, int, int]] = {}
TUNE_LOG: List[tuple] = []         # (signature, chosen config, {config: ms})


def _tune_key(d: GemmDesc, M: int, N: int, K: int, precision: str) -> tuple:
    return (M, N, K, precision, d.a_dtype, d.b_dtype, d.c_dtype, d.a_kmajor, d.b_kmajor,
            d.c_mmajor, d.a_conj, d.b_conj)


def tune_candidates(d: GemmDesc, M: int, N: int, K: int) -> List[Tuple[int, int, int, int]]:
    """(kind, bmb, bnb, ksplit) configurations worth timing for one launch."""
    cands = [(_lib.PTN_GEMM_RULE, 0, 0, 0)]
    if d.a_dtype * 4 + d.b_dtype * 2 + d.c_dtype in _WAVE_DTYPES:
        if 1 <= K <= SKINNY_MAX_K and 1 <= N <= SKINNY_MAX_N:
            cands.append((_lib.PTN_GEMM_SKINNY, 0, 0, 0))
        for bm, bn in _WAVE_BLOCKS:
            if -(-M // (16 * bm)) * -(-N // (16 * bn)) <= 1 << 15:
                cands.append((_lib.PTN_GEMM_WAVE, bm, bn, 0))
    cands.append((_lib.PTN_GEMM_TILE, 0, 0, 1))
    for ks in (2, 4, 8, 16, 32):
        if K // ks >= 64 and d.ws and d.ws_bytes >= 16 * ks * M * N:
            cands.append((_lib.PTN_GEMM_TILE, 0, 0, ks))
    return cands


def _launch_config(lib, d: GemmDesc, cfg) -> None:
    _lib.check(lib.ptn_gemm_gather_config(ctypes.c_void_p(stream_handle()), ctypes.byref(d),
                                          *cfg), "ptn_gemm_gather_config")


def _tuned_launch(lib, d: GemmDesc, M: int, N: int, K: int, precision: str) -> None:
    key = _tune_key(d, M, N, K, precision)
    cfg = _TUNED.get(key)
    if cfg is None:
        if d.beta_one:
            cfg = _TUNED[key] = (_lib.PTN_GEMM_RULE, 0, 0, 0)
        else:
            times = {}
            for c in tune_candidates(d, M, N, K):
                _launch_config(lib, d, c)          # warm-up (and first-touch of tables)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(TUNE_REPS):
                    _launch_config(lib, d, c)
                e1.record()
                e1.synchronize()
                times[c] = e0.elapsed_time(e1) / TUNE_REPS
            cfg = _TUNED[key] = min(times, key=times.get)
            TUNE_LOG.append((key, cfg, times))
    _launch_config(lib, d, cfg)


def _run_gemm(d: GemmDesc, M: int, N: int, K: int, precision: str):
    lib = _lib.load()
    cls = gemm_kernel_class(M, N, K, precision, bool(d.a_kmajor), bool(d.c_mmajor)) \
        if AUTOTUNE or PROFILER.active else None
    if AUTOTUNE and cls != "large":
        if PROFILER.active:
            PROFILER.flops += 8.0 * M * N * K
            PROFILER.launches += 1
        _tuned_launch(lib, d, M, N, K, precision)
        return
    prof = PROFILER
    if prof.active:
        fl = 8.0 * M * N * K
        prof.flops += fl
        prof.launches += 1
        if cls == "large":
            prof.order.append((M, N, K, precision, int(d.beta_one), int(d.c_mmajor),
                               prof.phase))
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            _lib.check(lib.ptn_gemm_gather(ctypes.c_void_p(stream_handle()), ctypes.byref(d)),
                       "ptn_gemm_gather")
            ev1.record()
            prof.events.setdefault((M, N, K, precision, prof.phase, "gather"), []).append(
                (ev0, ev1))
            return
    _lib.check(lib.ptn_gemm_gather(ctypes.c_void_p(stream_handle()), ctypes.byref(d)),
               "ptn_gemm_gather")


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
    if _RECORDER is not None:
        _RECORDER.ops.append(("fill", t, v))
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
class BlockSparse:
    """A tensor whose ``block_axes`` only take the index tuples listed in ``tuples``.

    The data are stored compactly as ``data[b, dense...]`` (block index first,
    the remaining axes in their logical order).  A TTNO node tensor is of this
    kind: only a few combinations of its virtual (bond-state) indices carry a
    non-zero physical operator, e.g. 12 of the 125 at the root of a ternary XYZ
    tree.  Contractions skip the absent blocks exactly.
    """

    def __init__(self, data: torch.Tensor, block_axes: Sequence[int], tuples: np.ndarray,
                 shape: Sequence[int]):
        self.data = data
        self.block_axes = tuple(block_axes)
        self.tuples = np.ascontiguousarray(np.asarray(tuples, dtype=np.int64).reshape(
            len(tuples), len(self.block_axes)))
        self.shape = tuple(shape)
        self.tkey = hash(self.tuples.tobytes()) ^ hash(self.tuples.shape)

    def dim(self) -> int:
        return len(self.shape)

    @property
    def dtype(self):
        return self.data.dtype

    @staticmethod
    def from_dense(t: torch.Tensor, block_axes: Sequence[int]) -> "BlockSparse":
        block_axes = list(block_axes)
        host = t.detach().cpu().numpy()
        rest = [a for a in range(host.ndim) if a not in block_axes]
        moved = np.transpose(host, block_axes + rest)
        bshape = moved.shape[:len(block_axes)]
        flat = moved.reshape(int(np.prod(bshape)) if bshape else 1, -1)
        nz = np.nonzero(np.any(flat != 0, axis=1))[0]
        if len(nz) == 0:
            nz = np.array([0])
        tuples = np.stack(np.unravel_index(nz, bshape), axis=1) if bshape else \
            np.zeros((1, 0), dtype=np.int64)
        data = torch.as_tensor(np.ascontiguousarray(
            flat[nz].reshape((len(nz),) + moved.shape[len(block_axes):]))).to(t.device)
        return BlockSparse(data.to(t.dtype), block_axes, tuples, host.shape)


class _Blk:
    """Block-sparse intermediate of a network contraction."""

    def __init__(self, data, dense_labels, blabels, tuples, bdims, conj=False, tkey=None,
                 source=None):
        self.data = data
        self.source = source      # the BlockSparse it was made from (None: intermediate)
        self.dense_labels = tuple(dense_labels)
        self.blabels = tuple(blabels)
        self.tuples = tuples
        self.bdims = dict(bdims)
        self.conj = conj
        self.tkey = tkey if tkey is not None else hash(tuples.tobytes()) ^ hash(tuples.shape)


_BLK_CACHE: "OrderedDict[tuple, list]" = OrderedDict()
_GROUPS: Dict[tuple, tuple] = {}


def _contract_block(X: _Blk, Y: torch.Tensor, ly: Sequence[Label], conj_y: bool,
                    res_labels: Optional[Sequence[Label]], precision: str,
                    out: Optional[torch.Tensor], out_dtype: torch.dtype):
    """Block-sparse X times dense Y; one gather-GEMM per output block group."""
    ly = tuple(ly)
    dims: Dict[Label, int] = dict(X.bdims)
    dims.update(zip(X.dense_labels, X.data.shape[1:]))
    for l, d in zip(ly, Y.shape):
        if dims.setdefault(l, d) != d:
            raise ValueError(f"dimension mismatch on label {l!r}")
    k_blk = [l for l in X.blabels if l in ly]
    m_blk = [l for l in X.blabels if l not in ly]
    k_d = [l for l in X.dense_labels if l in ly]
    m_d = [l for l in X.dense_labels if l not in ly]
    n_l = [l for l in ly if l not in X.dense_labels and l not in X.blabels]
    cols_k = [X.blabels.index(l) for l in k_blk]
    cols_m = [X.blabels.index(l) for l in m_blk]
    gk = (X.tkey, tuple(cols_m))
    hitg = _GROUPS.get(gk)
    if hitg is None:
        if cols_m:
            G, inv = np.unique(X.tuples[:, cols_m], axis=0, return_inverse=True)
            inv = inv.reshape(-1)
        else:
            G, inv = np.zeros((1, 0), dtype=np.int64), np.zeros(len(X.tuples), dtype=np.int64)
        _GROUPS[gk] = hitg = (G, inv)
    G, inv = hitg
    dev = X.data.device
    block_out = bool(m_blk) and res_labels is None
    if block_out:
        odt = out_dtype
        res = tuple(m_d + n_l)
        data = torch.empty((len(G),) + tuple(dims[l] for l in res), dtype=odt, device=dev)
        o_strides = dict(zip(res, data.stride()[1:]))
        o_blk_stride = data.stride(0)
        target = data
    else:
        res = tuple(res_labels) if res_labels is not None else tuple(m_d + n_l)
        if sorted(map(repr, res)) != sorted(map(repr, m_d + m_blk + n_l)):
            raise ValueError(f"output labels {res} do not match {m_d + m_blk + n_l}")
        if out is None:
            out = torch.empty(tuple(dims[l] for l in res), dtype=out_dtype, device=dev)
        full = math.prod(dims[l] for l in m_blk) if m_blk else 1
        if len(G) < full:
            fill_(out, 0.0)
        o_strides = dict(zip(res, out.stride()))
        target = out
    sx = dict(zip(X.dense_labels, X.data.stride()[1:]))
    sx_blk = X.data.stride(0)
    sy = dict(zip(ly, Y.stride()))
    key = (X.data.shape, X.data.stride(), X.dense_labels, X.blabels, X.tkey, tuple(Y.shape),
           tuple(Y.stride()), ly, tuple(target.shape), tuple(target.stride()), res, block_out)
    nj = math.prod(dims[l] for l in m_d)
    ni = math.prod(dims[l] for l in k_d)
    R = math.prod(dims[l] for l in n_l)
    if (nj <= 4 and ni <= 4 and len(X.tuples) <= 128 and len(G) <= 128 and R >= 4096):
        _block_apply(X, Y, conj_y, dims, m_d, k_d, n_l, k_blk, cols_k, cols_m, G, inv, sy,
                     target, o_strides, block_out, o_blk_stride if block_out else 0, key)
    else:
        groups = _BLK_CACHE.get(key)
        if groups is None:
            a_m = _offsets([dims[l] for l in m_d], [sx[l] for l in m_d])
            kd_x = _offsets([dims[l] for l in k_d], [sx[l] for l in k_d])
            kd_y = _offsets([dims[l] for l in k_d], [sy[l] for l in k_d])
            b_n = _offsets([dims[l] for l in n_l], [sy[l] for l in n_l])
            c_m0 = _offsets([dims[l] for l in m_d], [o_strides[l] for l in m_d])
            c_n = _offsets([dims[l] for l in n_l], [o_strides[l] for l in n_l])
            ystr_k = np.array([sy[l] for l in k_blk], dtype=np.int64)
            ostr_m = np.array([o_strides[l] for l in m_blk], dtype=np.int64) if not block_out \
                else None
            host_groups = []
            for g in range(len(G)):
                P = np.nonzero(inv == g)[0].astype(np.int64)
                a_k = (P[:, None] * sx_blk + kd_x[None, :]).reshape(-1)
                yb = X.tuples[P][:, cols_k] @ ystr_k if cols_k else np.zeros(len(P), np.int64)
                b_k = (yb[:, None] + kd_y[None, :]).reshape(-1)
                shift = g * o_blk_stride if block_out else int(G[g] @ ostr_m) if cols_m else 0
                host_groups.append((a_m, a_k, b_k, b_n, c_m0 + shift, c_n))
            groups = []
            for tabs in host_groups:
                dtabs, ptrs = _CACHE.get(("blk", key, len(groups)), lambda t=tabs: list(t), dev)
                aff = [_affine_table(t) for t in tabs]
                # the group keeps its device tables alive (dtabs) beyond cache eviction
                groups.append((ptrs, len(tabs[0]), len(tabs[1]), len(tabs[3]),
                               None if any(a is None for a in aff) else tuple(aff), dtabs))
            _BLK_CACHE[key] = groups
            if len(_BLK_CACHE) > 4096:
                _BLK_CACHE.popitem(last=False)
        _group_gemms(X, Y, conj_y, dims, m_d, k_d, n_l, sx, sy, groups, precision, target,
                     o_strides)
    if block_out:
        return _Blk(data, res, m_blk, G, {l: X.bdims[l] for l in m_blk})
    return out


def _group_gemms(X, Y, conj_y, dims, m_d, k_d, n_l, sx, sy, groups, precision, target,
                 o_strides):
    ka = _inner_is_unit([dims[l] for l in k_d], [sx[l] for l in k_d]) and \
        not _inner_is_unit([dims[l] for l in m_d], [sx[l] for l in m_d])
    kb = _inner_is_unit([dims[l] for l in k_d], [sy[l] for l in k_d]) and \
        not _inner_is_unit([dims[l] for l in n_l], [sy[l] for l in n_l])
    cm = _inner_is_unit([dims[l] for l in m_d], [o_strides[l] for l in m_d])
    cn = _inner_is_unit([dims[l] for l in n_l], [o_strides[l] for l in n_l])
    for ptrs, M, K, N, aff, dtabs in groups:
        _launch(X.data, Y, target, ptrs, M, N, K, ka, kb, X.conj, conj_y, precision, False,
                1.0, aff, dtabs, (cm and not cn, cn and not cm))


_BAPPLY_CACHE: "OrderedDict[tuple, tuple]" = OrderedDict()
BLOCK_TILE = 256                 # rows per workgroup of ptn_block_apply_staged
BLOCK_STAGE_BYTES = 60 * 1024    # its LDS for the operator blocks and the staged inputs
USE_BLOCK_STAGED = True


def _block_apply(X, Y, conj_y, dims, m_d, k_d, n_l, k_blk, cols_k, cols_m, G, inv, sy, target,
                 o_strides, block_out, o_blk_stride, key):
    """The TTNO tensor (few small blocks) meets a large tensor: one pass of
    ``ptn_block_apply`` (each row read once, every live output block written)."""
    dev = X.data.device
    hit = _BAPPLY_CACHE.get(key)
    if hit is None:
        order = np.argsort(inv, kind="stable")
        counts = np.bincount(inv, minlength=len(G))
        g_begin = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        ystr_k = np.array([sy[l] for l in k_blk], dtype=np.int64)
        y_t = (X.tuples[order][:, cols_k] @ ystr_k) if cols_k else np.zeros(len(order), np.int64)
        if block_out:
            o_g = np.arange(len(G), dtype=np.int64) * o_blk_stride
        elif cols_m:
            m_blk = [X.blabels[c] for c in cols_m]
            o_g = G @ np.array([o_strides[l] for l in m_blk], dtype=np.int64)
        else:
            o_g = np.zeros(1, dtype=np.int64)
        dl = list(X.dense_labels)
        perm = [0] + [1 + dl.index(l) for l in m_d] + [1 + dl.index(l) for l in k_d]
        nj = math.prod(dims[l] for l in m_d)
        ni = math.prod(dims[l] for l in k_d)
        y_r = _offsets([dims[l] for l in n_l], [sy[l] for l in n_l])
        o_r = _offsets([dims[l] for l in n_l], [o_strides[l] for l in n_l])
        small = np.concatenate([y_t.astype(np.int64), o_g.astype(np.int64),
                                g_begin.astype(np.int64)])
        big = torch.from_numpy(np.concatenate([y_r, o_r])).to(dev)
        sm = torch.from_numpy(small).to(dev)
        gb32 = torch.from_numpy(g_begin).to(dev)
        y_i = _offsets([dims[l] for l in k_d], [sy[l] for l in k_d])
        o_j = _offsets([dims[l] for l in m_d], [o_strides[l] for l in m_d])
        # LDS staging of the inputs: tiles of BLOCK_TILE rows along the innermost
        # row dimension, when the inputs of a tile fit one bounded span
        staged = None
        d0 = dims[n_l[-1]] if n_l else 0
        s0 = sy[n_l[-1]] if n_l else 0
        if n_l and d0 % BLOCK_TILE == 0 and s0 > 0:
            offs = (y_t[:, None] + y_i[None, :]).reshape(-1)
            lo, hi = int(offs.min()), int(offs.max())
            span = (BLOCK_TILE - 1) * s0 + hi - lo + 1
            ybytes = 16 if Y.dtype == torch.complex128 else 8
            if span * ybytes + len(y_t) * nj * ni * 16 <= BLOCK_STAGE_BYTES:
                staged = (s0, lo, span)
        hit = (big, sm, gb32, tuple(perm), order, len(y_r), len(order), len(G), nj, ni, y_i, o_j,
               len(y_t), staged)
        _BAPPLY_CACHE[key] = hit
        if len(_BAPPLY_CACHE) > 1024:
            _BAPPLY_CACHE.popitem(last=False)
    big, sm, gb32, perm, order, R, nt, ng, nj, ni, y_i, o_j, nyt, staged = hit
    # the operator values: cached on the persistent BlockSparse, never on structure alone
    wkey = (perm, order.tobytes())
    w = None
    if X.source is not None:
        wc = X.source.__dict__.setdefault("_apply_cache", {})
        w = wc.get(wkey)
    if w is None:
        w = X.data.permute(list(perm)).reshape(len(X.tuples), nj, ni)[
            torch.as_tensor(order).to(dev)].to(torch.complex128).contiguous()
        if X.source is not None:
            wc[wkey] = w
    arr = ctypes.c_int64 * 4
    args = [R, nt, ng, nj, ni, None, _DTYPE_CODE[Y.dtype], ctypes.c_void_p(big.data_ptr()),
            ctypes.c_void_p(sm.data_ptr()), arr(*[int(v) for v in y_i] + [0] * (4 - ni)),
            int(conj_y), ctypes.c_void_p(gb32.data_ptr()), ctypes.c_void_p(w.data_ptr()),
            int(X.conj), None, _DTYPE_CODE[target.dtype],
            ctypes.c_void_p(big.data_ptr() + 8 * R), ctypes.c_void_p(sm.data_ptr() + 8 * nyt),
            arr(*[int(v) for v in o_j] + [0] * (4 - nj))]
    if staged is not None and USE_BLOCK_STAGED:
        args += [staged[0], staged[1], staged[2]]
    keep = (big, sm, gb32, w)
    if _RECORDER is not None:
        _RECORDER.ops.append(("block", args, Y, target, keep))
    _run_block(args, Y, target)


def _run_block(args, Y, target):
    args = list(args)
    args[5] = ctypes.c_void_p(Y.data_ptr())
    args[14] = ctypes.c_void_p(target.data_ptr())
    lib = _lib.load()
    if len(args) > 19:   # with the LDS-staging parameters
        _lib.check(lib.ptn_block_apply_staged(ctypes.c_void_p(stream_handle()), *args),
                   "ptn_block_apply_staged")
    else:
        _lib.check(lib.ptn_block_apply(ctypes.c_void_p(stream_handle()), *args),
                   "ptn_block_apply")


@dataclass
class NetPlan:
    steps: List[Tuple[int, int, Tuple[Label, ...]]]   # (slot_a, slot_b, result labels)
    flops: float                                       # 8 * MACs (4M counting)
    peak: int                                          # largest intermediate (elements)


_PLANS: Dict[tuple, NetPlan] = {}


def _merge_cost(l1: frozenset, l2: frozenset, dims: Dict[Label, int]) -> int:
    return math.prod(dims[l] for l in (l1 | l2))


def plan_network(labels: Sequence[Sequence[Label]], dims: Dict[Label, int],
                 out_labels: Sequence[Label], block: Optional[tuple] = None) -> NetPlan:
    """Optimal pairwise order (min total flops) by DP over subsets.

    ``block = (i, block_labels, tuples)`` marks tensor i as block sparse: a merge
    involving it costs 8 * M_dense * K_dense * N * (live blocks) and its
    intermediates hold only the live blocks.
    """
    n = len(labels)
    key = (tuple(tuple(l) for l in labels), tuple(sorted((repr(k), v) for k, v in dims.items())),
           tuple(out_labels),
           None if block is None else (block[0], tuple(block[1]), block[2].tobytes()))
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
    bi = -1
    bset: frozenset = frozenset()
    if block is not None:
        bi, blabels, btuples = block[0], list(block[1]), block[2]
        bset = frozenset(blabels)
        bcol = {l: i for i, l in enumerate(blabels)}
    nb_memo: Dict[frozenset, int] = {}

    def nblocks(open_blk: frozenset) -> int:
        hit = nb_memo.get(open_blk)
        if hit is None:
            cols = sorted(bcol[l] for l in open_blk)
            hit = len(np.unique(btuples[:, cols], axis=0)) if cols else 1
            nb_memo[open_blk] = hit
        return hit

    def prod(ls) -> int:
        return math.prod(dims[l] for l in ls)

    def size_of(mask: int) -> int:
        L = leg_of[mask]
        if bi >= 0 and mask >> bi & 1:
            return prod(L - bset) * nblocks(L & bset)
        return prod(L)

    def merge_cost(u: int, v: int) -> float:
        Lu, Lv = leg_of[u], leg_of[v]
        if bi >= 0 and (u >> bi & 1 or v >> bi & 1):
            if v >> bi & 1:
                Lu, Lv = Lv, Lu
            du = Lu - bset
            return 8.0 * prod(du - Lv) * prod(du & Lv) * prod(Lv - Lu) * nblocks(Lu & bset)
        return 8.0 * prod(Lu | Lv)
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
                    c = merge_cost(sub, other)
                    tot = best[sub][0] + best[other][0] + c
                    size = size_of(mask)
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


# Identity blocks of environments (``contract_network(identity=...)``): when set,
# every identity block is checked against the identity matrix on the device
# (one host synchronisation per split network) and a deviation above this bound
# raises.  Off in the sweeps; the tests turn it on.
CHECK_IDENTITY = False
IDENTITY_CHECK_TOL = 1e-10


class _SplitPart:
    """One term group of a network split at its identity environment blocks."""

    def __init__(self, tensors, labels, out_labels, conj, fixed, full):
        self.tensors = tensors
        self.labels = labels
        self.out_labels = out_labels
        self.conj = conj
        self.fixed = fixed        # [(output axis, index)] the part writes a slice at
        self.full = full          # writes the whole output


def _split_identity(tensors, labels, out_labels, conj, identity) -> Optional[List[_SplitPart]]:
    """The network as a sum of smaller networks, using environment blocks that are
    identity matrices.

    ``identity[i]`` lists the states s of tensor i's middle (bond-state) leg for
    which ``tensors[i][:, s, :]`` is the identity (legs (ket, ham, bra), ket and
    bra of equal dimension).  The network's block-sparse operator tensor W holds
    live bond-state tuples t; the network is the sum over t of the product of
    W's block t with every environment's block t_i.  Tuples are grouped by the
    environments whose block is not an identity:

    * no such environment: the identities join the ket and bra legs directly
      (an environment update becomes the Gram matrix of the state tensor, an
      effective Hamiltonian the operator on the physical leg alone);
    * one environment i: that environment, restricted to the states the group
      uses, meets the group's operator blocks W'(s_i, open bond legs, out, in)
      as a dense three-tensor network (a single contraction over (leg i, s_i)
      for the large tensors);
    * two or more: the original network restricted to those tuples.

    Open bond-state legs (the new environment's middle leg) keep their full range
    in W', or are fixed when the group uses one value of them, in which case the
    group writes that slice of the output.  Returns None when nothing can be
    split (no block-sparse W, an environment leg of W without identity data, or
    every tuple in the last group)."""
    n = len(tensors)
    bi = [i for i, t in enumerate(tensors) if isinstance(t, BlockSparse)]
    if len(bi) != 1:
        return None
    bi = bi[0]
    W = tensors[bi]
    wl = list(labels[bi])
    blabels = [wl[a] for a in W.block_axes]
    env_col: Dict[int, int] = {}
    for i, states in identity.items():
        ket, ham, bra = labels[i]
        if ham not in blabels or tensors[i].shape[0] != tensors[i].shape[2]:
            return None
        env_col[i] = blabels.index(ham)
    open_cols = []
    for c, l in enumerate(blabels):
        owners = [i for i in range(n) if i != bi and l in labels[i]]
        if not owners:
            if l not in out_labels:
                return None
            open_cols.append(c)
        elif not (len(owners) == 1 and env_col.get(owners[0]) == c):
            return None
    key = (tuple(sorted((env_col[i], tuple(identity[i])) for i in env_col)),
           tuple(open_cols))
    cache = W.__dict__.setdefault("_split_cache", {})
    plan = cache.get(key)
    if plan is None:
        plan = cache[key] = _split_plan(W, env_col, identity, open_cols, blabels)
    if plan is False:
        return None
    dlabels = [l for a, l in enumerate(wl) if a not in W.block_axes]
    parts: List[_SplitPart] = []
    for g in plan:
        if g["residual"] is not None:
            ts = list(tensors)
            ts[bi] = g["residual"]
            parts.append(_SplitPart(ts, [list(l) for l in labels], list(out_labels),
                                    list(conj), [], True))
            continue
        ni = g["ni"]
        rename = {labels[j][2]: labels[j][0] for j in env_col if j not in ni}
        ts, ls, cs = [], [], []
        for j in range(n):
            if j == bi or (j in env_col and j not in ni):
                continue
            t = tensors[j]
            if j in ni:
                # the states from the group's first to its last (a view, so a
                # recorded replay reads the caller's tensor); W' is zero on the
                # states in between that the group does not use
                S = g["states"]
                t = t.narrow(1, S[0], S[-1] + 1 - S[0])
            ts.append(t)
            ls.append([rename.get(l, l) for l in labels[j]])
            cs.append(conj[j])
        w_labels = ([labels[tuple(ni)[0]][1]] if ni else []) + \
            [blabels[c] for c in g["kept"]] + dlabels
        ts.append(g["w"])
        ls.append(w_labels)
        cs.append(conj[bi])
        gone = [blabels[c] for c, _ in g["fixed"]]
        out_l = [rename.get(l, l) for l in out_labels if l not in gone]
        fx = [(list(out_labels).index(blabels[c]), v) for c, v in g["fixed"]]
        parts.append(_SplitPart(ts, ls, out_l, cs, fx, not fx))
    # a part that writes the whole output goes first; the others accumulate
    parts.sort(key=lambda p: not p.full)
    return parts


def _split_plan(W: "BlockSparse", env_col, identity, open_cols, blabels):
    """The operator-side pieces of ``_split_identity`` for one block-sparse tensor
    (cached on it): tuple groups, the residual tensor, the dense W' per group."""
    T = W.tuples
    groups: Dict[frozenset, List[int]] = {}
    for r in range(len(T)):
        ni = frozenset(i for i, c in env_col.items() if int(T[r, c]) not in identity[i])
        groups.setdefault(ni, []).append(r)
    if all(len(k) >= 2 for k in groups):
        return False
    host = W.data.detach().cpu().numpy()
    plan = []
    residual = sorted(r for k, rows in groups.items() if len(k) >= 2 for r in rows)
    if residual:
        plan.append({"residual": BlockSparse(
            W.data[torch.as_tensor(residual, device=W.data.device)], W.block_axes,
            T[residual], W.shape)})
    for ni, rows in sorted(groups.items(), key=lambda kv: sorted(kv[0])):
        if len(ni) >= 2:
            continue
        S, sidx = [], {}
        if ni:
            (i,) = tuple(ni)
            S = sorted({int(T[r, env_col[i]]) for r in rows})
            sidx = {s: s - S[0] for s in S}
        fixed, kept = [], []
        for c in open_cols:
            vals = {int(T[r, c]) for r in rows}
            if len(vals) == 1:
                fixed.append((c, vals.pop()))
            else:
                kept.append(c)
        shape = ([S[-1] + 1 - S[0]] if ni else []) + [W.shape[W.block_axes[c]] for c in kept] + \
            list(host.shape[1:])
        Wd = np.zeros(shape, dtype=host.dtype)
        for r in rows:
            idx = ([sidx[int(T[r, env_col[tuple(ni)[0]]])]] if ni else []) + \
                [int(T[r, c]) for c in kept]
            Wd[tuple(idx)] += host[r]
        plan.append({"residual": None, "ni": ni, "states": S, "fixed": fixed, "kept": kept,
                     "w": torch.as_tensor(Wd).to(W.data.device)})
    return plan


def _check_identity(tensors, identity) -> None:
    for i, states in identity.items():
        E_ = tensors[i]
        eye = torch.eye(E_.shape[0], dtype=E_.dtype, device=E_.device)
        for s in states:
            dev = float((E_[:, s, :] - eye).abs().max())
            if dev > IDENTITY_CHECK_TOL:
                raise ValueError(f"environment block {s} of network tensor {i} is not the "
                                 f"identity (max deviation {dev:.2e})")


def _contract_split(parts: List[_SplitPart], out_shape, precision, out, out_dtype,
                    accumulate) -> torch.Tensor:
    if out is None:
        out = torch.empty(out_shape, dtype=out_dtype, device=parts[0].tensors[0].device)
    acc = accumulate
    if not acc and not parts[0].full:
        fill_(out, 0.0)
        acc = True
    for p in parts:
        view = out
        for ax, v in sorted(p.fixed, reverse=True):
            view = view.select(ax, v)
        contract_network(p.tensors, p.labels, p.out_labels, conj=p.conj, precision=precision,
                         out=view, out_dtype=out_dtype, accumulate=acc)
        acc = True
    return out


def contract_network(tensors: Sequence, labels: Sequence[Sequence[Label]],
                     out_labels: Sequence[Label], *, conj: Optional[Sequence[bool]] = None,
                     precision: str = FP64, out: Optional[torch.Tensor] = None,
                     out_dtype: torch.dtype = torch.complex128,
                     identity: Optional[Dict[int, Sequence[int]]] = None,
                     accumulate: bool = False) -> torch.Tensor:
    """Contract a whole network; ``conj[i]`` conjugates tensor i on load.

    At most one tensor may be a ``BlockSparse`` (the TTNO node tensor).
    ``identity`` ({tensor index: states}) names environment blocks that are
    identity matrices (``_split_identity``); ``accumulate`` adds the result to
    ``out`` instead of overwriting it."""
    n = len(tensors)
    conj = list(conj) if conj is not None else [False] * n
    if identity:
        parts = _split_identity(tensors, labels, out_labels, conj, identity)
        if parts is not None:
            if CHECK_IDENTITY:
                _check_identity(tensors, identity)
            dims = {}
            for t, l in zip(tensors, labels):
                dims.update(zip(l, t.shape))
            return _contract_split(parts, tuple(dims[l] for l in out_labels), precision,
                                   out, out_dtype, accumulate)
    dims: Dict[Label, int] = {}
    block = None
    for i, (t, l) in enumerate(zip(tensors, labels)):
        if len(l) != t.dim():
            raise ValueError("label count does not match tensor rank")
        for x, d in zip(l, t.shape):
            if dims.setdefault(x, d) != d:
                raise ValueError(f"dimension mismatch on label {x!r}")
        if isinstance(t, BlockSparse):
            if block is not None:
                raise ValueError("at most one block-sparse tensor per network")
            block = (i, [l[a] for a in t.block_axes], t.tuples)
    if n == 1:
        if block is not None:
            raise ValueError("a lone block-sparse tensor cannot be contracted")
        if accumulate:
            raise ValueError("a lone tensor cannot be accumulated")
        perm = [list(labels[0]).index(l) for l in out_labels]
        return permute_copy(tensors[0], perm, out=out, conj=conj[0], dtype=out_dtype)
    plan = plan_network(labels, dims, out_labels, block)
    slots: list = []
    for i, (t, l, c) in enumerate(zip(tensors, labels, conj)):
        if isinstance(t, BlockSparse):
            dl = [x for a, x in enumerate(l) if a not in t.block_axes]
            bl = [l[a] for a in t.block_axes]
            slots.append(_Blk(t.data, dl, bl, t.tuples, {x: dims[x] for x in bl}, conj=c,
                              tkey=t.tkey, source=t))
        else:
            slots.append((t, tuple(l), c))
    inter = storage_dtype(precision)
    for idx, (sa, sb, res) in enumerate(plan.steps):
        A_, B_ = slots[sa], slots[sb]
        last = idx == len(plan.steps) - 1
        if isinstance(B_, _Blk):
            A_, B_ = B_, A_
        if isinstance(A_, _Blk):
            tb, lb_, cb = B_
            if last and accumulate:
                raise ValueError("the last step of an accumulating network is block sparse")
            r = _contract_block(A_, tb, lb_, cb, res if last else None, precision,
                                out if last else None, out_dtype if last else inter)
            slots.append(r if isinstance(r, _Blk) else
                         (r, tuple(res) if last else
                          tuple([l for l in A_.dense_labels if l not in lb_] +
                                [l for l in lb_ if l not in A_.dense_labels
                                 and l not in A_.blabels]), False))
        else:
            ta, la_, ca = A_
            tb, lb_, cb = B_
            if res is None:
                res = tuple([l for l in la_ if l not in lb_] + [l for l in lb_ if l not in la_])
                if LAYOUT_LOOKAHEAD:
                    res = _lookahead_layout(res, len(tensors) + idx, plan.steps, idx, slots,
                                            dims)
            r = contract(ta, la_, tb, lb_, res, conj_a=ca, conj_b=cb, precision=precision,
                         out=out if last else None,
                         out_dtype=out_dtype if last else inter,
                         accumulate=accumulate and last)
            slots.append((r, res, False))
        slots[sa] = None
        slots[sb] = None
    final = slots[-1]
    return final[0]


LAYOUT_LOOKAHEAD = True


def _lookahead_layout(res: tuple, slot: int, steps, idx: int, slots, dims) -> tuple:
    """Memory order of the intermediate made by step ``idx`` (slot ``slot``).

    When its consumer step is a long-K contraction with a small output (a
    split-K launch: < 256 output tiles, K >= 1024), the labels that step
    contracts go innermost, in the order of the consumer's other operand (block
    labels first, then dense labels, for a block-sparse partner): both operands
    then have the same contiguous innermost K mode, so the tile rows of
    consecutive k share cache lines.  Otherwise, or when the partner's layout is
    not known yet, the default order is kept (it keeps strided launches strided
    and the block-apply rows contiguous)."""
    for j in range(idx + 1, len(steps)):
        a, b, _ = steps[j]
        if slot in (a, b):
            pi = b if a == slot else a
            if pi >= len(slots):      # produced after this step: layout not known yet
                return res
            partner = slots[pi]
            break
    else:
        return res
    if partner is None:
        return res
    if isinstance(partner, _Blk):
        order = list(partner.blabels) + list(partner.dense_labels)
    else:
        order = list(partner[1])
    pset = set(order)
    rset = set(res)
    m = math.prod(dims[l] for l in res if l not in pset)
    n = math.prod(dims[l] for l in order if l not in rset)
    k = math.prod(dims[l] for l in res if l in pset)
    if -(-m // 64) * -(-n // 64) >= 256 or k < 1024:
        return res
    free = [l for l in res if l not in pset]
    shared = [l for l in order if l in set(res)]
    return tuple(free + shared)


class CompiledNetwork:
    """A network contraction recorded once and replayed with new variable inputs.

    The first call runs ``contract_network`` while recording every launch
    (gather-GEMM descriptors, block applies, fills) together with the tensors
    they touch.  Later calls allocate only the output, patch the pointers of the
    variable inputs and the output into the recorded descriptors, and re-issue
    the launches: no planning, labels or offset-table lookups on the host.
    Intermediates are kept and reused, so calls must be ordered on one stream
    (the Krylov loop calls one apply after another).
    """

    def __init__(self, tensors: Sequence, labels, out_labels, *, conj=None,
                 precision: str = FP64, out_dtype: torch.dtype = torch.complex128,
                 variable: Sequence[int] = (0,),
                 identity: Optional[Dict[int, Sequence[int]]] = None):
        self.identity = identity
        self.tensors = list(tensors)
        self.labels = labels
        self.out_labels = out_labels
        self.conj = conj
        self.precision = precision
        self.out_dtype = out_dtype
        self.variable = tuple(variable)
        self.ops = None
        self.sig = None

    @staticmethod
    def _sig(ts):
        return tuple((tuple(t.shape), tuple(t.stride()), t.dtype) for t in ts)

    def __call__(self, *var: torch.Tensor) -> torch.Tensor:
        sig = self._sig(var)
        if self.ops is None or sig != self.sig:
            return self._record(var, sig)
        out = torch.empty(self.out_shape, dtype=self.out_dtype, device=var[0].device)
        if AUTOTUNE:
            for op, roles in zip(self.ops, self.op_roles):
                self._replay_op(op, roles, var, out)
            return out
        prof = PROFILER
        for item in self.runs:
            if item[0] == "op":
                self._replay_op(self.ops[item[1]], self.op_roles[item[1]], var, out)
                continue
            _, arr, count, flops, patches = item
            for view, field, role in patches:
                t = var[role[1]] if role[0] == "var" else out
                setattr(view, field, t.data_ptr() + role[2])
            if prof.active:
                prof.flops += flops
                prof.launches += count
            _lib.check(_lib.load().ptn_gemm_gather_many(ctypes.c_void_p(stream_handle()), arr,
                                                        count), "ptn_gemm_gather_many")
        return out

    @staticmethod
    def _replay_op(op, roles, var, out) -> None:
        kind = op[0]
        if kind == "gemm" or kind == "blas":
            d = op[1]
            for field, role in roles:
                t = var[role[1]] if role[0] == "var" else out
                setattr(d, field, t.data_ptr() + role[2])
            if kind == "gemm":
                _run_gemm(d, op[5], op[6], op[7], op[8])
            else:
                _run_blas(d, op[5], op[6], op[7], op[8])
        elif kind == "block":
            Y = op[2] if roles[0] is None else (var[roles[0][1]] if roles[0][0] == "var"
                                                else out)
            T = op[3] if roles[1] is None else (var[roles[1][1]] if roles[1][0] == "var"
                                                else out)
            _run_block(op[1], Y, T)
        else:
            if roles[0] is None:
                T = op[1]
            else:
                e0 = roles[0][2] // out.element_size()
                T = out.view(-1).narrow(0, e0, op[1].numel())
            fill_(T, op[2])

    def _build_runs(self) -> None:
        """Group the recorded gather-GEMMs into runs replayed by one
        ptn_gemm_gather_many call each (descriptors copied into one array; the
        pointers of variable inputs and of the output are patched per call).
        LARGE launches stay single, so the launch profiler can time them; block
        applies, fills and strided library GEMMs keep their own calls."""
        self.runs = []
        seg = []

        def flush():
            if not seg:
                return
            arr = (GemmDesc * len(seg))()
            views, patches, flops = [], [], 0.0
            for i, k in enumerate(seg):
                op = self.ops[k]
                arr[i] = op[1]
                views.append(arr[i])
                flops += 8.0 * op[5] * op[6] * op[7]
                for field, role in self.op_roles[k]:
                    patches.append((views[i], field, role))
            self.runs.append(("many", arr, len(seg), flops, patches))
            seg.clear()

        for k, op in enumerate(self.ops):
            if op[0] == "gemm" and gemm_kernel_class(
                    op[5], op[6], op[7], op[8], bool(op[1].a_kmajor),
                    bool(op[1].c_mmajor)) != "large":
                seg.append(k)
            else:
                flush()
                self.runs.append(("op", k))
        flush()

    def release(self) -> None:
        """Drop the recording and the intermediates it keeps (a later call records
        again).  For callers that hold many networks alive at once."""
        self.ops = None
        self.sig = None
        self.rec_inputs = self.rec_out = None
        self.op_roles = None
        self.runs = None

    def _record(self, var, sig) -> torch.Tensor:
        global _RECORDER
        ts = list(self.tensors)
        for i, v in zip(self.variable, var):
            ts[i] = v
        rec = _Recorder()
        prev, _RECORDER = _RECORDER, rec
        try:
            out = contract_network(ts, self.labels, self.out_labels, conj=self.conj,
                                   precision=self.precision, out_dtype=self.out_dtype,
                                   identity=self.identity)
        finally:
            _RECORDER = prev
        self.out_shape = tuple(out.shape)
        self.sig = sig
        self.ops = rec.ops
        self.rec_inputs, self.rec_out = ts, out

        out_base = out.untyped_storage().data_ptr()

        var_base = [v.untyped_storage().data_ptr() for v in var]

        def role(t, off=0):
            for k, v in enumerate(var):
                if t is v:
                    return ("var", k, off)
            if t is out:
                return ("out", 0, off)
            # slices of a variable input or of the output (a network split at its
            # identity blocks reads environment slices and writes output slices)
            base = t.untyped_storage().data_ptr()
            for k, v in enumerate(var):
                if base == var_base[k]:
                    return ("var", k, off + t.data_ptr() - v.data_ptr())
            if base == out_base:
                return ("out", 0, off + t.data_ptr() - out.data_ptr())
            return None

        self.op_roles = []
        for op in self.ops:
            if op[0] == "gemm" or op[0] == "blas":
                offs = op[9] if op[0] == "blas" else (0, 0, 0)
                roles = []
                for field, t, off in (("A", op[2], offs[0]), ("B", op[3], offs[1]),
                                      ("C", op[4], offs[2])):
                    r = role(t, off)
                    if r is not None:
                        roles.append((field, r))
                self.op_roles.append(roles)
            elif op[0] == "block":
                self.op_roles.append((role(op[2]), role(op[3])))
            else:
                self.op_roles.append((role(op[1]),))
        self._build_runs()
        return out


class RelaxedNetwork:
    """One H_eff action held at two precisions, for relaxed-precision Krylov.

    ``high`` contracts complex128 inputs with FP64 GEMMs; ``low`` contracts
    complex64 copies of the same environments and operator with 3M FP32 GEMMs
    (FP32 accumulation), either on the FP32 MFMA or as bf16x6 products on the
    bf16 MFMA (the same accuracy).  Calling the object runs ``high``.  The Krylov loop runs
    ``low`` for an application whose error cannot move the result by more than
    ``relax_tol`` times the input's norm (``krylov.krylov_exp_action``), and the
    optional ``lower`` network (3M bf16 products, FP32 accumulation, relative error
    ``lower_eps``) where even that error is small enough: the error of the j-th
    application enters the result weighted by the j-th expansion coefficient,
    which the loop estimates as it goes.  This is the relaxation of
    matrix-vector accuracy in inexact Krylov methods (Simoncini and Szyld,
    SIAM J. Sci. Comput. 25 (2003) 454-477; for the exponential, van den Eshof
    and Hochbruck, SIAM J. Sci. Comput. 27 (2006) 1438-1457).  ``low_eps`` is
    the relative error of one ``low`` application.
    """

    LOW_EPS = 1e-6
    LOWER_EPS = 3e-2

    def __init__(self, high: "CompiledNetwork", low: "CompiledNetwork", relax_tol: float,
                 low_eps: Optional[float] = None, lower: Optional["CompiledNetwork"] = None,
                 lower_eps: Optional[float] = None):
        if high.precision != FP64 or low.precision not in (FP32, BF16X6) or \
                (lower is not None and lower.precision != BF16):
            raise ValueError("RelaxedNetwork pairs an FP64 network with an FP32 or BF16X6 "
                             "one (and optionally a BF16 one)")
        self.high, self.low, self.lower = high, low, lower
        self.relax_tol = float(relax_tol)
        self.low_eps = self.LOW_EPS if low_eps is None else float(low_eps)
        self.lower_eps = self.LOWER_EPS if lower_eps is None else float(lower_eps)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.high(x)

    def release(self) -> None:
        for net in (self.high, self.low, self.lower):
            if net is not None:
                net.release()


# --------------------------------------------------------------------------------------
# batched replay: one grouped launch per op over many same-shape instances
# --------------------------------------------------------------------------------------
_GEMM_PTR_FIELDS = {"A": GemmDesc.A.offset, "B": GemmDesc.B.offset, "C": GemmDesc.C.offset}


class _Program:
    """The ops of one recorded network with every device pointer expressed as
    (role, byte offset): role ("in", i) is input i of the network, ("out", 0)
    its output, ("tmp", k) the k-th intermediate buffer.  Instances whose
    inputs have the template's shapes, strides and dtypes run the same ops with
    their own pointers; the offset tables are shared."""

    def __init__(self, net: "CompiledNetwork"):
        inputs, out = net.rec_inputs, net.rec_out
        ranges = []
        for i, t in enumerate(inputs):
            if isinstance(t, BlockSparse):
                raise ValueError("batched replay needs dense operator tensors")
            nb = t.numel() * t.element_size()
            if nb:
                ranges.append((t.data_ptr(), t.data_ptr() + nb, "in", i))
        ranges.append((out.data_ptr(), out.data_ptr() + out.numel() * out.element_size(),
                       "out", 0))
        temps: Dict[int, int] = {}
        self.temp_bytes: List[int] = []
        self.tables: List[Optional[torch.Tensor]] = []

        def classify(t, ptr):
            for a, b, kind, idx in ranges:
                if a <= ptr < b:
                    return kind, idx, ptr - a
            st = t.untyped_storage()
            base = st.data_ptr()
            k = temps.get(base)
            if k is None:
                k = temps[base] = len(self.temp_bytes)
                self.temp_bytes.append(-(-st.nbytes() // 256) * 256)
            return "tmp", k, ptr - base

        self.ops = []
        for op in net.ops:
            n_before = len(self.temp_bytes)
            if op[0] == "gemm":
                d = op[1]
                fields = []
                for name, t in (("A", op[2]), ("B", op[3]), ("C", op[4])):
                    fields.append((_GEMM_PTR_FIELDS[name] // 8,
                                   classify(t, getattr(d, name))))
                tmpl = np.frombuffer(bytes(d), dtype=np.uint8).copy()
                self.ops.append(("gemm", tmpl, fields))
                self.tables.append(op[9])   # offset tables the template points into
            elif op[0] == "fill":
                t = op[1]
                self.ops.append(("fill", classify(t, t.data_ptr()), t.numel(),
                                 _DTYPE_CODE[t.dtype], op[2]))
            else:
                raise ValueError(f"batched replay does not support {op[0]!r} ops")
        # live range of every intermediate: allocated before its first op, released
        # after its last one (the caching allocator reuses it in stream order)
        first: Dict[int, int] = {}
        last: Dict[int, int] = {}
        for i, op in enumerate(self.ops):
            roles = [r for _, r in op[2]] if op[0] == "gemm" else [op[1]]
            for kind, idx, _ in roles:
                if kind == "tmp":
                    first.setdefault(idx, i)
                    last[idx] = i
        self.alloc_at = [[k for k, f in first.items() if f == i] for i in range(len(self.ops))]
        self.free_at = [[k for k, l in last.items() if l == i] for i in range(len(self.ops))]
        self.out_shape = tuple(out.shape)
        self.out_dtype = out.dtype
        self.n_in = len(inputs)
        self._work: Dict[tuple, torch.Tensor] = {}

    def _workspace(self, kind: str, count: int, dev) -> torch.Tensor:
        key = (kind, count, dev)
        w = self._work.get(key)
        if w is None:
            lib = _lib.load()
            nb = (lib.ptn_gemm_grouped_work_bytes(count) if kind == "gemm"
                  else lib.ptn_ptr_batch_work_bytes(count))
            w = self._work[key] = torch.empty(int(nb), dtype=torch.uint8, device=dev)
        return w

    def run(self, in_ptrs: np.ndarray, dev) -> torch.Tensor:
        """in_ptrs: (count, n_in) int64 data pointers of each instance's inputs.
        Returns the outputs stacked as (count, *out_shape)."""
        count = in_ptrs.shape[0]
        outs = torch.empty((count,) + self.out_shape, dtype=self.out_dtype, device=dev)
        ostride = outs.stride(0) * outs.element_size()
        out_ptr = outs.data_ptr() + ostride * np.arange(count, dtype=np.int64)
        tmp_ptr: Dict[int, np.ndarray] = {}
        bufs: Dict[int, torch.Tensor] = {}

        def base(role):
            kind, idx, off = role
            if kind == "in":
                return in_ptrs[:, idx] + off
            if kind == "out":
                return out_ptr + off
            return tmp_ptr[idx] + off

        lib = _lib.load()
        stream = ctypes.c_void_p(stream_handle())
        for i, op in enumerate(self.ops):
            for k in self.alloc_at[i]:
                nb = self.temp_bytes[k]
                bufs[k] = torch.empty(nb * count, dtype=torch.uint8, device=dev)
                tmp_ptr[k] = bufs[k].data_ptr() + nb * np.arange(count, dtype=np.int64)
            if op[0] == "gemm":
                tmpl, fields = op[1], op[2]
                buf = np.tile(tmpl, (count, 1))
                b64 = buf.view(np.int64)
                for col, role in fields:
                    b64[:, col] = base(role)
                w = self._workspace("gemm", count, dev)
                _lib.check(lib.ptn_gemm_gather_grouped(
                    stream, buf.ctypes.data_as(ctypes.POINTER(GemmDesc)), count,
                    ctypes.c_void_p(w.data_ptr())), "ptn_gemm_gather_grouped")
            else:
                _, role, numel, code, v = op
                ptrs = np.ascontiguousarray(base(role), dtype=np.int64)
                w = self._workspace("ptr", count, dev)
                _lib.check(lib.ptn_fill_batched(stream, count, ptrs.ctypes.data_as(ctypes.c_void_p),
                                                numel, code, v.real, v.imag,
                                                ctypes.c_void_p(w.data_ptr())),
                           "ptn_fill_batched")
            for k in self.free_at[i]:
                del bufs[k]
        return outs


class BatchedNetwork:
    """Many instances of one network contraction (one CompiledNetwork each, same
    labels) applied together.  Instances are grouped by the shapes, strides and
    dtypes of their inputs; the first call of a group records its first member
    (run on its own) as the template, every other member -- and every member on
    later calls -- runs as one grouped launch per op."""

    def __init__(self, nets: Sequence["CompiledNetwork"]):
        self.nets = list(nets)
        self.groups: Dict[tuple, List[int]] = {}
        for i, net in enumerate(self.nets):
            sig = (tuple((tuple(t.shape), tuple(t.stride()) if not isinstance(t, BlockSparse)
                          else None, t.dtype) for t in net.tensors[1:]),
                   net.precision, net.out_dtype)
            self.groups.setdefault(sig, []).append(i)
        self.programs: Dict[tuple, _Program] = {}
        self._fixed: Dict[tuple, np.ndarray] = {}

    def __call__(self, idxs: Sequence[int], xs: Sequence[torch.Tensor]) -> List[torch.Tensor]:
        global _BLAS_OFF
        pos = {i: k for k, i in enumerate(idxs)}
        res: List[Optional[torch.Tensor]] = [None] * len(idxs)
        for sig, members in self.groups.items():
            act = [i for i in members if i in pos]
            if not act:
                continue
            prog = self.programs.get(sig)
            if prog is None:
                first = act[0]
                _BLAS_OFF += 1
                try:
                    res[pos[first]] = self.nets[first](xs[pos[first]])
                finally:
                    _BLAS_OFF -= 1
                prog = self.programs[sig] = _Program(self.nets[first])
                act = act[1:]
                if not act:
                    continue
            dev = xs[pos[act[0]]].device
            key = (sig, tuple(act))
            fixed = self._fixed.get(key)
            if fixed is None:
                fixed = np.array([[0] + [t.data_ptr() for t in self.nets[i].tensors[1:]]
                                  for i in act], dtype=np.int64)
                self._fixed[key] = fixed
            ptrs = fixed.copy()
            ptrs[:, 0] = [xs[pos[i]].data_ptr() for i in act]
            outs = prog.run(ptrs, dev)
            for k, i in enumerate(act):
                res[pos[i]] = outs[k]
        return res


def network_live_peak(labels, dims: Dict[Label, int], out_labels) -> int:
    """Largest number of intermediate elements alive at once while the network
    is contracted in its planned order (the output counted, inputs not)."""
    plan = plan_network(labels, dims, out_labels)
    n = len(labels)
    members: List[frozenset] = [frozenset([i]) for i in range(n)]
    out_set = frozenset(out_labels)
    alive: Dict[int, int] = {}
    peak = 0
    for a, b, _ in plan.steps:
        cover = members[a] | members[b]
        inside = set()
        for i in cover:
            inside |= set(labels[i])
        outside = set()
        for i in range(n):
            if i not in cover:
                outside |= set(labels[i])
        res = [x for x in inside if x in out_set or x in outside]
        members.append(cover)
        size = math.prod(dims[x] for x in res)
        peak = max(peak, sum(alive.values()) + size)   # operands and result live together
        alive.pop(a, None)
        alive.pop(b, None)
        alive[len(members) - 1] = size
    return peak


# Small networks that recur with the same structure and input layouts (the
# environment updates and dense effective operators of the small nodes of a
# tree, every sweep) are recorded once and replayed: the host then issues the
# recorded launches without planning or offset-table lookups.  A recording
# keeps its intermediates, so only networks whose live intermediates stay
# within NET_CACHE_MAX_BYTES are kept, up to NET_CACHE_BUDGET bytes in all.
NET_CACHE_MAX_BYTES = 64 << 20
NET_CACHE_BUDGET = 2 << 30
_NET_CACHE: "OrderedDict[tuple, Tuple[CompiledNetwork, int]]" = OrderedDict()
_NET_CACHE_BYTES = [0]
_NET_TOO_BIG: set = set()


def contract_network_cached(tensors: Sequence, labels: Sequence[Sequence[Label]],
                            out_labels: Sequence[Label], *, conj: Optional[Sequence[bool]] = None,
                            precision: str = FP64,
                            out_dtype: torch.dtype = torch.complex128,
                            identity: Optional[Dict[int, Sequence[int]]] = None) -> torch.Tensor:
    """``contract_network`` through a recorded replay when the network is small.

    Dense inputs are the replay's variables (matched by shape, stride and dtype);
    a block-sparse operator tensor is part of the key, held by the recording."""
    var = [i for i, t in enumerate(tensors) if not isinstance(t, BlockSparse)]
    if not var or len(tensors) < 2:
        return contract_network(tensors, labels, out_labels, conj=conj, precision=precision,
                                out_dtype=out_dtype, identity=identity)
    key = (tuple(tuple(l) for l in labels), tuple(out_labels),
           tuple(conj) if conj is not None else None, precision, out_dtype,
           None if not identity else tuple(sorted(identity.items())),
           tuple(("blk", id(t)) if isinstance(t, BlockSparse) else
                 (tuple(t.shape), tuple(t.stride()), t.dtype, t.device) for t in tensors))
    hit = _NET_CACHE.get(key)
    if hit is None:
        if key in _NET_TOO_BIG:
            return contract_network(tensors, labels, out_labels, conj=conj,
                                    precision=precision, out_dtype=out_dtype,
                                    identity=identity)
        dims: Dict[Label, int] = {}
        for t, lab in zip(tensors, labels):
            for l, d in zip(lab, t.shape):
                dims[l] = d
        esize = torch.empty((), dtype=storage_dtype(precision)).element_size()
        nbytes = esize * network_live_peak(labels, dims, out_labels)
        if nbytes > NET_CACHE_MAX_BYTES:
            if len(_NET_TOO_BIG) > 65536:
                _NET_TOO_BIG.clear()
            _NET_TOO_BIG.add(key)
            return contract_network(tensors, labels, out_labels, conj=conj,
                                    precision=precision, out_dtype=out_dtype,
                                    identity=identity)
        net = CompiledNetwork(tensors, labels, out_labels, conj=conj, precision=precision,
                              out_dtype=out_dtype, variable=var, identity=identity)
        _NET_CACHE[key] = (net, nbytes)
        _NET_CACHE_BYTES[0] += nbytes
        while _NET_CACHE_BYTES[0] > NET_CACHE_BUDGET and len(_NET_CACHE) > 1:
            _, (_, nb) = _NET_CACHE.popitem(last=False)
            _NET_CACHE_BYTES[0] -= nb
    else:
        net = hit[0]
        _NET_CACHE.move_to_end(key)
    return net(*[tensors[i] for i in var])


def dense_operator(net: "CompiledNetwork", transposed: bool = False) -> torch.Tensor:
    """The linear map x -> net(x) of a network whose variable is input 0, as a dense
    complex128 matrix: rows over the output labels, columns over input 0's labels
    (both flattened row-major), i.e. the network contracted without x.  With
    ``transposed`` the (cols, rows) transpose is returned instead (the matrix in
    column-major order), written directly by the contraction."""
    lab_x = list(net.labels[0])
    x = net.tensors[0]
    tensors = list(net.tensors[1:])
    labels = [list(l) for l in net.labels[1:]]
    conj = list(net.conj[1:]) if net.conj is not None else None
    out = lab_x + list(net.out_labels) if transposed else list(net.out_labels) + lab_x
    t = contract_network_cached(tensors, labels, out, conj=conj, precision=net.precision,
                         out_dtype=torch.complex128)
    cols = math.prod(x.shape)
    rows = t.numel() // max(cols, 1)
    return t.reshape(cols, rows) if transposed else t.reshape(rows, cols)
