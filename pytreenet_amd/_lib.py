"""ctypes binding of libptnhip.so (declarations in include/ptn_hip.h).

The product path has no CPU fallback: if the library is missing or no ROCm
device is present, every compute call raises ``PtnBackendError``.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libptnhip.so"

PTN_C128, PTN_C64 = 0, 1
PTN_COMPUTE_F64_4M, PTN_COMPUTE_F32_3M, PTN_COMPUTE_BF16_3M, PTN_COMPUTE_BF16X6_3M = 0, 1, 2, 3
PTN_GEMM_RULE, PTN_GEMM_WAVE, PTN_GEMM_TILE, PTN_GEMM_SKINNY = 0, 1, 2, 3
PTN_OK, PTN_EINVAL, PTN_EHIP, PTN_ESOLVER, PTN_ENOCONV, PTN_ERANGE = 0, -1, -2, -3, -4, -5

EXPORTED = (
    "ptn_gemm_gather", "ptn_gemm_gather_many", "ptn_gemm_gather_config", "ptn_permute", "ptn_fill", "ptn_krylov_work_bytes",
    "ptn_krylov_orthogonalize", "ptn_krylov_orthonormalize",
    "ptn_krylov_orthonormalize_noalloc", "ptn_normalize",
    "ptn_krylov_combine", "ptn_norm_sq", "ptn_dotc", "ptn_krylov_small_work_bytes",
    "ptn_krylov_expm_small",
    "ptn_gemm_strided", "ptn_gemm_strided_ok", "ptn_qr", "ptn_svd",
    "ptn_gemm_grouped_work_bytes", "ptn_gemm_gather_grouped", "ptn_krylov_batched_work_bytes",
    "ptn_krylov_orthonormalize_batched", "ptn_normalize_batched", "ptn_krylov_combine_batched",
    "ptn_ptr_batch_work_bytes", "ptn_fill_batched", "ptn_permute_batched", "ptn_potrf_lower", "ptn_trtri_lower", "ptn_block_apply",
    "ptn_qr_split_small", "ptn_block_apply_staged",
    "ptn_svd_jacobi_work_bytes", "ptn_svd_jacobi", "ptn_svd_jacobi_batched_work_bytes",
    "ptn_svd_jacobi_batched",
    "ptn_last_error", "ptn_version",
)


class PtnBackendError(RuntimeError):
    """The HIP backend is unavailable or a kernel call failed."""


class StridedDesc(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_int64), ("N", ctypes.c_int64), ("K", ctypes.c_int64),
        ("A", ctypes.c_void_p), ("a_sm", ctypes.c_int64), ("a_sk", ctypes.c_int64),
        ("a_conj", ctypes.c_int32), ("_pad0", ctypes.c_int32),
        ("B", ctypes.c_void_p), ("b_sk", ctypes.c_int64), ("b_sn", ctypes.c_int64),
        ("b_conj", ctypes.c_int32), ("_pad1", ctypes.c_int32),
        ("C", ctypes.c_void_p), ("c_sm", ctypes.c_int64), ("c_sn", ctypes.c_int64),
        ("dtype", ctypes.c_int32), ("beta_one", ctypes.c_int32),
        ("alpha_re", ctypes.c_double), ("alpha_im", ctypes.c_double),
    ]


class GemmDesc(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_int64), ("N", ctypes.c_int64), ("K", ctypes.c_int64),
        ("A", ctypes.c_void_p), ("a_off_m", ctypes.c_void_p), ("a_off_k", ctypes.c_void_p),
        ("a_dtype", ctypes.c_int32), ("a_conj", ctypes.c_int32), ("a_kmajor", ctypes.c_int32),
        ("c_lower", ctypes.c_int32),
        ("B", ctypes.c_void_p), ("b_off_k", ctypes.c_void_p), ("b_off_n", ctypes.c_void_p),
        ("b_dtype", ctypes.c_int32), ("b_conj", ctypes.c_int32), ("b_kmajor", ctypes.c_int32),
        ("_pad1", ctypes.c_int32),
        ("C", ctypes.c_void_p), ("c_off_m", ctypes.c_void_p), ("c_off_n", ctypes.c_void_p),
        ("c_dtype", ctypes.c_int32), ("beta_one", ctypes.c_int32),
        ("alpha_re", ctypes.c_double), ("alpha_im", ctypes.c_double),
        ("compute", ctypes.c_int32), ("c_mmajor", ctypes.c_int32),
        ("ws", ctypes.c_void_p), ("ws_bytes", ctypes.c_int64),
    ]


_lib = None


def load(path: os.PathLike | None = None) -> ctypes.CDLL:
    """Load (once) and return the shared library, raising if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise PtnBackendError(
            f"{p} is missing: build it with `python -m pytreenet_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(p))
    vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    sig = {
        "ptn_gemm_gather": (ctypes.c_int, [vp, ctypes.POINTER(GemmDesc)]),
        "ptn_gemm_gather_many": (ctypes.c_int, [vp, ctypes.POINTER(GemmDesc), i32]),
        "ptn_gemm_gather_config": (ctypes.c_int, [vp, ctypes.POINTER(GemmDesc), i32, i32, i32,
                                                  i32]),
        "ptn_gemm_strided": (ctypes.c_int, [vp, ctypes.POINTER(StridedDesc)]),
        "ptn_gemm_strided_ok": (ctypes.c_int, [ctypes.POINTER(StridedDesc)]),
        "ptn_permute": (ctypes.c_int, [vp, i32, vp, vp, vp, i32, vp, vp, i32, i32]),
        "ptn_fill": (ctypes.c_int, [vp, vp, i64, i32, dbl, dbl]),
        "ptn_krylov_work_bytes": (i64, [i64, i32]),
        "ptn_krylov_orthogonalize": (ctypes.c_int, [vp, i64, vp, i64, i32, vp, vp, vp]),
        "ptn_krylov_orthonormalize": (ctypes.c_int, [vp, i64, vp, i64, i32, vp, vp, vp, vp]),
        "ptn_krylov_orthonormalize_noalloc": (ctypes.c_int, [vp, i64, vp, i64, i32, vp, vp, vp]),
        "ptn_svd_jacobi_work_bytes": (ctypes.c_int64, [i64, i64]),
        "ptn_svd_jacobi": (ctypes.c_int, [vp, i64, i64, vp, i64, i32, vp, vp, vp, vp, vp]),
        "ptn_svd_jacobi_batched_work_bytes": (ctypes.c_int64, [i32]),
        "ptn_svd_jacobi_batched": (ctypes.c_int, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                                  vp, vp]),
        "ptn_normalize": (ctypes.c_int, [vp, i64, vp, vp, vp, vp]),
        "ptn_krylov_combine": (ctypes.c_int, [vp, i64, vp, i64, i32, vp, vp]),
        "ptn_norm_sq": (ctypes.c_int, [vp, i64, vp, vp, vp]),
        "ptn_krylov_small_work_bytes": (i64, [i64, i32]),
        "ptn_krylov_expm_small": (ctypes.c_int, [vp, i64, vp, vp, vp, i32, dbl, dbl, dbl, dbl,
                                                 i32, vp, vp]),
        "ptn_dotc": (ctypes.c_int, [vp, i64, vp, vp, vp, vp]),
        "ptn_qr": (ctypes.c_int, [vp, i64, i64, vp, vp, vp]),
        "ptn_svd": (ctypes.c_int, [vp, i64, i64, vp, vp, vp, vp, vp]),
        "ptn_potrf_lower": (ctypes.c_int, [vp, i32, vp, i64, vp, dbl, dbl, vp]),
        "ptn_trtri_lower": (ctypes.c_int, [vp, i32, vp, vp]),
        "ptn_qr_split_small": (ctypes.c_int, [vp, i32, vp, vp, i32, vp, vp, vp]),
        "ptn_block_apply": (ctypes.c_int, [vp, i64, i32, i32, i32, i32, vp, i32, vp, vp, vp,
                                           i32, vp, vp, i32, vp, i32, vp, vp, vp]),
        "ptn_block_apply_staged": (ctypes.c_int, [vp, i64, i32, i32, i32, i32, vp, i32, vp, vp,
                                                  vp, i32, vp, vp, i32, vp, i32, vp, vp, vp,
                                                  i64, i64, i32]),
        "ptn_gemm_grouped_work_bytes": (i64, [i32]),
        "ptn_gemm_gather_grouped": (ctypes.c_int, [vp, ctypes.POINTER(GemmDesc), i32, vp]),
        "ptn_krylov_batched_work_bytes": (i64, [i32]),
        "ptn_krylov_orthonormalize_batched": (ctypes.c_int, [vp, i64, i32, i32, vp, vp, vp, vp,
                                                             vp, vp]),
        "ptn_normalize_batched": (ctypes.c_int, [vp, i64, i32, vp, vp, vp, vp]),
        "ptn_krylov_combine_batched": (ctypes.c_int, [vp, i64, i32, i32, vp, vp, vp, vp, vp]),
        "ptn_ptr_batch_work_bytes": (i64, [i32]),
        "ptn_fill_batched": (ctypes.c_int, [vp, i32, vp, i64, i32, dbl, dbl, vp]),
        "ptn_permute_batched": (ctypes.c_int, [vp, i32, i32, vp, vp, vp, i32, vp, vp, i32, i32,
                                               vp]),
        "ptn_last_error": (ctypes.c_char_p, []),
        "ptn_version": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().ptn_last_error().decode(errors="replace")
        raise PtnBackendError(f"{what} failed with code {rc}: {msg}")
