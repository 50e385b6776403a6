"""Skinny contractions (K <= 16, N <= 64, M >= 8192: the row-streaming kernel with B in
LDS) against numpy einsum: both layouts of A and C, conjugation, alpha / accumulate
on strided views, complex64 operands, and the launch-configuration entry point."""
import ctypes

import numpy as np
import pytest
import torch

from pytreenet_amd import _lib
from pytreenet_amd import engine as E

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dev(a, dtype=torch.complex128):
    return torch.as_tensor(a).to(dtype).to("cuda")


CASES = [
    # (shapeA, labelsA, shapeB, labelsB, out)            (M, N, K); A and C m-contiguous
    ((10, 20000), "km", (10, 10), "kn", "nm"),           # 20000 x 10 x 10
    ((16, 20000), "km", (16, 64), "kn", "nm"),           # N = 64, K = 16
    ((1, 9000), "km", (1, 1), "kn", "nm"),               # K = N = 1
    ((16, 9000), "km", (64, 16), "nk", "nm"),            # B stored n-major
    ((5, 2, 30, 300), "kaxy", (5, 2, 8), "kbn", "bnaxy"),  # gathered modes on both sides
    # A or C contiguous along k / n: the other kernels (checked all the same)
    ((100, 100, 10), "xyk", (10, 50), "kn", "xyn"),
    ((100, 100, 10), "xyk", (10, 50), "kn", "nxy"),
]
SKINNY = 5      # the first five run on the skinny kernel


def _check(got, a, b, la, lb, lo, conj, tol):
    ref = np.einsum(f"{la},{lb}->{lo}", a.conj() if conj[0] else a, b.conj() if conj[1] else b)
    np.testing.assert_allclose(got, ref, rtol=0, atol=tol * np.abs(ref).max())


@pytest.mark.parametrize("case", CASES, ids=[f"{c[1]},{c[3]}->{c[4]}" for c in CASES])
@pytest.mark.parametrize("conj", [(False, False), (True, False), (False, True)])
def test_skinny_matches_einsum(gpu, case, conj):
    sa, la, sb, lb, lo = case
    _say(f"skinny gemm {sa}{la} x {sb}{lb} -> {lo} conj={conj}")
    rng = np.random.default_rng(sum(sa) + 3 * sum(sb) + 7 * conj[0] + 11 * conj[1])
    a, b = _crand(rng, sa), _crand(rng, sb)
    M = int(np.prod([d for d, l in zip(sa, la) if l not in lb]))
    N = int(np.prod([d for d, l in zip(sb, lb) if l not in la]))
    K = int(np.prod([d for d, l in zip(sa, la) if l in lb]))
    classes = []
    orig = E._run_gemm

    def seen(d, M_, N_, K_, p):
        classes.append(E.gemm_kernel_class(M_, N_, K_, p, bool(d.a_kmajor), bool(d.c_mmajor)))
        orig(d, M_, N_, K_, p)
    E._run_gemm = seen
    try:
        got = E.contract(_dev(a), la, _dev(b), lb, lo, conj_a=conj[0], conj_b=conj[1])
    finally:
        E._run_gemm = orig
    _say(f"  M={M} N={N} K={K}: {classes}")
    assert (classes == ["skinny"]) == (CASES.index(case) < SKINNY)
    _check(got.cpu().numpy(), a, b, la, lb, lo, conj, 1e-14 * 4)


@pytest.mark.parametrize("idx", [0, 3, 4])
def test_skinny_accumulate_alpha_views(gpu, idx):
    sa, la, sb, lb, lo = CASES[idx]
    _say(f"skinny gemm accumulate {la},{lb}->{lo}")
    rng = np.random.default_rng(200 + idx)
    a, b = _crand(rng, sa), _crand(rng, sb)
    ta = _dev(np.concatenate([a, a], axis=0))[: sa[0]]            # strided view of A (k)
    ref0 = np.einsum(f"{la},{lb}->{lo}", a, b)
    big = _crand(rng, (2 * ref0.shape[0],) + ref0.shape[1:])
    out = _dev(big)[::2]                                          # strided output view
    base = out.cpu().numpy().copy()
    E.contract(ta, la, _dev(b), lb, lo, out=out, accumulate=True, alpha=0.25 + 1.5j)
    ref = base + (0.25 + 1.5j) * ref0
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=0, atol=4e-14 * np.abs(ref).max())


def test_skinny_complex64_operands(gpu):
    sa, la, sb, lb, lo = CASES[4]
    _say("skinny gemm complex64")
    rng = np.random.default_rng(7)
    a, b = _crand(rng, sa), _crand(rng, sb)
    got = E.contract(_dev(a, torch.complex64), la, _dev(b, torch.complex64), lb, lo,
                     precision=E.FP32, out_dtype=torch.complex64)
    _check(got.cpu().numpy(), a.astype(np.complex64), b.astype(np.complex64), la, lb, lo,
           (False, False), 1e-6)


def test_skinny_config_kind_equals_rule(gpu):
    """PTN_GEMM_SKINNY through the configuration entry point equals the rule's launch
    (here the same kernel) and is refused for a K beyond its limit."""
    _say("skinny gemm config kind")
    lib = _lib.load()
    rng = np.random.default_rng(9)
    sa, la, sb, lb, lo = CASES[0]
    a, b = _dev(_crand(rng, sa)), _dev(_crand(rng, sb))
    outs = []
    orig = E._run_gemm
    for cfg in ((_lib.PTN_GEMM_RULE, 0, 0, 0), (_lib.PTN_GEMM_SKINNY, 0, 0, 0),
                (_lib.PTN_GEMM_TILE, 0, 0, 1)):
        def run(d, M, N, K, p, cfg=cfg):
            E._launch_config(lib, d, cfg)
        E._run_gemm = run
        try:
            outs.append(E.contract(a, la, b, lb, lo).cpu().numpy())
        finally:
            E._run_gemm = orig
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_allclose(outs[2], outs[0], rtol=0, atol=1e-13 * np.abs(outs[0]).max())
    big = _dev(_crand(rng, (9000, 17)))
    bk = _dev(_crand(rng, (17, 8)))
    def run_bad(d, M, N, K, p):
        E._launch_config(lib, d, (_lib.PTN_GEMM_SKINNY, 0, 0, 0))
    E._run_gemm = run_bad
    try:
        with pytest.raises(_lib.PtnBackendError):
            E.contract(big, "mk", bk, "kn", "mn")
    finally:
        E._run_gemm = orig
