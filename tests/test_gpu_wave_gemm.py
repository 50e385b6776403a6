"""Small and thin FP64 contractions (the wave-level MFMA GEMM: M or N within one or
two 16-blocks) against numpy einsum: every block shape, K split over waves with
both partial-sum reductions, conjugation, alpha / accumulate and gathered
(permuted, strided) operands."""
import numpy as np
import pytest
import torch

from pytreenet_amd import engine as E

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dev(a):
    return torch.as_tensor(a, dtype=torch.complex128).to("cuda")


CASES = [
    # (shapeA, labelsA, shapeB, labelsB, out)          (M, N, K)      block shape, split
    ((16, 20000), "mk", (16, 20000), "nk", "mn"),      # 16 x 16 x 20000: (1,1), wide reduce
    ((2, 640), "mk", (640, 2), "kn", "nm"),            # 2 x 2 x 640: (1,1), 10 splits
    ((20, 3000), "mk", (3000, 30), "kn", "mn"),        # (2,2) split
    ((30, 500), "mk", (10, 500), "nk", "nm"),          # (2,1) split
    ((10, 500), "mk", (500, 30), "kn", "mn"),          # (1,2) split
    ((50, 100, 16), "xyk", (16, 16), "kn", "nyx"),     # 5000 x 16 x 16: (4,1), one split
    ((16, 80), "mk", (100, 80, 50), "xky", "ymx"),     # 16 x 5000 x 80: (1,4)
    ((50, 100, 200), "xyk", (200, 16), "kn", "yxn"),   # 5000 x 16 x 200: (4,1), 3 splits
    ((7, 3, 5), "abk", (5, 4), "kc", "cab"),           # 21 x 4 x 5
    ((3, 2), "mk", (2, 1), "kn", "mn"),                # tiny
    # C contiguous along the long side: stores from transposed accumulators
    ((50, 100, 16), "xyk", (16, 16), "kn", "nxy"),     # 5000 x 16 x 16, m-contiguous C
    ((16, 80), "mk", (100, 80, 50), "xky", "mxy"),     # operands swapped, then m-contiguous
    ((20, 30), "mk", (30, 25), "kn", "nm"),            # (2,2), m-contiguous C
    ((40, 30, 16), "xyk", (16, 80), "kn", "xyn"),      # 1200 x 80 x 16: (1,5)
    ((48, 200), "mk", (200, 30, 20), "kxy", "xym"),    # 600 x 48 x 200 after the swap: (1,3)
    ((40, 100), "mk", (100, 60), "kn", "mn"),          # 40 x 60 x 100: (3,1)
    ((40, 30, 100), "xyk", (100, 40), "kn", "nxy"),    # 1200 x 40 x 100: (1,3), m-contiguous C
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[1]},{c[3]}->{c[4]}:{i}" for i, c in
                                             enumerate(CASES)])
@pytest.mark.parametrize("conj", [(False, False), (True, False), (False, True)])
def test_wave_gemm_matches_einsum(gpu, case, conj):
    sa, la, sb, lb, lo = case
    _say(f"wave gemm {sa}{la} x {sb}{lb} -> {lo} conj={conj}")
    rng = np.random.default_rng(sum(sa) + 3 * sum(sb) + 7 * conj[0] + 11 * conj[1])
    a, b = _crand(rng, sa), _crand(rng, sb)
    got = E.contract(_dev(a), la, _dev(b), lb, lo, conj_a=conj[0], conj_b=conj[1]).cpu().numpy()
    ref = np.einsum(f"{la},{lb}->{lo}", a.conj() if conj[0] else a, b.conj() if conj[1] else b)
    k = max(1, int(np.prod([d for d, l in zip(sa, la) if l in lb])))
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * np.sqrt(k) * np.abs(ref).max())


@pytest.mark.parametrize("idx", [0, 3, 5, 7, 10, 11, 13, 15, 16])
def test_wave_gemm_accumulate_alpha_views(gpu, idx):
    """Strided operand views and output, C += alpha A B."""
    sa, la, sb, lb, lo = CASES[idx]
    _say(f"wave gemm accumulate {la},{lb}->{lo}")
    rng = np.random.default_rng(100 + idx)
    a, b = _crand(rng, sa), _crand(rng, sb)
    ta = _dev(np.concatenate([a, a], axis=-1))[..., : sa[-1]]     # strided view of A
    tb = _dev(b)
    ref0 = np.einsum(f"{la},{lb}->{lo}", a, b)
    big = _crand(rng, ref0.shape[::-1])
    out = _dev(big).permute(*reversed(range(ref0.ndim)))          # transposed output view
    base = out.cpu().numpy().copy()
    E.contract(ta, la, tb, lb, lo, out=out, accumulate=True, alpha=0.25 + 1.5j)
    ref = base + (0.25 + 1.5j) * ref0
    k = max(1, int(np.prod([d for d, l in zip(sa, la) if l in lb])))
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=0,
                               atol=1e-14 * np.sqrt(k) * np.abs(ref).max())


@pytest.mark.parametrize("idx", [0, 1, 2, 5, 6, 7, 10, 13, 14, 15])
@pytest.mark.parametrize("dts", [("c64", "c64", "c64"), ("c128", "c64", "c64"),
                                 ("c64", "c64", "c128"), ("c64", "c128", "c64")])
def test_wave_gemm_mixed_form(gpu, idx, dts):
    """Small and thin products of the mixed form run on the wave kernel too: the
    complex64 operands are widened on load and accumulated in FP64, so the result
    is the exact product of the given operands up to FP64 rounding (and the final
    rounding of a complex64 output)."""
    sa, la, sb, lb, lo = CASES[idx]
    _say(f"wave gemm mixed {la},{lb}->{lo} {dts}")
    rng = np.random.default_rng(200 + idx)
    tdt = {"c64": torch.complex64, "c128": torch.complex128}
    a = torch.as_tensor(_crand(rng, sa)).to(tdt[dts[0]])
    b = torch.as_tensor(_crand(rng, sb)).to(tdt[dts[1]])
    got = E.contract(a.cuda(), la, b.cuda(), lb, lo, precision=E.FP32,
                     out_dtype=tdt[dts[2]]).cpu().numpy()
    ref = np.einsum(f"{la},{lb}->{lo}", a.numpy().astype(complex), b.numpy().astype(complex))
    k = max(1, int(np.prod([d for d, l in zip(sa, la) if l in lb])))
    tol = (2e-7 if dts[2] == "c64" else 1e-14 * np.sqrt(k)) * np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=0, atol=tol)
