"""GPU kernel tests through the C ABI: gather-GEMM, permute, Krylov level-1, QR, SVD."""
import numpy as np
import pytest
import torch

from pytreenet_amd import engine as E
from pytreenet_amd import krylov as K
from pytreenet_amd import splitting as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dev(a, dtype=torch.complex128):
    return torch.as_tensor(a, dtype=dtype).to("cuda")


@pytest.mark.parametrize("mnk", [(1, 1, 1), (16, 16, 4), (70, 50, 33), (129, 65, 300), (5, 200, 2)])
def test_gemm_plain_fp64(gpu, mnk):
    m, n, k = mnk
    _say(f"gemm plain fp64 {mnk}")
    rng = np.random.default_rng(m * 1000 + n)
    a, b = _crand(rng, (m, k)), _crand(rng, (k, n))
    c = E.contract(_dev(a), "mk", _dev(b), "kn", "mn").cpu().numpy()
    ref = a @ b
    assert np.abs(c - ref).max() <= 1e-13 * np.abs(ref).max() * max(1, np.sqrt(k))


CASES = [
    # (shapeA, labelsA, shapeB, labelsB, out)
    ((6, 5, 7, 2), "abcp", (5, 3, 9), "bwx", "axcpw"),
    ((6, 5, 7, 2), "abcp", (7, 4, 6), "cwa", "wbp"),
    ((3, 4, 5), "ijk", (5, 4, 6), "kjl", "li"),
    ((40, 3, 30), "xyz", (30, 3, 50), "zyq", "qx"),
    ((2, 3), "ij", (4, 5), "kl", "likj"),            # outer product (K empty)
]


@pytest.mark.parametrize("case", CASES, ids=[c[4] for c in CASES])
@pytest.mark.parametrize("conj", [(False, False), (True, False), (False, True)])
def test_gather_gemm_permutations_fp64(gpu, case, conj):
    sa, la, sb, lb, lo = case
    _say(f"gather fp64 {la},{lb}->{lo} conj={conj}")
    rng = np.random.default_rng(len(lo) + 7 * conj[0] + 3 * conj[1])
    a, b = _crand(rng, sa), _crand(rng, sb)
    got = E.contract(_dev(a), la, _dev(b), lb, lo, conj_a=conj[0], conj_b=conj[1]).cpu().numpy()
    ref = np.einsum(f"{la},{lb}->{lo}", a.conj() if conj[0] else a, b.conj() if conj[1] else b)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_gather_gemm_views_accumulate_alpha(gpu):
    _say("gather strided views + accumulate + alpha")
    rng = np.random.default_rng(11)
    a, b = _crand(rng, (8, 9, 10)), _crand(rng, (10, 8, 7))
    ta = _dev(a).permute(2, 0, 1)          # non-contiguous view, labels "kij"
    tb = _dev(b)[:, ::2, :]                # strided slice (10, 4, 7)
    out = _dev(_crand(rng, (9, 4, 8, 7))).permute(0, 1, 3, 2)   # strided output view
    base = out.cpu().numpy().copy()
    E.contract(ta, "kij", tb, "klm", "jlmi", out=out, accumulate=True, alpha=0.5 - 2j)
    ref = base + (0.5 - 2j) * np.einsum("ijk,klm->jlmi", a, b[:, ::2, :])
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-11)


@pytest.mark.parametrize("dt", [torch.complex128, torch.complex64])
def test_gather_gemm_fp32_3m(gpu, dt):
    _say(f"gather fp32 3M inputs {dt}")
    rng = np.random.default_rng(12)
    a, b = _crand(rng, (50, 64, 3)), _crand(rng, (64, 3, 70))
    got = E.contract(_dev(a, dt), "mkw", _dev(b, dt), "kwn", "nm", precision=E.FP32,
                     out_dtype=torch.complex64).cpu().numpy()
    ref = np.einsum("mkw,kwn->nm", a, b)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err


def test_permute_fill(gpu):
    _say("permute/fill")
    rng = np.random.default_rng(13)
    a = _crand(rng, (3, 4, 5, 6))
    got = E.permute_copy(_dev(a), [2, 0, 3, 1], conj=True).cpu().numpy()
    np.testing.assert_array_equal(got, a.conj().transpose(2, 0, 3, 1))
    low = E.permute_copy(_dev(a), [0, 1, 2, 3], dtype=torch.complex64).cpu().numpy()
    np.testing.assert_allclose(low, a, rtol=1e-6)
    z = torch.empty(1000, dtype=torch.complex128, device="cuda")
    E.fill_(z, 1.5 - 0.5j)
    assert torch.all(z == (1.5 - 0.5j))


@pytest.mark.parametrize("n,nvec", [(1, 1), (37, 3), (100000, 9), (300001, 17)])
def test_krylov_orthogonalize(gpu, n, nvec):
    _say(f"krylov orth n={n} nvec={nvec}")
    rng = np.random.default_rng(n + nvec)
    nv = min(nvec, n)
    q, _ = np.linalg.qr(_crand(rng, (n, nv)))
    Q = np.ascontiguousarray(q.T)
    w = _crand(rng, n)
    tQ, tw = _dev(Q), _dev(w)
    packed = torch.empty(nv + 1, dtype=torch.complex128, device="cuda")
    import ctypes
    from pytreenet_amd import _lib
    _lib.check(_lib.load().ptn_krylov_orthogonalize(
        ctypes.c_void_p(E.stream_handle()), n, ctypes.c_void_p(tQ.data_ptr()), n, nv,
        ctypes.c_void_p(tw.data_ptr()), ctypes.c_void_p(packed.data_ptr()),
        ctypes.c_void_p(K._work(tQ.device).data_ptr())), "orth")
    h = packed.cpu().numpy()
    h_ref = Q.conj() @ w
    w1 = w - Q.T @ h_ref
    h2 = Q.conj() @ w1
    w2 = w1 - Q.T @ h2
    np.testing.assert_allclose(h[:nv], h_ref + h2, atol=1e-10 * np.linalg.norm(w))
    assert abs(h[nv].real - np.vdot(w, w).real) < 1e-10 * np.vdot(w, w).real
    assert abs(h[nv].imag - np.vdot(w2, w2).real) < 1e-9 * np.vdot(w, w).real + 1e-300
    np.testing.assert_allclose(tw.cpu().numpy(), w2, atol=1e-10 * np.linalg.norm(w))


@pytest.mark.parametrize("n,nvec,pad", [(4, 2, 0), (8192, 9, 3), (8193, 3, 0), (100000, 11, 5),
                                         (100000, 16, 0), (70000, 40, 2)])
def test_krylov_orthonormalize(gpu, n, nvec, pad):
    """One-workgroup (n <= 8192) and multi-block forms (three sweeps with beta from
    |w'|^2 - |h2|^2 for bases up to 16 vectors, four sweeps above), basis rows with
    stride > n."""
    _say(f"krylov orthonormalize n={n} nvec={nvec} pad={pad}")
    import ctypes
    from pytreenet_amd import _lib
    rng = np.random.default_rng(7 * n + nvec)
    nv = min(nvec, n)
    q, _ = np.linalg.qr(_crand(rng, (n, nv)))
    ld = n + pad
    Qbuf = np.zeros((nv + 1, ld), dtype=complex)
    Qbuf[:nv, :n] = q.T
    w = _crand(rng, n)
    tQ, tw = _dev(Qbuf), _dev(w)
    packed = torch.empty(nv + 1, dtype=torch.complex128, device="cuda")
    vp = ctypes.c_void_p
    _lib.check(_lib.load().ptn_krylov_orthonormalize(
        vp(E.stream_handle()), n, vp(tQ.data_ptr()), ld, nv, vp(tw.data_ptr()),
        vp(tQ[nv].data_ptr()), vp(packed.data_ptr()), vp(K._work(tQ.device).data_ptr())),
        "orthonormalize")
    h = packed.cpu().numpy()
    Q = q.T
    h1 = Q.conj() @ w
    w1 = w - Q.T @ h1
    h2 = Q.conj() @ w1
    w2 = w1 - Q.T @ h2
    scale = np.linalg.norm(w)
    np.testing.assert_allclose(h[:nv], h1 + h2, atol=1e-12 * scale)
    assert abs(h[nv].real - scale ** 2) < 1e-12 * scale ** 2
    assert abs(h[nv].imag - np.vdot(w2, w2).real) < 1e-12 * scale ** 2
    # (w's content on return is left unspecified when qnext is given: the
    # multi-block form writes the second projection straight into qnext)
    got = tQ.cpu().numpy()
    np.testing.assert_allclose(got[nv, :n], w2 / np.linalg.norm(w2), atol=1e-12)
    np.testing.assert_array_equal(got[:nv], Qbuf[:nv])       # basis untouched
    # the new vector is orthonormal to the basis
    assert np.max(np.abs(Q.conj() @ got[nv, :n])) < 1e-12


@pytest.mark.parametrize("n", [1, 8192, 8193, 300001])
def test_normalize(gpu, n):
    _say(f"normalize n={n}")
    import ctypes
    from pytreenet_amd import _lib
    rng = np.random.default_rng(n)
    x = _crand(rng, n)
    tx = _dev(x)
    y = torch.empty_like(tx)
    out = torch.empty(1, dtype=torch.complex128, device="cuda")
    vp = ctypes.c_void_p
    lib, work = _lib.load(), K._work(tx.device)
    _lib.check(lib.ptn_normalize(vp(E.stream_handle()), n, vp(tx.data_ptr()), vp(y.data_ptr()),
                                 vp(out.data_ptr()), vp(work.data_ptr())), "normalize")
    nrm2 = np.vdot(x, x).real
    assert abs(out.cpu().item() - nrm2) < 1e-12 * nrm2
    np.testing.assert_allclose(y.cpu().numpy(), x / np.sqrt(nrm2), atol=1e-14)
    # in place, and the zero vector maps to zero
    _lib.check(lib.ptn_normalize(vp(E.stream_handle()), n, vp(tx.data_ptr()), vp(tx.data_ptr()),
                                 vp(out.data_ptr()), vp(work.data_ptr())), "normalize")
    np.testing.assert_allclose(tx.cpu().numpy(), x / np.sqrt(nrm2), atol=1e-14)
    z = torch.zeros(n, dtype=torch.complex128, device="cuda")
    _lib.check(lib.ptn_normalize(vp(E.stream_handle()), n, vp(z.data_ptr()), vp(y.data_ptr()),
                                 vp(out.data_ptr()), vp(work.data_ptr())), "normalize")
    assert out.cpu().item() == 0 and not torch.any(y != 0)


def test_dot_norm_combine(gpu):
    _say("dot/norm/combine")
    rng = np.random.default_rng(21)
    x, y = _crand(rng, 5000), _crand(rng, 5000)
    assert abs(K.dotc(_dev(x), _dev(y)) - np.vdot(x, y)) < 1e-10 * abs(np.vdot(x, y))
    assert abs(K.norm_sq(_dev(x)) - np.vdot(x, x).real) < 1e-10 * np.vdot(x, x).real
    Q = _crand(rng, (3, 5000))
    c = np.array([1 - 1j, 0.5, 2j])
    out = torch.empty(5000, dtype=torch.complex128, device="cuda")
    K.combine(_dev(Q), c, out)
    np.testing.assert_allclose(out.cpu().numpy(), c @ Q, atol=1e-12)


@pytest.mark.parametrize("shape,q_legs,r_legs", [((4, 3, 5), [0, 1], [2]), ((2, 2, 8), [0, 1], [2]),
                                                 ((6, 7, 2), [1, 2], [0])])
def test_qr_keep(gpu, shape, q_legs, r_legs):
    _say(f"qr keep {shape}")
    rng = np.random.default_rng(31)
    t = _crand(rng, shape)
    q, r = S.tensor_qr_decomposition(_dev(t), q_legs, r_legs, mode=S.SplitMode.KEEP)
    q, r = q.cpu().numpy(), r.cpu().numpy()
    n = int(np.prod([shape[i] for i in r_legs]))
    assert q.shape[-1] == n and r.shape[0] == n
    rec = np.tensordot(q, r, axes=([q.ndim - 1], [0]))
    perm = np.argsort(q_legs + r_legs)
    np.testing.assert_allclose(rec.transpose(perm), t, atol=1e-12)
    qm = q.reshape(-1, n)
    g = qm.conj().T @ qm
    k = min(qm.shape[0], n)
    np.testing.assert_allclose(g, np.diag([1.0] * k + [0.0] * (n - k)), atol=1e-12)


@pytest.mark.parametrize("shape", [(3, 4, 5), (30, 2, 40), (64, 2, 2)])
def test_svd(gpu, shape):
    _say(f"svd {shape}")
    rng = np.random.default_rng(41)
    t = _crand(rng, shape)
    u, s, vh = S.tensor_svd(_dev(t), [0, 2], [1])
    mat = t.transpose(0, 2, 1).reshape(shape[0] * shape[2], shape[1])
    np.testing.assert_allclose(s, np.linalg.svd(mat, compute_uv=False), rtol=1e-12)
    rec = np.einsum("aek,k,kb->aeb", u.cpu().numpy(), s, vh.cpu().numpy()).transpose(0, 2, 1)
    np.testing.assert_allclose(rec, t, atol=1e-11)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gather_gemm_split_k(gpu, prec):
    """Few output tiles, long K: the K range is split across workgroups."""
    _say(f"split-K {prec}")
    rng = np.random.default_rng(51)
    a, b = _crand(rng, (40, 5, 3000)), _crand(rng, (3000, 5, 16))
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    out = _dev(_crand(rng, (16, 40)), dt)
    base = out.cpu().numpy().copy()
    E.contract(_dev(a, dt), "mwk", _dev(b, dt), "kwn", "nm", precision=prec, out=out,
               accumulate=True, alpha=2.0)
    ref = base + 2.0 * np.einsum("mwk,kwn->nm", a, b)
    err = np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < (1e-13 if prec == "fp64" else 2e-6), err


@pytest.mark.parametrize("shape,leg", [((8, 8, 2), 0), ((16, 4, 4, 2), 2), ((64, 64, 2), 1),
                                       ((300, 40), 1)])
def test_cholesky_qr3_split(gpu, shape, leg):
    _say(f"cholqr3 {shape} leg {leg}")
    rng = np.random.default_rng(61)
    t = _crand(rng, shape)
    q, r, ok = S.cholesky_qr3_leg(_dev(t), leg)
    assert ok
    q, r = q.cpu().numpy(), r.cpu().numpy()
    n = shape[leg]
    rec = np.moveaxis(np.tensordot(np.moveaxis(q, leg, -1), r, axes=([q.ndim - 1], [0])), -1, leg)
    np.testing.assert_allclose(rec, t, atol=1e-12 * np.abs(t).max())
    qm = np.moveaxis(q, leg, -1).reshape(-1, n)
    np.testing.assert_allclose(qm.conj().T @ qm, np.eye(n), atol=1e-13)


def test_cholesky_qr3_flags_rank_deficient(gpu):
    _say("cholqr3 rank deficient")
    rng = np.random.default_rng(62)
    a = _crand(rng, (20, 3)) @ _crand(rng, (3, 6))       # rank 3, 6 columns
    q, r, ok = S.cholesky_qr3_leg(_dev(a), 1)
    assert not ok


@pytest.mark.parametrize("n", [200, 17, 256, 512, 600])
def test_potrf_trtri(gpu, n):
    """Cholesky factor and its inverse (row-blocked inverse up to n = 512, one
    workgroup per column above) against numpy; X is exactly lower triangular."""
    _say(f"potrf/trtri n={n}")
    rng = np.random.default_rng(63 + n)
    a = _crand(rng, (n + 100, n))
    g = a.conj().T @ a
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    L, X = S._potrf(_dev(g), 0.0, 0.0, status)
    L, X = L.cpu().numpy(), X.cpu().numpy()
    assert status.item() == 0
    np.testing.assert_allclose(L @ L.conj().T, g, atol=1e-10 * np.abs(g).max())
    assert np.abs(np.triu(L, 1)).max() == 0
    assert np.abs(np.triu(X, 1)).max() == 0
    np.testing.assert_allclose(X @ L, np.eye(n), atol=1e-11)
    np.testing.assert_allclose(X, np.linalg.inv(L), atol=1e-11 * np.abs(X).max())


def _layout(rng, rows, cols, colmajor, dtype):
    """A rows x cols complex matrix stored column-major (unit row stride) or row-major."""
    x = _crand(rng, (rows, cols)).astype(dtype)
    if colmajor:
        t = torch.from_numpy(np.ascontiguousarray(x.T)).to("cuda").t()
    else:
        t = torch.from_numpy(np.ascontiguousarray(x)).to("cuda")
    return x, t


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_gemm_strided_layouts(gpu, dtype):
    """ptn_gemm_strided over every (A, B, C) layout and conjugation, against numpy;
    ptn_gemm_strided_ok against the rule that a conjugated operand must be transposed
    in the column-major call."""
    import ctypes
    from itertools import product
    from pytreenet_amd import _lib
    lib = _lib.load()
    M, N, K = 96, 80, 72
    rng = np.random.default_rng(5)
    tol = 1e-12 if dtype == np.complex128 else 2e-5
    code = _lib.PTN_C128 if dtype == np.complex128 else _lib.PTN_C64
    ran = 0
    for acol, bcol, ccol, ca, cb, acc in product([0, 1], repeat=6):
        a, tA = _layout(rng, M, K, acol, dtype)
        b, tB = _layout(rng, K, N, bcol, dtype)
        c0, tC = _layout(rng, M, N, ccol, dtype)
        alpha = 0.75 - 0.5j
        sd = _lib.StridedDesc()
        sd.M, sd.N, sd.K = M, N, K
        sd.A, sd.a_sm, sd.a_sk, sd.a_conj = tA.data_ptr(), tA.stride(0), tA.stride(1), ca
        sd.B, sd.b_sk, sd.b_sn, sd.b_conj = tB.data_ptr(), tB.stride(0), tB.stride(1), cb
        sd.C, sd.c_sm, sd.c_sn = tC.data_ptr(), tC.stride(0), tC.stride(1)
        sd.dtype, sd.beta_one = code, acc
        sd.alpha_re, sd.alpha_im = alpha.real, alpha.imag
        # column-major C: A is op N (col-major, no conj) or op T/C (row-major);
        # row-major C runs C^T = B^T A^T: the roles of the layouts flip
        f1 = ccol and (not ca or not acol) and (not cb or not bcol)
        f2 = (not ccol) and (not cb or bcol) and (not ca or acol)
        ok = lib.ptn_gemm_strided_ok(ctypes.byref(sd))
        assert ok == int(bool(f1 or f2)), (acol, bcol, ccol, ca, cb)
        if not ok:
            continue
        _lib.check(lib.ptn_gemm_strided(ctypes.c_void_p(E.stream_handle()), ctypes.byref(sd)),
                   "strided")
        ref = alpha * ((a.conj() if ca else a) @ (b.conj() if cb else b)) + (c0 if acc else 0)
        got = tC.cpu().numpy()
        np.testing.assert_allclose(got, ref, atol=tol * np.abs(ref).max())
        ran += 1
    assert ran == 36


@pytest.fixture
def blas_fp64():
    E.BLAS_FP64 = True
    yield
    E.BLAS_FP64 = False


def test_contract_blas_matches_gather(gpu, blas_fp64):
    """Engine contractions routed to rocBLAS (FP64 strided dispatch switched on)
    equal the gather-GEMM ones."""
    rng = np.random.default_rng(17)
    x = _dev(_crand(rng, (64, 64, 64, 2)))
    env = _dev(_crand(rng, (64, 5, 64)))
    cases = [((x, "abcp"), (env, "awB"), "bcpwB", False),
             ((x, "abcp"), (env, "cwB"), "abpwB", True),
             ((env, "awB"), (x, "abcp"), "wBbcp", False)]
    for (A, la), (B, lb), lo, cj in cases:
        try:
            E.USE_BLAS = False
            ref = E.contract(A, la, B, lb, lo, conj_b=cj).cpu().numpy()
            E.USE_BLAS = True
            got = E.contract(A, la, B, lb, lo, conj_b=cj).cpu().numpy()
        finally:
            E.USE_BLAS = True
        np.testing.assert_allclose(got, ref, atol=1e-11 * np.abs(ref).max())
    # the compiled replay of a network that routes through rocBLAS
    net = E.CompiledNetwork([x, env, env], [["a", "b", "c", "p"], ["a", "w", "A"],
                                            ["b", "v", "B"]], ["A", "w", "B", "v", "c", "p"])
    first = net(x).cpu().numpy()
    y = _dev(_crand(rng, (64, 64, 64, 2)))
    again = net(y).cpu().numpy()
    E.USE_BLAS = False
    try:
        ref = E.contract_network([y, env, env], [["a", "b", "c", "p"], ["a", "w", "A"],
                                                 ["b", "v", "B"]],
                                 ["A", "w", "B", "v", "c", "p"]).cpu().numpy()
    finally:
        E.USE_BLAS = True
    np.testing.assert_allclose(again, ref, atol=1e-11 * np.abs(ref).max())
    assert not np.allclose(first, again)


@pytest.mark.parametrize("m,n,rank", [(2000, 64, None), (48, 3000, None), (40, 640, 12),
                                      (256, 4096, None)])
def test_svd_qr_preconditioned(gpu, m, n, rank):
    """Rectangular panels go through CholeskyQR3 + small SVD (or, rank deficient, the
    direct SVD): singular values equal numpy's to O(u ||A||), U S V^H = A, U and V
    orthonormal."""
    _say(f"svd via qr {m}x{n} rank={rank}")
    rng = np.random.default_rng(m + n)
    A = _crand(rng, (m, n))
    if rank is not None:
        A = _crand(rng, (m, rank)) @ _crand(rng, (rank, n))
    panel = _dev(np.asfortranarray(A).reshape(-1, order="F").copy())
    before = dict(S.SVD_STATS)
    s, U, Vh, k = S.svd_colmajor(panel, m, n)
    ref = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(s, ref, atol=1e-13 * ref[0])
    Um = U.cpu().numpy().reshape(k, m).T
    Vm = Vh.cpu().numpy().reshape(n, k).T
    np.testing.assert_allclose((Um * s) @ Vm, A, atol=1e-12 * ref[0])
    if rank is None:
        np.testing.assert_allclose(Um.conj().T @ Um, np.eye(k), atol=1e-12)
        np.testing.assert_allclose(Vm @ Vm.conj().T, np.eye(k), atol=1e-12)
        assert S.SVD_STATS["qr_tall" if m > n else "qr_wide"] == before["qr_tall" if m > n
                                                                          else "qr_wide"] + 1
    else:
        assert S.SVD_STATS["qr_fallback"] == before["qr_fallback"] + 1


QR_SMALL_CASES = [((2, 2), 0), ((2, 2), 1), ((16, 2, 2, 2, 2), 0), ((16, 2, 2, 2, 2), 1),
                  ((3, 8), 1), ((48, 64), 1), ((12, 256), 1), ((5, 7, 3), 1)]


@pytest.mark.parametrize("shape,leg", QR_SMALL_CASES)
def test_qr_split_small(gpu, shape, leg):
    """One-launch Householder QR(KEEP): numpy's QR in the gauge diag(R) >= 0, the
    KEEP zero padding, and the same factors as CholeskyQR3 on full-rank tall panels."""
    _say(f"qr split small {shape} leg={leg}")
    rng = np.random.default_rng(sum(shape) + leg)
    t = _crand(rng, shape)
    q, r = S.qr_split_small(_dev(t), leg)
    q, r = q.cpu().numpy(), r.cpu().numpy()
    n = shape[leg]
    others = [i for i in range(len(shape)) if i != leg]
    A = np.transpose(t, others + [leg]).reshape(-1, n)
    m = A.shape[0]
    k = min(m, n)
    Qm = np.transpose(q, others + [leg]).reshape(-1, n)
    qn, rn = np.linalg.qr(A)
    d = np.diag(rn) / np.abs(np.diag(rn))
    np.testing.assert_allclose(Qm[:, :k], qn * d[None, :], atol=1e-12)
    np.testing.assert_allclose(r[:k], rn * d.conj()[:, None], atol=1e-12 * np.abs(A).max())
    assert np.all(Qm[:, k:] == 0) and np.all(r[k:] == 0)
    assert np.all(np.abs(np.tril(r, -1)) == 0)
    np.testing.assert_allclose(Qm @ r, A, atol=1e-12 * np.abs(A).max())
    if m >= n:
        qc, rc, ok = S.cholesky_qr3_leg(_dev(t), leg)
        assert ok
        np.testing.assert_allclose(q, qc.cpu().numpy(), atol=1e-11)
        np.testing.assert_allclose(r, rc.cpu().numpy(), atol=1e-11 * np.abs(A).max())


def test_qr_split_small_rank_deficient_and_views(gpu):
    _say("qr split small rank-deficient / strided view")
    rng = np.random.default_rng(77)
    u, v = _crand(rng, (12, 1)), _crand(rng, (1, 6))
    t = (u @ v).reshape(3, 4, 6)                       # rank 1 over (3*4) x 6
    q, r = S.qr_split_small(_dev(t), 2)
    Qm, R = q.cpu().numpy().reshape(12, 6), r.cpu().numpy()
    np.testing.assert_allclose(Qm @ R, t.reshape(12, 6), atol=1e-12 * np.abs(t).max())
    # Householder Q stays orthonormal on a rank-deficient panel
    np.testing.assert_allclose(Qm.conj().T @ Qm, np.eye(6), atol=1e-12)
    assert np.all(np.diag(R).real >= 0) and np.all(np.diag(R).imag == 0)
    base = _crand(rng, (6, 5, 4))
    view = _dev(base).permute(2, 0, 1)                 # (4, 6, 5), non-contiguous
    q, r = S.qr_split_small(view, 0)
    tv = base.transpose(2, 0, 1)
    rec = np.einsum("kab,kj->jab", q.cpu().numpy(), r.cpu().numpy())
    np.testing.assert_allclose(rec, tv, atol=1e-12 * np.abs(tv).max())


@pytest.mark.parametrize("conj", [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize("case", [
    ((200, 70), "mk", (70, 150), "kn", "nm"),            # 64x64 tiles, m-contiguous C
    ((4, 50, 9, 7), "abkc", (7, 9, 130), "ckn", "nab"),   # gathered, m-contiguous C
    ((100, 3000), "mk", (3000, 90), "kn", "nm"),          # split-K grid: no transposed stores
])
def test_direct_gemm_m_contiguous_output(gpu, case, conj):
    """The 64 x 64 kernel with the C-layout hint set (lanes along m in the
    epilogue), plain and accumulating with alpha."""
    sa, la, sb, lb, lo = case
    _say(f"direct gemm m-contiguous C {la},{lb}->{lo} conj={conj}")
    rng = np.random.default_rng(sum(sa) + sum(sb) + 3 * conj[0] + 5 * conj[1])
    a, b = _crand(rng, sa), _crand(rng, sb)
    ref = np.einsum(f"{la},{lb}->{lo}", a.conj() if conj[0] else a, b.conj() if conj[1] else b)
    got = E.contract(_dev(a), la, _dev(b), lb, lo, conj_a=conj[0], conj_b=conj[1]).cpu().numpy()
    k = int(np.prod([d for d, l in zip(sa, la) if l in lb]))
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * np.sqrt(k) * np.abs(ref).max())
    base = _crand(rng, ref.shape)
    out = _dev(base)
    E.contract(_dev(a), la, _dev(b), lb, lo, conj_a=conj[0], conj_b=conj[1], out=out,
               accumulate=True, alpha=-0.5 + 0.25j)
    np.testing.assert_allclose(out.cpu().numpy(), base + (-0.5 + 0.25j) * ref, rtol=0,
                               atol=1e-14 * np.sqrt(k) * np.abs(ref).max())


@pytest.mark.parametrize("out_dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("conj", [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize("case", [
    ((4000, 37), "mk", (37, 2100), "kn", "mn"),               # n-contiguous C
    ((4000, 37), "mk", (37, 2100), "kn", "nm"),               # m-contiguous C
    ((40, 100, 3, 13), "abkc", (13, 3, 2100), "ckn", "nab"),  # gathered m and k
])
def test_gemm_fp32_large_c64(gpu, case, conj, out_dt):
    """LARGE FP32 products of complex64 operands (tile grid >= 2048, ragged edge
    tiles, odd K): rocBLAS cgemm for the strided layouts, the gather kernel for
    the gathered one; plain and accumulating with alpha, against einsum."""
    sa, la, sb, lb, lo = case
    _say(f"large gemm fp32 c64 {la},{lb}->{lo} conj={conj} out={out_dt}")
    rng = np.random.default_rng(sum(sa) + 3 * conj[0] + 5 * conj[1])
    a, b = _crand(rng, sa), _crand(rng, sb)
    a32, b32 = a.astype(np.complex64), b.astype(np.complex64)
    ref = np.einsum(f"{la},{lb}->{lo}", a32.conj() if conj[0] else a32,
                    b32.conj() if conj[1] else b32).astype(np.complex128)
    M = int(np.prod([d for d, l in zip(sa, la) if l not in lb]))
    N = int(np.prod([d for d, l in zip(sb, lb) if l not in la]))
    K = int(np.prod([d for d, l in zip(sa, la) if l in lb]))
    assert E.gemm_kernel_class(M, N, K, E.FP32) == "large"
    ta, tb = _dev(a, torch.complex64), _dev(b, torch.complex64)
    got = E.contract(ta, la, tb, lb, lo, conj_a=conj[0], conj_b=conj[1], precision=E.FP32,
                     out_dtype=out_dt).cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err
    base = _crand(rng, ref.shape)
    out = _dev(base, out_dt)
    base = out.cpu().numpy().astype(np.complex128)
    E.contract(ta, la, tb, lb, lo, conj_a=conj[0], conj_b=conj[1], precision=E.FP32, out=out,
               accumulate=True, alpha=-0.5 + 0.25j)
    want = base + (-0.5 + 0.25j) * ref
    err = np.abs(out.cpu().numpy() - want).max() / np.abs(want).max()
    assert err < 2e-6, err


@pytest.mark.parametrize("case", [
    ((3000, 40), "mk", (40, 24), "kn", "mn", torch.complex128, E.FP64),    # thin N
    ((24, 700), "mk", (700, 900), "kn", "nm", torch.complex128, E.FP64),   # thin M, m-contiguous C
    ((300, 260, 2), "mkp", (260, 2, 200), "kpn", "mn", torch.complex128, E.FP64),  # mid, long K
    ((70, 33), "mk", (33, 50), "kn", "mn", torch.complex128, E.FP64),      # small, ragged
    ((500, 300), "mk", (300, 40), "kn", "nm", torch.complex64, E.FP32),    # mixed form
])
def test_gemm_configurations_and_autotuner(gpu, case):
    """Every launch configuration the autotuner times (library rule, wave kernel
    block shapes, 64 x 64 kernel with and without K splits) gives the product;
    the tuned launch replays the fastest one."""
    sa, la, sb, lb, lo, dt, prec = case
    _say(f"gemm configurations {la},{lb}->{lo} {prec}")
    rng = np.random.default_rng(sum(sa) + sum(sb))
    a, b = _crand(rng, sa), _crand(rng, sb)
    ta = _dev(a, torch.complex128)               # complex128 A: the mixed form's Krylov side
    tb = _dev(b, dt)
    bb = b.astype(np.complex64) if dt == torch.complex64 else b
    ref = np.einsum(f"{la},{lb}->{lo}", a, bb)
    k = int(np.prod([d for d, l in zip(sa, la) if l in lb]))
    tol = (1e-14 * np.sqrt(k) if prec == E.FP64 else 1e-5) * np.abs(ref).max()
    E.AUTOTUNE = True
    try:
        n0 = len(E.TUNE_LOG)
        got = E.contract(ta, la, tb, lb, lo, precision=prec).cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=0, atol=tol)
        assert len(E.TUNE_LOG) == n0 + 1
        key, chosen, times = E.TUNE_LOG[-1]
        assert chosen == min(times, key=times.get) and len(times) >= 3
        for cfg in times:
            E._TUNED[key] = cfg
            got = E.contract(ta, la, tb, lb, lo, precision=prec).cpu().numpy()
            np.testing.assert_allclose(got, ref, rtol=0, atol=tol, err_msg=str(cfg))
        E._TUNED[key] = chosen
    finally:
        E.AUTOTUNE = False


def test_gemm_config_rejects_unknown(gpu):
    """A wave block shape without a kernel instantiation, and K splits beyond the
    workspace, are errors, not launches."""
    import ctypes
    from pytreenet_amd import _lib
    rng = np.random.default_rng(3)
    a, b = _dev(_crand(rng, (40, 300))), _dev(_crand(rng, (300, 30)))
    E.AUTOTUNE = True
    try:
        E.contract(a, "mk", b, "kn", "mn")
        key, chosen, _ = E.TUNE_LOG[-1]
        for bad in ((_lib.PTN_GEMM_WAVE, 3, 3, 0), (_lib.PTN_GEMM_TILE, 0, 0, 0),
                    (_lib.PTN_GEMM_TILE, 0, 0, 1 << 20), (7, 0, 0, 0)):
            E._TUNED[key] = bad
            with pytest.raises(Exception):
                E.contract(a, "mk", b, "kn", "mn")
        E._TUNED[key] = chosen
    finally:
        E.AUTOTUNE = False


@pytest.mark.parametrize("case", [
    ((4000, 37), "mk", (37, 2100), "kn", "mn", torch.complex64, (False, False)),  # LARGE grid
    ((40, 100, 3, 13), "abkc", (13, 3, 2100), "ckn", "nab", torch.complex64, (True, False)),
    ((100, 3000), "mk", (3000, 90), "kn", "nm", torch.complex128, (False, True)),  # split K
    ((70, 33), "mk", (33, 50), "kn", "mn", torch.complex64, (False, False)),       # ragged
])
def test_gather_gemm_bf16_3m(gpu, case):
    """The BF16 compute mode (3M bf16 products, FP32 accumulation) against einsum:
    within bf16 accuracy, and measurably less accurate than FP32 (so the bf16
    kernel, not an FP32 one, ran); accumulation with alpha."""
    sa, la, sb, lb, lo, adt, conj = case
    _say(f"gather bf16 3M {la},{lb}->{lo} A={adt} conj={conj}")
    rng = np.random.default_rng(sum(sa) + 7)
    a, b = _crand(rng, sa), _crand(rng, sb)
    ref = np.einsum(f"{la},{lb}->{lo}", a.conj() if conj[0] else a, b.conj() if conj[1] else b)
    ta, tb = _dev(a, adt), _dev(b, torch.complex64)
    M = int(np.prod([d for d, l in zip(sa, la) if l not in lb]))
    N = int(np.prod([d for d, l in zip(sb, lb) if l not in la]))
    K = int(np.prod([d for d, l in zip(sa, la) if l in lb]))
    assert E.gemm_kernel_class(M, N, K, E.BF16) != "wave"
    got = E.contract(ta, la, tb, lb, lo, conj_a=conj[0], conj_b=conj[1], precision=E.BF16,
                     out_dtype=torch.complex64).cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert 1e-4 < err < 3e-2, err
    base = _crand(rng, ref.shape)
    out = _dev(base, torch.complex64)
    base = out.cpu().numpy().astype(np.complex128)
    E.contract(ta, la, tb, lb, lo, conj_a=conj[0], conj_b=conj[1], precision=E.BF16, out=out,
               accumulate=True, alpha=-0.5 + 0.25j)
    want = base + (-0.5 + 0.25j) * ref
    err = np.abs(out.cpu().numpy() - want).max() / np.abs(want).max()
    assert err < 3e-2, err


@pytest.mark.parametrize("case", [
    ((4000, 37), "mk", (37, 2100), "kn", "mn", torch.complex64, (False, False)),  # LARGE grid
    ((40, 100, 3, 13), "abkc", (13, 3, 2100), "ckn", "nab", torch.complex64, (True, False)),
    ((100, 3000), "mk", (3000, 90), "kn", "nm", torch.complex128, (False, True)),  # split K
    ((70, 33), "mk", (33, 50), "kn", "mn", torch.complex64, (False, False)),       # ragged
    ((2048, 1280), "mk", (1280, 256), "kn", "mn", torch.complex64, (False, False)),  # k-major A
    ((1280, 2048), "km", (1280, 256), "kn", "mn", torch.complex64, (True, True)),    # m-major A
])
def test_gather_gemm_bf16x6_3m(gpu, case):
    """The BF16X6 compute mode (three-piece bf16 splits, six piece products per real
    product, FP32 accumulation) against einsum: within the accuracy of the FP32
    mode on the same complex64 operands (so FP32 accuracy), yet not bitwise the
    FP32 kernel's result; accumulation with alpha."""
    sa, la, sb, lb, lo, adt, conj = case
    _say(f"gather bf16x6 3M {la},{lb}->{lo} A={adt} conj={conj}")
    rng = np.random.default_rng(sum(sa) + 11)
    a, b = _crand(rng, sa), _crand(rng, sb)
    ta, tb = _dev(a, adt), _dev(b, torch.complex64)
    a_in = ta.cpu().numpy().astype(np.complex128)
    b_in = tb.cpu().numpy().astype(np.complex128)
    ref = np.einsum(f"{la},{lb}->{lo}", a_in.conj() if conj[0] else a_in,
                    b_in.conj() if conj[1] else b_in)
    M = int(np.prod([d for d, l in zip(sa, la) if l not in lb]))
    N = int(np.prod([d for d, l in zip(sb, lb) if l not in la]))
    K = int(np.prod([d for d, l in zip(sa, la) if l in lb]))
    assert E.gemm_kernel_class(M, N, K, E.BF16X6) != "wave"
    got = E.contract(ta, la, tb, lb, lo, conj_a=conj[0], conj_b=conj[1], precision=E.BF16X6,
                     out_dtype=torch.complex128).cpu().numpy()
    f32 = E.contract(ta, la, tb, lb, lo, conj_a=conj[0], conj_b=conj[1], precision=E.FP32,
                     out_dtype=torch.complex128).cpu().numpy()
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max() / scale
    err32 = np.abs(f32 - ref).max() / scale
    assert err < max(4 * err32, 1e-7) and err < 2e-6, (err, err32)
    assert not np.array_equal(got, f32)
    base = _crand(rng, ref.shape)
    out = _dev(base, torch.complex64)
    base = out.cpu().numpy().astype(np.complex128)
    E.contract(ta, la, tb, lb, lo, conj_a=conj[0], conj_b=conj[1], precision=E.BF16X6, out=out,
               accumulate=True, alpha=-0.5 + 0.25j)
    want = base + (-0.5 + 0.25j) * ref
    err = np.abs(out.cpu().numpy() - want).max() / np.abs(want).max()
    assert err < 2e-6, err


@pytest.mark.parametrize("n,rows", [(256, 20000), (200, 3000), (40, 5000)])
def test_gram_lower_tiles(gpu, n, rows):
    """A Gram product with the lower-triangle hint: every element on and below the
    diagonal equals the full product bit for bit; above it an element is either
    the full product's or left as it was (tiles strictly above the diagonal may be
    skipped, and are on the split-K tile kernel)."""
    _say(f"gram lower tiles n={n} rows={rows}")
    rng = np.random.default_rng(n + rows)
    a = _dev(_crand(rng, (rows, n)))
    full = E.contract(a, "kc", a, "kd", "cd", conj_a=True).cpu().numpy()
    sentinel = 7.0 - 3.0j
    out = torch.full((n, n), sentinel, dtype=torch.complex128, device="cuda")
    E.contract(a, "kc", a, "kd", "cd", conj_a=True, out=out, lower=True)
    got = out.cpu().numpy()
    low = np.tril(np.ones((n, n), dtype=bool))
    assert np.array_equal(got[low], full[low])
    assert np.all((got == full) | (got == sentinel))
    if n == 256:
        # split-K on the 64 x 64 tile kernel: the tiles above the diagonal were skipped
        assert np.all(got[:64, 64:] == sentinel)
    ref = a.cpu().numpy().conj().T @ a.cpu().numpy()
    np.testing.assert_allclose(got[low], ref[low], rtol=0, atol=1e-12 * np.abs(ref).max())
