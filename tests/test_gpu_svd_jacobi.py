"""One-sided Jacobi SVD kernel (ptn_svd_jacobi, ptn_svd_jacobi_batched) against
numpy's LAPACK SVD: singular values to 1e-13 |A|, reconstruction, unitary factors,
graded and rank-deficient spectra, both orientations, and the batched launch."""
import ctypes

import numpy as np
import pytest
import torch

from pytreenet_amd import _lib

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _rand(rng, m, n, spectrum=None):
    A = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    if spectrum is not None:
        u, _, vh = np.linalg.svd(A, full_matrices=False)
        A = (u * spectrum) @ vh
    return A


def _dev_colmajor(A):
    return torch.as_tensor(np.asfortranarray(A).ravel(order="F")).cuda()


def _run(lib, A, on_rows):
    m, n = A.shape
    k = min(m, n)
    dA = _dev_colmajor(A)
    S = torch.empty(k, dtype=torch.float64, device="cuda")
    U = torch.empty(m * k, dtype=torch.complex128, device="cuda")
    Vh = torch.empty(k * n, dtype=torch.complex128, device="cuda")
    ws = torch.empty(int(lib.ptn_svd_jacobi_work_bytes(m, n)), dtype=torch.uint8, device="cuda")
    info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.ptn_svd_jacobi(stream, m, n, _vp(dA), m, on_rows, _vp(S), _vp(U), _vp(Vh),
                                  _vp(ws), _vp(info)), "ptn_svd_jacobi")
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert np.array_equal(dA.cpu().numpy(), np.asfortranarray(A).ravel(order="F"))
    return (S.cpu().numpy(), U.cpu().numpy().reshape(k, m).T,
            Vh.cpu().numpy().reshape(n, k).T)


def _check(A, s, U, Vh, tol=1e-13):
    k = min(A.shape)
    ref = np.linalg.svd(A, compute_uv=False)
    scale = max(ref[0], 1e-300)
    assert np.all(np.diff(s) <= 0)
    np.testing.assert_allclose(s, ref, rtol=0, atol=tol * scale * max(A.shape) ** 0.5)
    rec = (U * s) @ Vh
    assert np.linalg.norm(rec - A) <= tol * 10 * np.linalg.norm(A)
    live = s > 1e-12 * scale
    Ul = U[:, live]
    np.testing.assert_allclose(Ul.conj().T @ Ul, np.eye(live.sum()), atol=1e-12)
    np.testing.assert_allclose(Vh @ Vh.conj().T, np.eye(k), atol=1e-12)


@pytest.mark.parametrize("m,n,on_rows", [(1, 1, 0), (7, 3, 0), (3, 7, 1), (64, 64, 0),
                                         (64, 64, 1), (200, 17, 0), (17, 200, 1),
                                         (256, 256, 1), (300, 256, 0), (4096, 8, 0),
                                         (129, 128, 0)])
def test_jacobi_svd_random(gpu, m, n, on_rows):
    print(f"jacobi svd {m}x{n} on_rows={on_rows}", flush=True)
    lib = _lib.load()
    A = _rand(np.random.default_rng(m * 1000 + n), m, n)
    _check(A, *_run(lib, A, on_rows))


def test_jacobi_svd_graded_and_rank_deficient(gpu):
    print("jacobi svd graded / rank deficient", flush=True)
    lib = _lib.load()
    rng = np.random.default_rng(5)
    n = 96
    A = _rand(rng, n, n, np.logspace(0, -14, n))
    # the QR-preconditioned form: rows of the triangle R
    R = np.linalg.qr(A)[1]
    s, U, Vh = _run(lib, R, 1)
    _check(R, s, U, Vh)
    ref = np.linalg.svd(R, compute_uv=False)
    # singular values above the 1e-10 cut agree to high relative accuracy
    big = ref > 1e-10 * ref[0]
    np.testing.assert_allclose(s[big], ref[big], rtol=1e-9)
    B = _rand(rng, 80, 40)
    B[:, 10:20] = 0.0                      # ten zero columns
    s, U, Vh = _run(lib, B, 0)
    _check(B, s, U, Vh)
    assert np.all(s[30:] == 0.0)


def test_jacobi_svd_batched_mixed_shapes(gpu):
    print("jacobi svd batched, mixed shapes", flush=True)
    lib = _lib.load()
    rng = np.random.default_rng(7)
    shapes = [(40, 12, 0), (12, 40, 1), (128, 128, 1), (5, 5, 0), (256, 33, 0)] * 3
    mats = [_rand(rng, m, n) for m, n, _ in shapes]
    outs = []
    keep = []
    m_a, n_a, lda, rows, A_p, S_p, U_p, V_p, W_p, I_p = ([] for _ in range(10))
    for (m, n, r), A in zip(shapes, mats):
        k = min(m, n)
        dA = _dev_colmajor(A)
        S = torch.empty(k, dtype=torch.float64, device="cuda")
        U = torch.empty(m * k, dtype=torch.complex128, device="cuda")
        Vh = torch.empty(k * n, dtype=torch.complex128, device="cuda")
        ws = torch.empty(int(lib.ptn_svd_jacobi_work_bytes(m, n)), dtype=torch.uint8,
                         device="cuda")
        info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
        keep += [dA, ws]
        outs.append((S, U, Vh, info))
        m_a.append(m); n_a.append(n); lda.append(m); rows.append(r)
        for lst, t in ((A_p, dA), (S_p, S), (U_p, U), (V_p, Vh), (W_p, ws), (I_p, info)):
            lst.append(t.data_ptr())
    cnt = len(shapes)
    work = torch.empty(int(lib.ptn_svd_jacobi_batched_work_bytes(cnt)), dtype=torch.uint8,
                       device="cuda")
    arr = lambda v, dt: np.ascontiguousarray(np.array(v, dtype=dt))
    host = [arr(m_a, np.int64), arr(n_a, np.int64), arr(A_p, np.int64), arr(lda, np.int64),
            arr(rows, np.int32), arr(S_p, np.int64), arr(U_p, np.int64), arr(V_p, np.int64),
            arr(W_p, np.int64), arr(I_p, np.int64)]
    cp = [ctypes.c_void_p(h.ctypes.data) for h in host]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.ptn_svd_jacobi_batched(stream, cnt, *cp, _vp(work)), "ptn_svd_jacobi_batched")
    torch.cuda.synchronize()
    for (m, n, r), A, (S, U, Vh, info) in zip(shapes, mats, outs):
        k = min(m, n)
        assert int(info.item()) == 0
        _check(A, S.cpu().numpy(), U.cpu().numpy().reshape(k, m).T,
               Vh.cpu().numpy().reshape(n, k).T)


def test_jacobi_svd_rejects_beyond_limits(gpu):
    lib = _lib.load()
    t = torch.zeros(1, dtype=torch.complex128, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.ptn_svd_jacobi(stream, 300, 300, _vp(t), 300, 0, _vp(t), _vp(t), _vp(t), _vp(t),
                            _vp(t))
    assert rc == _lib.PTN_ERANGE
    rc = lib.ptn_svd_jacobi(stream, 10, 20, _vp(t), 10, 0, _vp(t), _vp(t), _vp(t), _vp(t),
                            _vp(t))
    assert rc == _lib.PTN_EINVAL
