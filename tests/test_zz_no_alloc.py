"""The allocation-free orthonormalisation step against the one with a caller-owned
work buffer, through the C ABI: two calls in a row on one GPU, the second with a
shorter vector and fewer basis vectors, each bit-identical to the allocating entry
point on the same inputs (same kernels and grid)."""
import ctypes

import pytest
import torch

from pytreenet_amd import _lib

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _case(lib, stream, n, nvec, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    # an orthonormal basis (rows), as the Krylov recurrence holds it
    Q = torch.linalg.qr(torch.randn((n, nvec), dtype=torch.complex128, generator=g))[0]
    Q = Q.T.contiguous().cuda()
    w0 = torch.randn(n, dtype=torch.complex128, generator=g).cuda()
    # reference: the allocating entry point with a torch-owned work buffer
    work = torch.empty(int(lib.ptn_krylov_work_bytes(n, nvec)), dtype=torch.uint8, device="cuda")
    w_ref, q_ref = w0.clone(), torch.empty_like(w0)
    out_ref = torch.empty(nvec + 1, dtype=torch.complex128, device="cuda")
    _lib.check(lib.ptn_krylov_orthonormalize(stream, n, _vp(Q), Q.stride(0), nvec, _vp(w_ref),
                                             _vp(q_ref), _vp(out_ref), _vp(work)),
               "ptn_krylov_orthonormalize")
    w, q = w0.clone(), torch.empty_like(w0)
    out = torch.empty(nvec + 1, dtype=torch.complex128, device="cuda")
    _lib.check(lib.ptn_krylov_orthonormalize_noalloc(stream, n, _vp(Q), Q.stride(0), nvec,
                                                     _vp(w), _vp(q), _vp(out)),
               "ptn_krylov_orthonormalize_noalloc")
    torch.cuda.synchronize()
    assert torch.equal(out, out_ref)
    assert torch.equal(w, w_ref)
    assert torch.equal(q, q_ref)
    # and it is an orthonormalisation: q is a unit vector orthogonal to Q
    assert abs(torch.linalg.vector_norm(q).item() - 1.0) < 1e-12
    assert (Q.conj() @ q).abs().max().item() < 1e-10 * n ** 0.5


def test_noalloc_orthonormalize_matches_allocating_entry_point(gpu):
    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("noalloc orthonormalize: n=300000 nvec=20, then n=50000 nvec=5", flush=True)
    _case(lib, stream, 300000, 20, 1)
    _case(lib, stream, 50000, 5, 2)
    _case(lib, stream, 4000, 3, 3)           # the one-workgroup form (no scratch used)


def test_noalloc_orthonormalize_rejects_too_many_vectors(gpu):
    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = 10000
    Q = torch.zeros((1, n), dtype=torch.complex128, device="cuda")
    w = torch.zeros(n, dtype=torch.complex128, device="cuda")
    out = torch.empty(130, dtype=torch.complex128, device="cuda")
    rc = lib.ptn_krylov_orthonormalize_noalloc(stream, n, _vp(Q), n, 129, _vp(w),
                                               ctypes.c_void_p(0), _vp(out))
    assert rc == _lib.PTN_ERANGE
    assert b"beyond the maximum" in lib.ptn_last_error()
