"""One-launch Krylov exponential of short vectors (ptn_krylov_expm_small) against the
oracle's krylov_exp_action on the same dense operator and against the exact
exponential."""
import ctypes

import numpy as np
import pytest
import torch
from scipy.linalg import expm

from oracle.krylov import krylov_exp_action as oracle_krylov
from pytreenet_amd import _lib
from pytreenet_amd import krylov as K

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _herm(rng, n, scale):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = 0.5 * (a + a.conj().T)
    return scale * h / max(np.linalg.norm(h, 2), 1e-300)


def _run(psi, H, dt, f, tol, maxdim, hermitian):
    out = K.krylov_exp_small(torch.as_tensor(psi).to("cuda"), torch.as_tensor(H).to("cuda"),
                             dt, f, krylov_tol=tol, max_krylov_dim=maxdim, hermitian=hermitian)
    return out


@pytest.mark.parametrize("n", [1, 2, 4, 17, 64, 256, 1024])
def test_small_krylov_hermitian_matches_oracle(gpu, n):
    _say(f"krylov_small hermitian n={n}")
    rng = np.random.default_rng(n)
    H = _herm(rng, n, 3.0)
    psi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dt, f = 0.1, -1j
    stats = {}
    ref = oracle_krylov(psi, lambda y: H @ y, dt, f, krylov_tol=1e-10, stats=stats)
    calls0 = K.STATS.calls
    out = _run(psi, H, dt, f, 1e-10, None, True)
    assert out is not None
    assert K.STATS.calls == calls0 + 1
    assert K.STATS.dims[-1] == stats["m"]
    got = out.cpu().numpy()
    scale = np.linalg.norm(psi)
    assert np.linalg.norm(got - ref) <= 1e-12 * scale
    exact = expm(f * dt * H) @ psi
    assert np.linalg.norm(got - exact) <= 1e-9 * scale


@pytest.mark.parametrize("n,maxdim", [(64, 8), (4, None), (24, None)])
def test_small_krylov_fixed_depth(gpu, n, maxdim):
    """krylov_tol = 0: the recurrence runs to max_krylov_dim or breakdown."""
    _say(f"krylov_small tol=0 n={n} maxdim={maxdim}")
    rng = np.random.default_rng(100 + n)
    H = _herm(rng, n, 2.0)
    psi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    stats = {}
    ref = oracle_krylov(psi, lambda y: H @ y, 0.2, 1j, krylov_tol=0.0,
                        max_krylov_dim=maxdim, stats=stats)
    out = _run(psi, H, 0.2, 1j, 0.0, maxdim, True)
    assert out is not None
    assert K.STATS.dims[-1] == stats["m"]
    assert np.linalg.norm(out.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(psi)
    if maxdim is None:
        exact = expm(1j * 0.2 * H) @ psi
        assert np.linalg.norm(out.cpu().numpy() - exact) <= 1e-12 * np.linalg.norm(psi)


def test_small_krylov_non_hermitian_imaginary_time(gpu):
    _say("krylov_small non-hermitian")
    rng = np.random.default_rng(7)
    n = 48
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A /= np.linalg.norm(A, 2)
    psi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    stats = {}
    ref = oracle_krylov(psi, lambda y: A @ y, 0.3, -1.0, krylov_tol=1e-10, hermitian=False,
                        stats=stats)
    out = _run(psi, A, 0.3, -1.0, 1e-10, None, False)
    assert out is not None
    assert K.STATS.dims[-1] == stats["m"]
    assert np.linalg.norm(out.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(psi)
    exact = expm(-0.3 * A) @ psi
    assert np.linalg.norm(out.cpu().numpy() - exact) <= 1e-9 * np.linalg.norm(psi)


def test_small_krylov_zero_vector(gpu):
    _say("krylov_small zero")
    H = _herm(np.random.default_rng(1), 8, 1.0)
    out = _run(np.zeros(8, dtype=complex), H, 0.1, -1j, 1e-10, None, True)
    assert out is not None and torch.count_nonzero(out).item() == 0


def test_small_krylov_defers_deep_solves(gpu):
    """A solve that needs more than SMALL_MAX_STEPS steps is handed back (None)
    and TimeEvoMode.KRYLOV.time_evolve then gives the general form's result."""
    _say("krylov_small fallback")
    rng = np.random.default_rng(3)
    n = 200
    H = _herm(rng, n, 40.0)
    psi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert _run(psi, H, 1.0, -1j, 1e-12, None, True) is None
    stats = {}
    ref = oracle_krylov(psi, lambda y: H @ y, 1.0, -1j, krylov_tol=1e-12, stats=stats)
    assert stats["m"] > K.SMALL_MAX_STEPS
    out = K.TimeEvoMode.KRYLOV.time_evolve(torch.as_tensor(psi).to("cuda"),
                                           torch.as_tensor(H).to("cuda"), 1.0,
                                           forward=True, krylov_tol=1e-12)
    assert np.linalg.norm(out.cpu().numpy() - ref) <= 1e-9 * np.linalg.norm(psi)


def test_small_krylov_rejects_out_of_range(gpu):
    _say("krylov_small range")
    lib = _lib.load()
    n = 2048
    H = torch.zeros((n, n), dtype=torch.complex128, device="cuda")
    x = torch.ones(n, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    work = torch.empty(n * 25, dtype=torch.complex128, device="cuda")
    info = torch.empty(3, dtype=torch.int32, device="cuda")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = lib.ptn_krylov_expm_small(None, n, vp(H), vp(x), vp(y), 24, 0.1, 0.0, -1.0, 1e-10,
                                   1, vp(work), vp(info))
    assert rc == -5
    rc = lib.ptn_krylov_expm_small(None, 64, vp(H), vp(x), vp(y), 25, 0.1, 0.0, -1.0, 1e-10,
                                   1, vp(work), vp(info))
    assert rc == -5


def test_dense_operator_transposed(gpu):
    """dense_operator(transposed=True) is the transpose of the row-major matrix,
    and both act like the network on a vector."""
    from pytreenet_amd import engine as E
    _say("dense_operator transposed")
    rng = np.random.default_rng(5)
    c = lambda *s: torch.as_tensor(rng.standard_normal(s) + 1j * rng.standard_normal(s)).cuda()
    r, ea, eb = c(3, 5), c(3, 4, 6), c(5, 4, 2)
    net = E.CompiledNetwork([r, ea, eb], [["a0", "a1"], ["a0", "w", "b0"], ["a1", "w", "b1"]],
                            ["b0", "b1"])
    Hr = E.dense_operator(net)
    Ht = E.dense_operator(net, transposed=True)
    assert Hr.shape == (12, 15) and Ht.shape == (15, 12)
    assert torch.allclose(Ht, Hr.t(), rtol=1e-13, atol=1e-13)
    v = c(3, 5)
    assert torch.allclose(Hr @ v.reshape(-1), net(v).reshape(-1), rtol=1e-12, atol=1e-12)
