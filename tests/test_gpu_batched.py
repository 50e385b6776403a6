"""Batched form: grouped launches and lockstep TDVP over independent instances.

Every batched result is checked against the single-instance path of the same
library (itself pinned to the oracle in the other GPU tests): the kernels
element by element, the TDVP runs through gauge-invariant outputs (energy,
<Z_i>, norm) and identical truncation decisions per instance.
"""
import ctypes

import numpy as np
import pytest
import torch

import pytreenet_amd as ptn
from pytreenet_amd import _lib
from pytreenet_amd import engine as E
from pytreenet_amd import krylov as K
from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
from pytreenet_amd.ttno import pauli_matrices, random_field_heisenberg

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def _say(msg):
    print(f"[batched] {msg}", flush=True)


def _crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to("cuda")


def test_grouped_gemm_mixed_shapes(gpu):
    """One grouped launch over contractions of different shapes equals the single launches."""
    _say("grouped gemm")
    rng = np.random.default_rng(3)
    shapes = [(70, 33, 19), (64, 64, 64), (1, 5, 3), (130, 7, 200)]
    descs, refs, outs = [], [], []
    recs = []
    for (m, n, k) in shapes:
        A = _dev(_crand(rng, (m, k)))
        B = _dev(_crand(rng, (k, n)))
        C = torch.empty((m, n), dtype=torch.complex128, device="cuda")
        rec = E._Recorder()
        E._RECORDER = rec
        try:
            E.USE_BLAS = False
            E.contract(A, "mk", B, "kn", "mn", out=C)
        finally:
            E._RECORDER = None
            E.USE_BLAS = True
        op = rec.ops[0]
        assert op[0] == "gemm"
        recs.append((A, B, C))
        refs.append(C.cpu().numpy().copy())
        C.zero_()
        descs.append(E.GemmDesc.from_buffer_copy(op[1]))
    arr = (E.GemmDesc * len(descs))(*descs)
    work = torch.empty(int(_lib.load().ptn_gemm_grouped_work_bytes(len(descs))),
                       dtype=torch.uint8, device="cuda")
    _lib.check(_lib.load().ptn_gemm_gather_grouped(vp(E.stream_handle()), arr, len(descs),
                                                   vp(work.data_ptr())), "grouped")
    for (A, B, C), ref in zip(recs, refs):
        np.testing.assert_allclose(C.cpu().numpy(), ref, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("n", [300, 20000])
def test_batched_krylov_steps(gpu, n):
    """Batched normalize / orthonormalize / combine equal numpy per instance."""
    _say(f"batched krylov n={n}")
    lib = _lib.load()
    rng = np.random.default_rng(n)
    count, nv = 5, 4
    Qs, ws, xs = [], [], []
    for b in range(count):
        q, _ = np.linalg.qr(_crand(rng, (n, nv)))
        buf = np.zeros((nv + 1, n), dtype=complex)
        buf[:nv] = q.T
        Qs.append(_dev(buf))
        ws.append(_crand(rng, n))
        xs.append(_crand(rng, n))
    tw = [_dev(w) for w in ws]
    out = torch.empty((count, nv + 1), dtype=torch.complex128, device="cuda")
    work = torch.empty(int(lib.ptn_krylov_batched_work_bytes(count)), dtype=torch.uint8,
                       device="cuda")
    P = lambda ts: np.array([t.data_ptr() for t in ts], dtype=np.int64)
    qp, wp = P(Qs), P(tw)
    ld = np.array([n] * count, dtype=np.int64)
    qn = P([Q[nv] for Q in Qs])
    op = np.array([out.data_ptr() + b * out.stride(0) * 16 for b in range(count)],
                  dtype=np.int64)
    _lib.check(lib.ptn_krylov_orthonormalize_batched(
        vp(E.stream_handle()), n, nv, count, vp(qp.ctypes.data), vp(ld.ctypes.data),
        vp(wp.ctypes.data), vp(qn.ctypes.data), vp(op.ctypes.data), vp(work.data_ptr())), "o")
    h = out.cpu().numpy()
    for b in range(count):
        Q = Qs[b].cpu().numpy()[:nv]
        w = ws[b]
        h1 = Q.conj() @ w
        w1 = w - Q.T @ h1
        h2 = Q.conj() @ w1
        w2 = w1 - Q.T @ h2
        sc = np.linalg.norm(w)
        np.testing.assert_allclose(h[b, :nv], h1 + h2, atol=1e-12 * sc)
        assert abs(h[b, nv].real - sc ** 2) < 1e-12 * sc ** 2
        assert abs(h[b, nv].imag - np.vdot(w2, w2).real) < 1e-12 * sc ** 2
        np.testing.assert_allclose(Qs[b].cpu().numpy()[nv], w2 / np.linalg.norm(w2), atol=1e-12)
    # normalize
    tx = [_dev(x) for x in xs]
    ty = [torch.empty_like(t) for t in tx]
    nrm = torch.empty((count, 1), dtype=torch.complex128, device="cuda")
    npp = np.array([nrm.data_ptr() + 16 * b for b in range(count)], dtype=np.int64)
    xp, yp = P(tx), P(ty)
    _lib.check(lib.ptn_normalize_batched(vp(E.stream_handle()), n, count, vp(xp.ctypes.data),
                                         vp(yp.ctypes.data), vp(npp.ctypes.data),
                                         vp(work.data_ptr())), "n")
    for b in range(count):
        nn = np.vdot(xs[b], xs[b]).real
        assert abs(nrm[b, 0].item() - nn) < 1e-12 * nn
        np.testing.assert_allclose(ty[b].cpu().numpy(), xs[b] / np.sqrt(nn), atol=1e-14)
    # combine with per-instance coefficients
    coef = _crand(rng, (count, nv)).astype(np.complex128)
    co = [torch.empty(n, dtype=torch.complex128, device="cuda") for _ in range(count)]
    cp = P(co)
    _lib.check(lib.ptn_krylov_combine_batched(vp(E.stream_handle()), n, nv, count,
                                              vp(qp.ctypes.data), vp(ld.ctypes.data),
                                              vp(coef.ctypes.data), vp(cp.ctypes.data),
                                              vp(work.data_ptr())), "c")
    for b in range(count):
        Q = Qs[b].cpu().numpy()[:nv]
        np.testing.assert_allclose(co[b].cpu().numpy(), coef[b] @ Q, atol=1e-12)


def test_batched_fill_permute(gpu):
    _say("batched fill/permute")
    lib = _lib.load()
    rng = np.random.default_rng(9)
    count = 3
    srcs = [_crand(rng, (4, 5, 6)) for _ in range(count)]
    ts = [_dev(a) for a in srcs]
    ds = [torch.empty((6, 4, 5), dtype=torch.complex128, device="cuda") for _ in range(count)]
    work = torch.empty(int(lib.ptn_ptr_batch_work_bytes(count)), dtype=torch.uint8,
                       device="cuda")
    dims = np.array([4, 5, 6], dtype=np.int64)
    sst = np.array([30, 6, 1], dtype=np.int64)
    dst = np.array([5, 1, 20], dtype=np.int64)
    sp = np.array([t.data_ptr() for t in ts], dtype=np.int64)
    dp = np.array([t.data_ptr() for t in ds], dtype=np.int64)
    _lib.check(lib.ptn_permute_batched(vp(E.stream_handle()), count, 3, vp(dims.ctypes.data),
                                       vp(sp.ctypes.data), vp(sst.ctypes.data), 0,
                                       vp(dp.ctypes.data), vp(dst.ctypes.data), 0, 1,
                                       vp(work.data_ptr())), "perm")
    for a, d in zip(srcs, ds):
        np.testing.assert_array_equal(d.cpu().numpy(), a.conj().transpose(2, 0, 1))
    _lib.check(lib.ptn_fill_batched(vp(E.stream_handle()), count, vp(dp.ctypes.data), 120, 0,
                                    0.5, -2.0, vp(work.data_ptr())), "fill")
    for d in ds:
        assert torch.all(d == (0.5 - 2.0j))


def _instances(count, depth, chis, seed0, two_site=True, precision="fp64", max_bond_dim=12):
    ids, parent, children = complete_tree(2, depth)
    algs = []
    for b in range(count):
        chi = chis[b % len(chis)]
        st = build_state(ids, parent, children,
                         random_state_tensors(ids, parent, children, chi, seed=seed0 + b),
                         device="cuda")
        H = random_field_heisenberg(st, 2.0, seed=seed0 + 100 + b)
        if two_site:
            cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV,
                                        max_bond_dim=max_bond_dim,
                                        rel_tol=1e-10, total_tol=1e-10, precision=precision)
            cls = ptn.SecondOrderTwoSiteTDVP
        else:
            cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision=precision)
            cls = ptn.SecondOrderOneSiteTDVP
        a = cls(st, H, 0.05, 0.1, [], config=cfg,
                solver_options={"max_krylov_dim": 8, "krylov_tol": 1e-10})
        c = a.state.orthogonality_center_id
        a.state.tensors[c] = a.state.tensors[c] / a.state.norm()
        algs.append(a)
    return algs


def _observables(a):
    st = a.state
    e = ptn.expectation_value(st, a.hamiltonian).real
    _, _, Z = pauli_matrices()
    zs = [st.single_site_operator_expectation_value(n, Z).real for n in list(st.nodes)[:4]]
    return np.array([e, st.norm()] + zs)


@pytest.mark.parametrize("two_site,seq_n,group", [(True, None, None), (False, None, None),
                                                   (True, 200, None), (True, None, 2)])
def test_batched_tdvp_matches_single(gpu, two_site, seq_n, group):
    """Lockstep batched TDVP = each instance run alone: observables to 1e-12 and the
    same kept bond dimensions; instances start from two different bond dims (uneven
    shapes, several shape groups).  seq_n: requests with local spaces from this size
    on are solved one instance at a time, the others batched.  group: instances per
    lockstep group (5 instances in groups of 2, 2 and 1)."""
    _say(f"batched tdvp two_site={two_site} seq_n={seq_n} group={group}")
    count, steps = 5, 2
    single = _instances(count, 3, [3, 6], 11, two_site)
    batch = _instances(count, 3, [3, 6], 11, two_site)
    for a in single:
        for _ in range(steps):
            a.run_one_time_step()
    bt = ptn.BatchedTDVP(batch)
    if seq_n is not None:
        bt.SEQUENTIAL_N = seq_n
    bt.lockstep_group = group
    bt.run(steps)
    for a, b in zip(single, batch):
        oa, ob = _observables(a), _observables(b)
        np.testing.assert_allclose(ob, oa, rtol=0, atol=1e-12 * max(1.0, np.abs(oa).max()))
        if two_site:
            assert a.kept_dims == b.kept_dims
        assert [x[2] for x in a.apply_log] == [x[2] for x in b.apply_log]


def test_batched_config4_shape_matches_single(gpu):
    """The batching configuration's instance shape: 31-node binary tree,
    random-field Heisenberg (W=2), second-order two-site TDVP with truncation at
    1e-10 and chi_max=128, starting at bond dimension 8 so the bonds grow unevenly:
    three instances in lockstep equal their single runs after one step."""
    _say("batched config C4 shape")
    single = _instances(3, 4, [8], 41, True, max_bond_dim=128)
    batch = _instances(3, 4, [8], 41, True, max_bond_dim=128)
    for a in single:
        a.run_one_time_step()
    ptn.BatchedTDVP(batch).run(1)
    for a, b in zip(single, batch):
        oa, ob = _observables(a), _observables(b)
        np.testing.assert_allclose(ob, oa, rtol=0, atol=1e-12 * max(1.0, np.abs(oa).max()))
        assert a.kept_dims == b.kept_dims
        assert max(a.kept_dims) > 8


def test_batched_deferred_qr_checks_redo_a_collapsed_group(gpu):
    """The lockstep reads the CholeskyQR pivot statuses of a group once, at the end
    of its step.  One-site instances on a ternary tree at chi 16 (root and level-1
    splits above the one-launch QR size), one of them zero-padded beyond bond
    dimension 3 so its splits collapse: the group is restored and stepped again
    with checked splits, and every instance equals its single run (which redoes
    its own step the same way)."""
    _say("batched deferred QR checks with a collapsing instance")
    from pytreenet_amd.time_evolution import TDVPAlgorithm
    ids, parent, children = complete_tree(3, 2)

    def make(b):
        tensors = random_state_tensors(ids, parent, children, 16, seed=70 + b)
        if b == 1:
            for nid, t in tensors.items():
                t = t.copy()
                for ax in range(t.ndim - 1):
                    idx = [slice(None)] * t.ndim
                    idx[ax] = slice(3, None)
                    t[tuple(idx)] = 0
                tensors[nid] = t
        st = build_state(ids, parent, children, tensors, device="cuda")
        H = random_field_heisenberg(st, 2.0, seed=170 + b)
        cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV)
        a = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                       solver_options={"max_krylov_dim": 8, "krylov_tol": 1e-10})
        return a

    single = [make(b) for b in range(3)]
    batch = [make(b) for b in range(3)]
    for a in single:
        a.run_one_time_step()
    n0 = TDVPAlgorithm.steps_redone
    ptn.BatchedTDVP(batch).run(1)
    assert TDVPAlgorithm.steps_redone > n0
    for a, b in zip(single, batch):
        oa, ob = _observables(a), _observables(b)
        np.testing.assert_allclose(ob, oa, rtol=0, atol=1e-12 * max(1.0, np.abs(oa).max()))
        assert [x[2] for x in a.apply_log] == [x[2] for x in b.apply_log]
