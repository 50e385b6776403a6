"""Engine invariants on the device: recorded launches own what they point at."""
import numpy as np
import pytest
import torch

from pytreenet_amd import engine as E

pytestmark = pytest.mark.gpu


def _crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to("cuda")


def _flood_cache(n):
    """Push n fresh offset-table entries through the LRU cache."""
    for i in range(n):
        E._CACHE.get(("flood", i), lambda i=i: [np.arange(3 + i % 5, dtype=np.int64)], "cuda")


def test_recorded_tables_survive_cache_eviction(gpu):
    """A recorded network and a cached block-sparse group keep their offset tables:
    evicting every cache entry and releasing torch's cached blocks must not change
    a replay (before the fix such replays read freed tables)."""
    print("[engine] table ownership", flush=True)
    rng = np.random.default_rng(4)
    cap = E._CACHE.cap
    E._CACHE.cap = 4
    try:
        x = _dev(_crand(rng, (16, 12, 2)))
        env = _dev(_crand(rng, (16, 5, 16)))
        W = np.zeros((5, 5, 2, 2), dtype=complex)
        W[0, 0] = np.eye(2)
        W[1, 2] = _crand(rng, (2, 2))
        W[3, 4] = _crand(rng, (2, 2))
        Wb = E.BlockSparse.from_dense(_dev(W), [0, 1])
        labels = [["a", "c", "p"], ["a", "w", "b"], ["w", "v", "q", "p"]]
        out = ["b", "c", "v", "q"]
        net = E.CompiledNetwork([x, env, Wb], labels, out)
        first = net(x).cpu().numpy()
        ref_block = E.contract_network([x, env, Wb], labels, out).cpu().numpy()
        for rep in range(3):
            _flood_cache(16)
            torch.cuda.empty_cache()
            y = _dev(_crand(rng, (16, 12, 2)))
            got = net(y).cpu().numpy()
            ref = E.contract_network([y, env, _dev(W)], labels, out).cpu().numpy()
            np.testing.assert_allclose(got, ref, atol=1e-12 * np.abs(ref).max())
            again = E.contract_network([x, env, Wb], labels, out).cpu().numpy()
            np.testing.assert_allclose(again, ref_block, atol=1e-12 * np.abs(ref_block).max())
        assert np.abs(first).max() > 0
    finally:
        E._CACHE.cap = cap


@pytest.mark.parametrize("blocked", [False, True])
def test_cached_network_replay_matches_direct(gpu, blocked):
    """contract_network_cached replays a recorded environment-update network: every
    call, with new input values (the ket used twice, once conjugated), equals a
    fresh contract_network; a new input layout records a new network."""
    print(f"cached network replay blocked={blocked}", flush=True)
    rng = np.random.default_rng(5 + blocked)
    c = lambda *s: torch.as_tensor(rng.standard_normal(s) + 1j * rng.standard_normal(s)).cuda()
    w = np.zeros((3, 4, 2, 2), dtype=complex)           # (w_parent, w_child, out, in)
    w[0, 0] = np.eye(2)
    w[1, 2] = [[0, 1], [1, 0]]
    w[2, 3] = [[1, 0], [0, -1]]
    W = torch.as_tensor(w).cuda()
    op = E.BlockSparse.from_dense(W, [0, 1]) if blocked else W
    labels = [["a0", "a1", "p"], ["w0", "w1", "q", "p"], ["b0", "b1", "q"], ["a1", "w1", "b1"]]
    out_l = ["a0", "w0", "b0"]
    conj = [False, False, True, False]
    E.clear_caches()
    for it in range(3):
        ket, env = c(6, 5, 2), c(5, 4, 5)
        got = E.contract_network_cached([ket, op, ket, env], labels, out_l, conj=conj)
        ref = E.contract_network([ket, op, ket, env], labels, out_l, conj=conj)
        assert torch.allclose(got, ref, rtol=0, atol=1e-13 * ref.abs().max().item()), it
    assert len(E._NET_CACHE) == 1
    ket2 = c(5, 6, 2).permute(1, 0, 2)                  # same shape, other strides
    env = c(5, 4, 5)
    got = E.contract_network_cached([ket2, op, ket2, env], labels, out_l, conj=conj)
    ref = E.contract_network([ket2, op, ket2, env], labels, out_l, conj=conj)
    assert torch.allclose(got, ref, rtol=0, atol=1e-13 * ref.abs().max().item())
    assert len(E._NET_CACHE) == 2


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
@pytest.mark.parametrize("layout", ["abwp", "bawp"])
def test_block_apply_staged_matches_unstaged_and_dense(gpu, dtype, layout):
    """The operator-tensor step with rows along a 256-long innermost dimension runs
    ptn_block_apply_staged (inputs of each 256-row tile staged through LDS); it
    equals ptn_block_apply and the dense contraction."""
    print(f"block apply staged {dtype} {layout}", flush=True)
    rng = np.random.default_rng(9)
    w = np.zeros((5, 2, 2), dtype=complex)          # (w, out, in): a few live blocks
    w[0] = [[1, 0], [0, -1]]
    w[2] = [[0, 1], [1, 0]]
    w[4] = [[0.5, -1j], [1j, 0.25]]
    W = torch.as_tensor(w).cuda()
    Wb = E.BlockSparse.from_dense(W, [0])
    dims = {"a": 16, "b": 256, "w": 5, "p": 2}
    y = rng.standard_normal([dims[c] for c in layout]) + \
        1j * rng.standard_normal([dims[c] for c in layout])
    Y = torch.as_tensor(y).to(dtype).cuda()
    labels = [list(layout), ["w", "q", "p"]]
    out_l = ["a", "b", "q"]
    prec = E.FP64 if dtype == torch.complex128 else E.FP32
    od = torch.complex128 if dtype == torch.complex128 else torch.complex64
    try:
        E.USE_BLOCK_STAGED = True
        E._BAPPLY_CACHE.clear()
        got = E.contract_network([Y, Wb], labels, out_l, precision=prec, out_dtype=od)
        E.USE_BLOCK_STAGED = False
        E._BAPPLY_CACHE.clear()
        ref = E.contract_network([Y, Wb], labels, out_l, precision=prec, out_dtype=od)
    finally:
        E.USE_BLOCK_STAGED = True
        E._BAPPLY_CACHE.clear()
    assert torch.equal(got, ref)
    dense = np.einsum(f"{layout},wqp->abq", Y.cpu().numpy().astype(complex), w)
    tol = 1e-13 if dtype == torch.complex128 else 1e-6
    np.testing.assert_allclose(got.cpu().numpy(), dense, rtol=0, atol=tol * np.abs(dense).max())


def test_network_replay_runs_as_one_native_call(gpu):
    """A recorded network's small gather-GEMMs replay through ptn_gemm_gather_many
    (one call per run of launches): the result equals the per-launch replay and a
    fresh contraction, call after call with new inputs and output buffers; a bad
    descriptor inside a run is reported with its index."""
    import ctypes
    from pytreenet_amd import _lib
    print("[engine] replay through ptn_gemm_gather_many", flush=True)
    rng = np.random.default_rng(6)
    labels = [["a", "b", "p"], ["a", "w", "A"], ["b", "v", "B"], ["w", "v", "q", "p"]]
    out = ["A", "B", "q"]
    envs = [_dev(_crand(rng, (9, 4, 7))), _dev(_crand(rng, (6, 3, 5)))]
    W = _dev(_crand(rng, (4, 3, 2, 2)))
    x = _dev(_crand(rng, (9, 6, 2)))
    net = E.CompiledNetwork([x, envs[0], envs[1], W], labels, out)
    net(x)
    runs = [r for r in net.runs if r[0] == "many"]
    assert runs and sum(r[2] for r in runs) >= 2
    for rep in range(3):
        y = _dev(_crand(rng, (9, 6, 2)))
        res = net(y)
        got = res.cpu().numpy()
        ref = E.contract_network([y, envs[0], envs[1], W], labels, out).cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        per_op = torch.empty_like(res)
        for op, roles in zip(net.ops, net.op_roles):
            net._replay_op(op, roles, (y,), per_op)
        assert np.array_equal(per_op.cpu().numpy(), got)
    # (res and y stay alive: the run's descriptors point at them.)  The second
    # descriptor of a two-launch run carries an unknown compute mode
    arr = (_lib.GemmDesc * 2)()
    arr[0] = runs[0][1][0]
    arr[1] = runs[0][1][0]
    arr[1].compute = 77
    stream = ctypes.c_void_p(E.stream_handle())
    rc = _lib.load().ptn_gemm_gather_many(stream, arr, 2)
    torch.cuda.synchronize()
    assert rc != 0
    assert "descriptor 1" in _lib.load().ptn_last_error().decode()
    assert _lib.load().ptn_gemm_gather_many(stream, arr, 0) == 0


@pytest.mark.parametrize("swap", [False, True])
def test_contract_replays_descriptor_of_same_layout(gpu, swap):
    """A second allocating contract() with the same labels, layouts, dtypes and
    options reuses the first call's launch descriptor (M < N swaps the operands'
    roles): new data give the new product, a different layout or option a new
    descriptor, and a recording network captures the replayed launch."""
    print(f"[engine] contract descriptor replay swap={swap}", flush=True)
    rng = np.random.default_rng(7 + swap)
    sa, sb = ((30, 7, 2), (7, 2, 90)) if swap else ((90, 7, 2), (7, 2, 30))
    E.clear_caches()
    n0 = len(E._FAST)
    for rep in range(3):
        a, b = _crand(rng, sa), _crand(rng, sb)
        got = E.contract(_dev(a), "mkp", _dev(b), "kpn", "nm").cpu().numpy()
        ref = np.einsum("mkp,kpn->nm", a, b)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        assert len(E._FAST) == n0 + 1
        if swap:
            assert next(iter(E._FAST.values()))[3]
    # conjugation is part of the key: a new entry, the conjugated product
    got = E.contract(_dev(a), "mkp", _dev(b), "kpn", "nm", conj_a=True).cpu().numpy()
    ref = np.einsum("mkp,kpn->nm", a.conj(), b)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    assert len(E._FAST) == n0 + 2
    # a network recorded through contract() replays correctly with new inputs
    x = _dev(_crand(rng, sa))
    net = E.CompiledNetwork([x, _dev(b)], [["m", "k", "p"], ["k", "p", "n"]], ["n", "m"])
    net(x)
    y = _crand(rng, sa)
    got = net(_dev(y)).cpu().numpy()
    ref = np.einsum("mkp,kpn->nm", y, b)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
