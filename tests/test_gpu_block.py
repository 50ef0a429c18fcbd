"""The passes of the block PCG recurrence (dfmi_block_gram / _combine / _update, csrc/block.hip) against
NumPy in FP64: Gram matrices over one to three 16-column tiles with and without a second right operand,
DOF counts off the 16-DOF block and the 256-row chunk, combinations in place and out of place, and the paired
update with its column norms.  FP32 operands; the sums are FP32 inside a wave's share of the DOFs, so the
tolerance is relative to the sum of absolute products."""
from ctypes import byref, c_int64

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _work(torch, n, ka, kb):
    from deep_frame_amd import _native
    count = c_int64()
    _native.call("dfmi_block_work_doubles", n, ka, kb, byref(count))
    return torch.empty(count.value, dtype=torch.float64, device="cuda")


def _stream(torch):
    from deep_frame_amd import _native
    return _native.stream_handle(torch)


def _coef(torch, m):
    """FP32 device matrix with row stride 48 (zero past the used columns)."""
    out = np.zeros((m.shape[0], 48), np.float32)
    out[:, :m.shape[1]] = m
    return torch.as_tensor(out, device="cuda")


@pytest.mark.parametrize("ka,kb1,kb2,n", [(42, 42, 42, 100003), (42, 42, 0, 65537), (20, 17, 32, 4099),
                                          (7, 7, 0, 15), (1, 1, 1, 300), (48, 48, 48, 257), (33, 40, 0, 1 << 16)])
def test_block_gram_matches_numpy(torch_gpu, ka, kb1, kb2, n):
    from deep_frame_amd import _native
    torch = torch_gpu
    rng = np.random.default_rng(ka * 1000 + kb1 * 10 + kb2)
    a = rng.normal(size=(ka, n)).astype(np.float32)
    b1 = rng.normal(size=(kb1, n)).astype(np.float32)
    b2 = rng.normal(size=(max(kb2, 1), n)).astype(np.float32)
    ta, tb1, tb2 = (torch.as_tensor(v, device="cuda") for v in (a, b1, b2))
    out = torch.empty((ka, kb1 + kb2), dtype=torch.float64, device="cuda")
    work = _work(torch, n, ka, kb1 + kb2)
    _native.call("dfmi_block_gram", _native.ptr(ta), ka, _native.ptr(tb1), kb1,
                 _native.ptr(tb2) if kb2 else None, kb2, n, _native.ptr(out), _native.ptr(work), _stream(torch))
    got = out.cpu().numpy()
    a64 = a.astype(np.float64)
    want = a64 @ np.concatenate([b1, b2[:kb2]]).astype(np.float64).T
    scale = np.abs(a64) @ np.abs(np.concatenate([b1, b2[:kb2]]).astype(np.float64)).T
    assert np.max(np.abs(got - want) / scale) < 2e-6
    # deterministic: a second call gives the same bits
    again = torch.empty_like(out)
    _native.call("dfmi_block_gram", _native.ptr(ta), ka, _native.ptr(tb1), kb1,
                 _native.ptr(tb2) if kb2 else None, kb2, n, _native.ptr(again), _native.ptr(work), _stream(torch))
    assert torch.equal(out, again)


def test_block_gram_rejects_mixed_tile_counts(torch_gpu):
    from deep_frame_amd import _native
    torch = torch_gpu
    t = torch.zeros((48, 64), dtype=torch.float32, device="cuda")
    out = torch.empty(48 * 48, dtype=torch.float64, device="cuda")
    work = _work(torch, 64, 48, 48)
    with pytest.raises(RuntimeError, match="tile count"):
        _native.call("dfmi_block_gram", _native.ptr(t), 40, _native.ptr(t), 10, None, 0, 64, _native.ptr(out),
                     _native.ptr(work), _stream(torch))


@pytest.mark.parametrize("kin,kout,n,inplace", [(42, 42, 100003, False), (42, 42, 4097, True), (5, 33, 1000, False),
                                                (48, 48, 257, True), (1, 1, 7, False)])
def test_block_combine_matches_numpy(torch_gpu, kin, kout, n, inplace):
    from deep_frame_amd import _native
    torch = torch_gpu
    rng = np.random.default_rng(kin * 100 + kout)
    u = rng.normal(size=(kin, n)).astype(np.float32)
    x = rng.normal(size=(kout, n)).astype(np.float32)
    m = rng.normal(size=(kin, kout)).astype(np.float32)
    tu, tx = torch.as_tensor(u, device="cuda"), torch.as_tensor(x, device="cuda")
    ty = tx if inplace else torch.full((kout, n), np.nan, dtype=torch.float32, device="cuda")
    _native.call("dfmi_block_combine", _native.ptr(tu), _native.ptr(tx), _native.ptr(ty), _native.ptr(_coef(torch, m)),
                 kin, kout, n, _stream(torch))
    want = x.astype(np.float64) + m.astype(np.float64).T @ u.astype(np.float64)
    scale = np.abs(x) + np.abs(m.astype(np.float64)).T @ np.abs(u.astype(np.float64))
    assert np.max(np.abs(ty.cpu().numpy() - want) / scale) < 1e-6


@pytest.mark.parametrize("k,n", [(42, 100003), (20, 513), (48, 65536)])
def test_block_update_matches_numpy(torch_gpu, k, n):
    from deep_frame_amd import _native
    torch = torch_gpu
    rng = np.random.default_rng(k)
    v, w, d, r = (rng.normal(size=(k, n)).astype(np.float32) for _ in range(4))
    m = (0.1 * rng.normal(size=(k, k))).astype(np.float32)
    tv, tw, td, tr = (torch.as_tensor(a, device="cuda") for a in (v, w, d, r))
    rr = torch.empty(k, dtype=torch.float64, device="cuda")
    work = _work(torch, n, k, k)
    _native.call("dfmi_block_update", _native.ptr(tv), _native.ptr(td), _native.ptr(tw), _native.ptr(tr),
                 _native.ptr(_coef(torch, m)), k, n, _native.ptr(rr), _native.ptr(work), _stream(torch))
    m64 = m.astype(np.float64)
    want_d = d + m64.T @ v.astype(np.float64)
    want_r = r - m64.T @ w.astype(np.float64)
    sd = np.abs(d) + np.abs(m64).T @ np.abs(v.astype(np.float64))
    sr = np.abs(r) + np.abs(m64).T @ np.abs(w.astype(np.float64))
    got_r = tr.cpu().numpy()
    assert np.max(np.abs(td.cpu().numpy() - want_d) / sd) < 1e-6
    assert np.max(np.abs(got_r - want_r) / sr) < 1e-6
    # squares summed in FP32 per lane, FP64 across lanes and workgroups
    np.testing.assert_allclose(rr.cpu().numpy(), np.sum(got_r.astype(np.float64) ** 2, axis=1), rtol=1e-6)


@pytest.mark.parametrize("k,n", [(42, 100003), (20, 4097), (7, 300), (48, 65536), (1, 5)])
def test_block_combine2_matches_numpy(torch_gpu, k, n):
    """The fused pass of the recurrence: Y0 = X0 + U M0 out of place (the next search block) and Y1 = X1 + U M1 in
    place (the deferred solution update), against NumPy in FP64; U, X0 and X1 are left unchanged where not
    written."""
    from deep_frame_amd import _native
    torch = torch_gpu
    rng = np.random.default_rng(100 + k)
    u, x0, x1 = (rng.normal(size=(k, n)).astype(np.float32) for _ in range(3))
    m0, m1 = ((0.2 * rng.normal(size=(k, k))).astype(np.float32) for _ in range(2))
    tu, tx0, tx1 = (torch.as_tensor(a, device="cuda") for a in (u, x0, x1))
    ty0 = torch.full((k, n), np.nan, dtype=torch.float32, device="cuda")
    tm0, tm1 = _coef(torch, m0), _coef(torch, m1)  # both alive across the call
    _native.call("dfmi_block_combine2", _native.ptr(tu), _native.ptr(tx0), _native.ptr(ty0), _native.ptr(tm0),
                 _native.ptr(tx1), _native.ptr(tx1), _native.ptr(tm1), k, n, _stream(torch))
    u64 = u.astype(np.float64)
    for got, x, m in ((ty0, x0, m0), (tx1, x1, m1)):
        want = x + m.astype(np.float64).T @ u64
        scale = np.abs(x) + np.abs(m.astype(np.float64)).T @ np.abs(u64)
        assert np.max(np.abs(got.cpu().numpy() - want) / scale) < 1e-6
    assert np.array_equal(tu.cpu().numpy(), u) and np.array_equal(tx0.cpu().numpy(), x0)


@pytest.mark.parametrize("k,n", [(42, 100003), (20, 4097), (7, 300), (1, 5)])
def test_block_combine_flush_matches_numpy(torch_gpu, k, n):
    """The reliable update of the block recurrence in one pass: D += U M on every column; the flagged columns then
    move into the FP64 solution (x += s D) and restart at D = 0, the others keep x unchanged (bit for bit)."""
    from deep_frame_amd import _native
    torch = torch_gpu
    rng = np.random.default_rng(200 + k)
    u, d = (rng.normal(size=(k, n)).astype(np.float32) for _ in range(2))
    m = (0.2 * rng.normal(size=(k, k))).astype(np.float32)
    x = rng.normal(size=(k, n))
    sc = rng.uniform(0.5, 2.0, size=k)
    flag = (rng.uniform(size=k) < 0.5).astype(np.int32)
    flag[0] = 1
    if k > 1:
        flag[-1] = 0
    tu, td, tx = (torch.as_tensor(a, device="cuda") for a in (u, d, x))
    ts, tf = torch.as_tensor(sc, device="cuda"), torch.as_tensor(flag, device="cuda")
    tm = _coef(torch, m)
    _native.call("dfmi_block_combine_flush", _native.ptr(tu), _native.ptr(td), _native.ptr(tm), k, n,
                 _native.ptr(tx), _native.ptr(ts), _native.ptr(tf), _stream(torch))
    u64 = u.astype(np.float64)
    sums = d + m.astype(np.float64).T @ u64
    scale = np.abs(d) + np.abs(m.astype(np.float64)).T @ np.abs(u64)
    got_d, got_x = td.cpu().numpy(), tx.cpu().numpy()
    on = flag.astype(bool)
    assert np.all(got_d[on] == 0.0)
    assert np.max(np.abs(got_d[~on] - sums[~on]) / scale[~on], initial=0.0) < 1e-6
    assert np.array_equal(got_x[~on], x[~on])
    want_x = x[on] + sc[on, None] * sums[on]
    assert np.max(np.abs(got_x[on] - want_x) / (sc[on, None] * scale[on])) < 1e-6
    assert np.array_equal(tu.cpu().numpy(), u)


def test_block_combine_flush_rejects_aliased_input(torch_gpu):
    from deep_frame_amd import _native
    torch = torch_gpu
    t = torch.zeros((4, 64), dtype=torch.float32, device="cuda")
    x = torch.zeros((4, 64), dtype=torch.float64, device="cuda")
    sc = torch.ones(4, dtype=torch.float64, device="cuda")
    fl = torch.ones(4, dtype=torch.int32, device="cuda")
    m = _coef(torch, np.zeros((4, 4), np.float32))
    with pytest.raises(RuntimeError, match="alias"):
        _native.call("dfmi_block_combine_flush", _native.ptr(t), _native.ptr(t), _native.ptr(m), 4, 64,
                     _native.ptr(x), _native.ptr(sc), _native.ptr(fl), _stream(torch))


def test_block_combine2_rejects_aliased_input(torch_gpu):
    from deep_frame_amd import _native
    torch = torch_gpu
    t = torch.zeros((4, 64), dtype=torch.float32, device="cuda")
    m = _coef(torch, np.zeros((4, 4), np.float32))
    with pytest.raises(RuntimeError, match="alias"):
        _native.call("dfmi_block_combine2", _native.ptr(t), _native.ptr(t), _native.ptr(t), _native.ptr(m), _native.ptr(t),
                     _native.ptr(t), _native.ptr(m), 4, 64, _stream(torch))
