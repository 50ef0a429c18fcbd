"""dfmi_gram_static, the allocation-free Gram entry point, against dfmi_gram through the C ABI: two calls in
a row on the same GPU (the second smaller, so it reuses the scratch array that the first one filled), each
bit-identical to dfmi_gram with the same 256-workgroup partition, and sizes beyond the scratch are refused."""
import ctypes

import numpy as np
import pytest

from deep_frame_amd import _native

pytestmark = pytest.mark.gpu


def _gram(torch, a, b, static):
    ka, n = a.shape
    kb = b.shape[0]
    out = torch.full((ka * kb,), np.nan, dtype=torch.float64, device="cuda")
    stream = _native.stream_handle(torch)
    if static:
        _native.call("dfmi_gram_static", _native.ptr(a), _native.ptr(b), n, ka, kb, _native.ptr(out), stream)
    else:
        work = torch.empty(256 * ka * kb, dtype=torch.float64, device="cuda")
        _native.call("dfmi_gram", _native.ptr(a), _native.ptr(b), n, ka, kb, _native.ptr(out), _native.ptr(work),
                     256, stream)
    torch.cuda.synchronize()
    return out.reshape(ka, kb).cpu().numpy()


def test_static_gram_matches_gram_on_two_calls(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(0)
    for ka, kb, n in ((42, 42, 1_000_003), (5, 9, 70_001)):  # the second call is smaller
        print(f"gram_static ka={ka} kb={kb} n={n}", flush=True)
        a = torch.as_tensor(rng.normal(size=(ka, n)), device="cuda")
        b = torch.as_tensor(rng.normal(size=(kb, n)), device="cuda")
        got = _gram(torch, a, b, static=True)
        ref = _gram(torch, a, b, static=False)
        assert np.array_equal(got, ref)
        host = a.cpu().numpy() @ b.cpu().numpy().T
        assert np.max(np.abs(got - host)) < 1e-10 * np.abs(host).max()
    # 64 x 64 (the vector-unit path) and a one-column case
    for ka, kb, n in ((64, 64, 20_000), (1, 1, 513)):
        a = torch.as_tensor(rng.normal(size=(ka, n)), device="cuda")
        b = torch.as_tensor(rng.normal(size=(kb, n)), device="cuda")
        assert np.array_equal(_gram(torch, a, b, static=True), _gram(torch, a, b, static=False))


def test_static_gram_refuses_sizes_beyond_the_scratch(torch_gpu):
    torch = torch_gpu
    a = torch.zeros((65, 16), dtype=torch.float64, device="cuda")
    out = torch.zeros(65 * 65, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="1 <= ka, kb <= 64"):
        _native.call("dfmi_gram_static", _native.ptr(a), _native.ptr(a), 16, 65, 65, _native.ptr(out),
                     _native.stream_handle(torch))
    with pytest.raises(RuntimeError, match="null operand"):
        _native.call("dfmi_gram_static", None, ctypes.c_void_p(0), 16, 1, 1, _native.ptr(out),
                     _native.stream_handle(torch))
