"""Learned coarse correction for the fp32 V-cycle: host side of the U-Net in libdfmi.

No counterpart in the reference.  The nearest pattern is its hand-written NumPy MLP with
forward, backward and Adam (deep_frame/topology_neural.py:59-113).  Here forward, backward
and Adam are hand-written HIP kernels (BF16 weights, FP32 accumulation), and training data are
pairs that the solver itself produces: for each fine V-cycle, the level-1 residual g left after
the coarse-grid correction of a density field's hierarchy.  The loss is the energy of the
prolonged correction, 0.5 delta.A_1 delta - delta.g, which is exact because A_1 = P^T A P.

Inside flexible PCG the correction is scaled per column by the energy-optimal step
max(0, delta.g / delta.A_1 delta).  So it never increases the A-norm error of the V-cycle's
iterate, and plain GMG-PCG (no network attached) remains both the fallback and the reference
for iteration counts.
"""
import ctypes
from ctypes import byref, c_double, c_int32, c_int64, c_void_p

import numpy as np

from deep_frame_amd import _native

# (name, kind, c_in, c_out); up layers store weights [c_coarse][c_fine][27]
LAYERS = [
    ("enc1a", "conv1", 6, 16), ("enc1b", "conv1", 16, 16),
    ("enc2a", "conv2", 16, 32), ("enc2b", "conv1", 32, 32),
    ("enc3a", "conv2", 32, 64), ("enc3b", "conv1", 64, 64),
    ("enc4a", "conv2", 64, 64), ("enc4b", "conv1", 64, 64),
    ("up3", "up", 64, 64), ("dec3", "conv1", 64, 32),
    ("up2", "up", 32, 32), ("dec2", "conv1", 32, 16),
    ("up1", "up", 16, 16), ("out", "conv1", 16, 3),
]


def _torch():
    import torch
    return torch


def initial_params(seed=0):
    """Flat FP32 parameters: He-normal weights, zero biases, zero output layer (the untrained network
    is then the Jacobi correction delta = Dinv g)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, kind, cin, cout in LAYERS:
        fan = cin * 27 / (8 if kind == "up" else 1)
        w = rng.normal(0.0, np.sqrt(2.0 / fan), cin * cout * 27)
        if name == "out":
            w[:] = 0.0
        parts += [w, np.zeros(cout)]
    return np.concatenate(parts).astype(np.float32)


def unflatten(flat):
    """Dict name -> {'w': array, 'b': array} in the layout of the flat parameter vector."""
    flat = np.asarray(flat)
    out, off = {}, 0
    for name, kind, cin, cout in LAYERS:
        shape = (cin, cout, 27) if kind == "up" else (cout, cin, 27)
        n = cin * cout * 27
        out[name] = {"w": flat[off:off + n].reshape(shape), "b": flat[off + n:off + n + cout]}
        off += n + cout
    return out


class CoarseCorrection:
    """The network, its parameters (device) and its training state."""

    def __init__(self, params=None, seed=0):
        torch = self.torch = _torch()
        handle = c_void_p()
        _native.call("dfmi_net_create", byref(handle))
        self.handle = handle
        count = c_int64()
        _native.call("dfmi_net_param_count", self.handle, byref(count))
        self.count = count.value
        self.set_params(initial_params(seed) if params is None else params)
        self.history = []

    def _stream(self):
        _native.call("dfmi_net_set_stream", self.handle, _native.stream_handle(self.torch))

    def set_params(self, params):
        t = self.torch.as_tensor(np.asarray(params, np.float32) if not isinstance(params, self.torch.Tensor) else params,
                                 dtype=self.torch.float32, device="cuda").contiguous()
        if t.numel() != self.count:
            raise ValueError(f"expected {self.count} parameters, got {t.numel()}")
        self._stream()
        _native.call("dfmi_net_set_params", self.handle, _native.ptr(t))

    def _get(self, which):
        out = self.torch.empty(self.count, dtype=self.torch.float32, device="cuda")
        self._stream()
        _native.call("dfmi_net_get", self.handle, which, _native.ptr(out))
        self.torch.cuda.synchronize()
        return out

    def params(self):
        return self._get(0)

    def grads(self):
        return self._get(1)

    def forward(self, x):
        """u (n, 3, X, Y, Z) = N(x (n, 6, X, Y, Z)), FP32 device tensors."""
        x = x.to(device="cuda", dtype=self.torch.float32).contiguous()
        n, c, *dims = x.shape
        if c != 6:
            raise ValueError("the network takes 6 input channels")
        u = self.torch.empty((n, 3, *dims), dtype=self.torch.float32, device="cuda")
        self._stream()
        _native.call("dfmi_net_forward", self.handle, _native.ptr(x), n, (c_int32 * 3)(*dims), _native.ptr(u))
        return u

    def backward(self, du):
        du = du.to(device="cuda", dtype=self.torch.float32).contiguous()
        self._stream()
        _native.call("dfmi_net_backward", self.handle, _native.ptr(du))

    def train_step(self, mg, key, samples, lr=1e-3):
        """One Adam step on harvested level-1 residuals samples (n, 3 * nodes_1) of hierarchy `key`."""
        slot = mg.hierarchies[key]["slot"]
        samples = samples.to(device="cuda", dtype=self.torch.float32).contiguous()
        loss = c_double()
        mg._stream()
        _native.call("dfmi_net_train_step", self.handle, mg.handle, slot, _native.ptr(samples), samples.shape[0],
                     float(lr), byref(loss))
        return loss.value

    def attach(self, mg):
        _native.call("dfmi_attach_net", mg.handle, self.handle)
        mg.net = self

    @staticmethod
    def detach(mg):
        _native.call("dfmi_attach_net", mg.handle, None)
        mg.net = None

    def save(self, path):
        np.savez(path, params=self.params().cpu().numpy())

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as f:
            return cls(params=f["params"])

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            _native.call("dfmi_net_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def level1_nodes(mg):
    """Node count of the level-1 grid (padded fine grid coarsened once)."""
    return int(np.prod([p // 2 + 1 for p in mg.padded]))


def harvest(mg, key, fixed, rhs, capacity):
    """Solve rhs (k, ndof) once with the plain V-cycle and record the level-1 residual of every fine
    V-cycle: returns (samples (count, 3 * nodes_1) FP32 device tensor, solve reports)."""
    torch = _torch()
    buf = torch.zeros((capacity, 3 * level1_nodes(mg)), dtype=torch.float32, device="cuda")
    _native.call("dfmi_harvest", mg.handle, _native.ptr(buf), capacity)
    try:
        x, reports = mg.solve_device(key, fixed, rhs)
    finally:
        count = c_int64()
        _native.call("dfmi_harvest_count", mg.handle, byref(count))
        _native.call("dfmi_harvest", mg.handle, None, 0)
    samples = buf[:count.value]
    keep = samples.abs().amax(dim=1) > 0  # inactive columns were recorded as zeros
    return samples[keep].contiguous(), reports
