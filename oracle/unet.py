"""Learned coarse correction: a 4-level 3D U-Net on the level-1 node grid (FP64 oracle).

There is no counterpart in the reference. The nearest pattern is its hand-written
NumPy MLP with forward, backward and Adam (deep_frame/topology_neural.py:59-113),
which is checked by finite differences (tests/test_topology_neural.py:15-40).
This module is the CPU restatement of the network that libdfmi runs on the GPU. It
is used only by the tests, to check the GPU forward pass and gradients.

Definitions (shared with the HIP kernels):
  * activations a[n][c][X][Y][Z] over a node grid; a 3x3x3 convolution with stride s
    reads input node s*v + t - 1 (t = 0, 1, 2 per axis; zero outside the grid), so
    stride 2 maps n nodes to (n - 1) // 2 + 1 nodes (the multigrid coarsening);
  * an up layer is the adjoint (transpose) of a stride-2 convolution, taking the
    coarse grid back to the fine grid size it came from;
  * weights are rounded to BF16 (round to nearest even) before use; the master
    copy and the Adam state are FP32 on the GPU.
Layers (name, kind, c_in, c_out): see LAYERS.  The network works in the Jacobi-symmetrised
space: input channels are h / s (3), with h = Dinv^(1/2) g the scaled level-1 residual and s the
per-sample max |h|, and the stiffness feature f (3).  The output u (3 channels) gives the
correction delta = Dinv^(1/2) (h + s u) = Dinv g + Dinv^(1/2) s u, so the untrained network
(zero output layer) is the Jacobi correction and the energy of the learned part,
0.5 s^2 u^T (D^-1/2 A_1 D^-1/2) u, has the scale of u whatever the stiffness contrast.

Loss (energy of the prolonged correction, exact for Galerkin coarse operators,
because P^T A P = A_1):  L = sum_n [0.5 delta_n^T A_1 delta_n - delta_n^T g_n] / (g_n^T Dinv g_n);
the Jacobi correction alone gives L >= -0.5 per sample only when it is a descent step, and the
minimum over all delta is -0.5 g^T A_1^-1 g / g^T Dinv g.
Parity of the forward pass and the gradients is pinned by finite differences of this
module (tests/test_oracle.py); the GPU is checked against this module.
"""
import numpy as np

LAYERS = [
    ("enc1a", "conv1", 6, 16), ("enc1b", "conv1", 16, 16),
    ("enc2a", "conv2", 16, 32), ("enc2b", "conv1", 32, 32),
    ("enc3a", "conv2", 32, 64), ("enc3b", "conv1", 64, 64),
    ("enc4a", "conv2", 64, 64), ("enc4b", "conv1", 64, 64),
    ("up3", "up", 64, 64), ("dec3", "conv1", 64, 32),
    ("up2", "up", 32, 32), ("dec2", "conv1", 32, 16),
    ("up1", "up", 16, 16), ("out", "conv1", 16, 3),
]


def bf16(w):
    """Round float32 values to bfloat16 (nearest even) and back to float32."""
    w = np.asarray(w, np.float32)
    bits = w.view(np.uint32).astype(np.uint64)
    rounding = 0x7FFF + ((bits >> 16) & 1)
    out = ((bits + rounding) >> 16) << 16
    return out.astype(np.uint32).view(np.float32)


def coarse(n):
    return (n - 1) // 2 + 1


def init(seed=0, scale=1.0):
    """He-style initial weights [c_out][c_in][27] (up layers: [c_coarse][c_fine][27]); zero 'out' layer."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, kind, cin, cout in LAYERS:
        if kind == "up":
            w = rng.normal(0, scale * np.sqrt(2.0 / (cin * 27 / 8)), (cin, cout, 27))
        else:
            w = rng.normal(0, scale * np.sqrt(2.0 / (cin * 27)), (cout, cin, 27))
        if name == "out":
            w[:] = 0.0
        params[name] = {"w": w.astype(np.float32), "b": np.zeros(cout, np.float32)}
    return params


def _taps():
    return [(t // 9, (t // 3) % 3, t % 3) for t in range(27)]


def conv(x, w, b, stride):
    """x (N, Ci, X, Y, Z), w (Co, Ci, 27) -> (N, Co, X', Y', Z')."""
    n, ci, *shape = x.shape
    out_shape = shape if stride == 1 else [coarse(s) for s in shape]
    pad = np.zeros((n, ci) + tuple(s + 3 for s in shape), x.dtype)
    pad[:, :, 1:shape[0] + 1, 1:shape[1] + 1, 1:shape[2] + 1] = x
    y = np.zeros((n, w.shape[0]) + tuple(out_shape))
    for t, (a, bb, c) in enumerate(_taps()):
        sl = pad[:, :, a:a + stride * out_shape[0]:stride, bb:bb + stride * out_shape[1]:stride,
                 c:c + stride * out_shape[2]:stride]
        y += np.einsum("oi,nixyz->noxyz", w[:, :, t], sl)
    return y + b[None, :, None, None, None]


def conv_adjoint(y, w, fine_shape, stride):
    """Adjoint of conv (without bias) with respect to its input: y (N, Co, ...) -> (N, Ci, fine_shape)."""
    n, co, *cs = y.shape
    ci = w.shape[1]
    pad = np.zeros((n, ci) + tuple(s + 3 for s in fine_shape))
    for t, (a, bb, c) in enumerate(_taps()):
        pad[:, :, a:a + stride * cs[0]:stride, bb:bb + stride * cs[1]:stride, c:c + stride * cs[2]:stride] += \
            np.einsum("oi,noxyz->nixyz", w[:, :, t], y)
    return pad[:, :, 1:fine_shape[0] + 1, 1:fine_shape[1] + 1, 1:fine_shape[2] + 1]


def conv_weight_grad(x, dy, stride):
    """dL/dw for y = conv(x, w): (Co, Ci, 27)."""
    n, ci, *shape = x.shape
    out_shape = dy.shape[2:]
    pad = np.zeros((n, ci) + tuple(s + 3 for s in shape))
    pad[:, :, 1:shape[0] + 1, 1:shape[1] + 1, 1:shape[2] + 1] = x
    g = np.zeros((dy.shape[1], ci, 27))
    for t, (a, bb, c) in enumerate(_taps()):
        sl = pad[:, :, a:a + stride * out_shape[0]:stride, bb:bb + stride * out_shape[1]:stride,
                 c:c + stride * out_shape[2]:stride]
        g[:, :, t] = np.einsum("noxyz,nixyz->oi", dy, sl)
    return g


def weights(params, rounded=True):
    """Weights as used by the arithmetic: BF16-rounded (the GPU path) or exact (finite-difference checks)."""
    return {k: (bf16(v["w"]) if rounded else np.asarray(v["w"])).astype(np.float64) for k, v in params.items()}


def forward(params, x, keep=False, rounded=True):
    """x (N, 6, X, Y, Z) -> u (N, 3, X, Y, Z); with keep=True also the activations for backward."""
    w = weights(params, rounded)
    b = {k: np.asarray(v["b"], np.float64) for k, v in params.items()}
    acts = {"x": x}
    relu = lambda v: np.maximum(v, 0.0)
    acts["enc1a"] = relu(conv(x, w["enc1a"], b["enc1a"], 1))
    s1 = acts["s1"] = relu(conv(acts["enc1a"], w["enc1b"], b["enc1b"], 1))
    acts["enc2a"] = relu(conv(s1, w["enc2a"], b["enc2a"], 2))
    acts["s2"] = relu(conv(acts["enc2a"], w["enc2b"], b["enc2b"], 1))
    acts["enc3a"] = relu(conv(acts["s2"], w["enc3a"], b["enc3a"], 2))
    acts["s3"] = relu(conv(acts["enc3a"], w["enc3b"], b["enc3b"], 1))
    acts["enc4a"] = relu(conv(acts["s3"], w["enc4a"], b["enc4a"], 2))
    acts["b4"] = relu(conv(acts["enc4a"], w["enc4b"], b["enc4b"], 1))
    acts["u3"] = relu(conv_adjoint(acts["b4"], w["up3"], acts["s3"].shape[2:], 2)
                      + b["up3"][None, :, None, None, None] + acts["s3"])
    acts["d3"] = relu(conv(acts["u3"], w["dec3"], b["dec3"], 1))
    acts["u2"] = relu(conv_adjoint(acts["d3"], w["up2"], acts["s2"].shape[2:], 2)
                      + b["up2"][None, :, None, None, None] + acts["s2"])
    acts["d2"] = relu(conv(acts["u2"], w["dec2"], b["dec2"], 1))
    acts["u1"] = relu(conv_adjoint(acts["d2"], w["up1"], acts["s1"].shape[2:], 2)
                      + b["up1"][None, :, None, None, None] + s1)
    out = conv(acts["u1"], w["out"], b["out"], 1)
    return (out, acts) if keep else out


def backward(params, acts, dout, rounded=True):
    """Gradients of a scalar loss with respect to every weight and bias, given dL/du (the gradient with
    respect to the rounded weights is applied to the master weights)."""
    w = weights(params, rounded)
    grads = {}
    relu_back = lambda d, a: d * (a > 0)

    def conv_layer(name, x, dy, stride):
        grads[name] = {"w": conv_weight_grad(x, dy, stride), "b": dy.sum(axis=(0, 2, 3, 4))}
        return conv_adjoint(dy, w[name], x.shape[2:], stride)

    def up_layer(name, coarse_act, dy):
        # y = conv_adjoint(c, W) + b: dL/dc = conv(dy, W) (no bias); dL/dW[co][ci][t] = sum c[co] dy[ci @ tap]
        grads[name] = {"w": conv_weight_grad(dy, coarse_act, 2), "b": dy.sum(axis=(0, 2, 3, 4))}
        return conv(dy, w[name], np.zeros(w[name].shape[0]), 2)

    d = conv_layer("out", acts["u1"], dout, 1)
    d = relu_back(d, acts["u1"])
    ds1 = d.copy()
    d = up_layer("up1", acts["d2"], d)
    d = relu_back(d, acts["d2"])
    d = conv_layer("dec2", acts["u2"], d, 1)
    d = relu_back(d, acts["u2"])
    ds2 = d.copy()
    d = up_layer("up2", acts["d3"], d)
    d = relu_back(d, acts["d3"])
    d = conv_layer("dec3", acts["u3"], d, 1)
    d = relu_back(d, acts["u3"])
    ds3 = d.copy()
    d = up_layer("up3", acts["b4"], d)
    d = relu_back(d, acts["b4"])
    d = conv_layer("enc4b", acts["enc4a"], d, 1)
    d = relu_back(d, acts["enc4a"])
    d = conv_layer("enc4a", acts["s3"], d, 2) + ds3
    d = relu_back(d, acts["s3"])
    d = conv_layer("enc3b", acts["enc3a"], d, 1)
    d = relu_back(d, acts["enc3a"])
    d = conv_layer("enc3a", acts["s2"], d, 2) + ds2
    d = relu_back(d, acts["s2"])
    d = conv_layer("enc2b", acts["enc2a"], d, 1)
    d = relu_back(d, acts["enc2a"])
    d = conv_layer("enc2a", acts["s1"], d, 2) + ds1
    d = relu_back(d, acts["s1"])
    d = conv_layer("enc1b", acts["enc1a"], d, 1)
    d = relu_back(d, acts["enc1a"])
    conv_layer("enc1a", acts["x"], d, 1)
    return grads


def features(g, diag, dinv):
    """Network input (N, 6, ...) and per-sample scales from level-1 residuals g (N, 3, ...) and diag(A_1) (3, ...)."""
    h = g * np.sqrt(dinv)[None]
    s = np.max(np.abs(h), axis=(1, 2, 3, 4))
    s = np.where(s > 0, s, 1.0)
    dmax = diag.max()
    f = np.where(diag > 0, np.log10(np.maximum(diag, 1e-30) / dmax) / 9.0 + 1.0, 0.0)
    x = np.concatenate([h / s[:, None, None, None, None], np.broadcast_to(f, g.shape)], axis=1)
    return x, s


def correction(params, g, diag, dinv):
    """delta (N, 3, ...) = Dinv g + Dinv^(1/2) s u."""
    x, s = features(g, diag, dinv)
    u = forward(params, x)
    return dinv[None] * g + np.sqrt(dinv)[None] * s[:, None, None, None, None] * u


def energy_loss(params, g, diag, dinv, apply_a, keep=False, rounded=True):
    """Normalised energy loss and dL/du; apply_a maps (3, X, Y, Z) -> A_1 times it."""
    x, s = features(g, diag, dinv)
    u, acts = forward(params, x, keep=True, rounded=rounded)
    sq = np.sqrt(dinv)
    delta = dinv[None] * g + sq[None] * s[:, None, None, None, None] * u
    loss, du = 0.0, np.zeros_like(u)
    for n in range(g.shape[0]):
        ad = apply_a(delta[n])
        norm = float(np.sum(g[n] * dinv * g[n]))
        norm = norm if norm > 0 else 1.0
        loss += (0.5 * np.sum(delta[n] * ad) - np.sum(delta[n] * g[n])) / norm
        du[n] = (ad - g[n]) * sq * s[n] / norm
    return (loss, du, acts) if keep else loss
