"""Learned coarse correction on the GPU against the CPU restatement (oracle/unet.py): forward pass,
backward pass (every layer's weight and bias gradients), the energy-loss training step with the
level-1 Galerkin operator, and what the trained network gains inside PCG."""
import numpy as np
import pytest

from deep_frame_amd import element, problems
from oracle import gmg, unet

pytestmark = pytest.mark.gpu


def random_params(seed):
    p = unet.init(seed)
    rng = np.random.default_rng(seed + 1)
    p["out"]["w"] = rng.normal(0, 0.1, p["out"]["w"].shape).astype(np.float32)
    for k in p:
        p[k]["b"] = rng.normal(0, 0.05, p[k]["b"].shape).astype(np.float32)
    return p


def flatten(p):
    from deep_frame_amd.learned import LAYERS
    return np.concatenate([np.concatenate([p[n]["w"].ravel(), p[n]["b"].ravel()]) for n, *_ in LAYERS]).astype(np.float32)


def rel(a, b):
    return np.linalg.norm(np.asarray(a, float) - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_matches_oracle(torch_gpu):
    from deep_frame_amd.learned import CoarseCorrection
    print("net forward (2, 6, 9, 7, 5)", flush=True)
    p = random_params(0)
    net = CoarseCorrection(flatten(p))
    x = np.random.default_rng(3).normal(size=(2, 6, 9, 7, 5)).astype(np.float32)
    u = net.forward(torch_gpu.as_tensor(x, device="cuda")).cpu().numpy()
    ref = unet.forward(p, x.astype(np.float64))
    assert rel(u, ref) < 1e-5
    net.close()


def test_backward_matches_oracle(torch_gpu):
    from deep_frame_amd.learned import CoarseCorrection, unflatten
    print("net backward (3, 6, 11, 9, 6)", flush=True)
    p = random_params(1)
    net = CoarseCorrection(flatten(p))
    rng = np.random.default_rng(4)
    x = rng.normal(size=(3, 6, 11, 9, 6)).astype(np.float32)
    du = rng.normal(size=(3, 3, 11, 9, 6)).astype(np.float32)
    net.forward(torch_gpu.as_tensor(x, device="cuda"))
    net.backward(torch_gpu.as_tensor(du, device="cuda"))
    got = unflatten(net.grads().cpu().numpy())
    _, acts = unet.forward(p, x.astype(np.float64), keep=True)
    ref = unet.backward(p, acts, du.astype(np.float64))
    for name in ref:
        for which in ("w", "b"):
            assert rel(got[name][which], ref[name][which]) < 1e-4, (name, which)
    net.close()


def small_frame(seed=0, shape=(16, 16, 8), levels=3, rhs=12):
    from deep_frame_amd.multigrid import GeometricMultigrid
    prob = problems.frame(shape, seed=seed, rhs=rhs)
    ke = element.stiffness((1, 1, 1), element.isotropic(0.3))
    mod, _ = problems.moduli(prob["density"])
    mg = GeometricMultigrid(shape, (1, 1, 1), ke, {"levels": levels, "precision": "float32"})
    mg.update(mod)
    return prob, ke, mod, mg


def test_energy_train_step_matches_oracle(torch_gpu):
    from deep_frame_amd.learned import CoarseCorrection, harvest, unflatten
    print("net train step on frame 16x16x8", flush=True)
    torch = torch_gpu
    prob, ke, mod, mg = small_frame()
    rhs = torch.as_tensor(np.ascontiguousarray(prob["forces"].T), device="cuda")
    samples, _ = harvest(mg, b"t", prob["fixed"], rhs, 64)
    samples = samples[:5]
    p = random_params(2)
    net = CoarseCorrection(flatten(p))
    loss = net.train_step(mg, b"t", samples, lr=0.0)
    got = unflatten(net.grads().cpu().numpy())
    # oracle level-1 Galerkin operator on the padded grid
    padded = mg.padded
    pts = element.node_points(prob["shape"])
    i, j, k = np.unravel_index(np.arange(len(pts)), tuple(n + 1 for n in prob["shape"]))
    pidx = np.ravel_multi_index((i, j, k), tuple(n + 1 for n in padded))
    fixed = 3 * pidx[prob["fixed"] // 3] + prob["fixed"] % 3
    levels = gmg.hierarchy(prob["shape"], ke, mod, fixed, (3, padded))
    a1 = levels[1]["A"]
    dims = tuple(n // 2 + 1 for n in padded)
    diag = a1.diagonal().reshape(dims + (3,)).transpose(3, 0, 1, 2)
    dinv = np.where(diag > 0, 1.0 / np.where(diag > 0, diag, 1.0), 0.0)

    def apply_a(v):
        return (a1 @ v.transpose(1, 2, 3, 0).ravel()).reshape(dims + (3,)).transpose(3, 0, 1, 2)

    g = samples.cpu().numpy().astype(np.float64).reshape((-1,) + dims + (3,)).transpose(0, 4, 1, 2, 3)
    ref_loss, du, acts = unet.energy_loss(p, g, diag, dinv, apply_a, keep=True)
    ref = unet.backward(p, acts, du)
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    for name in ref:
        assert rel(got[name]["w"], ref[name]["w"]) < 2e-3, name
    net.close()
    mg.close()


def test_trained_network_reduces_pcg_iterations(torch_gpu):
    """Train on residuals harvested from three densities, then solve a held-out density with and
    without the network; the network must not cost iterations and its energy loss must fall."""
    from deep_frame_amd.learned import CoarseCorrection, harvest
    torch = torch_gpu
    net = CoarseCorrection(seed=0)
    data = []
    for seed in (1, 2, 3):
        prob, ke, mod, mg = small_frame(seed, (32, 32, 8), 3, 42)
        rhs = torch.as_tensor(np.ascontiguousarray(prob["forces"].T), device="cuda")
        samples, _ = harvest(mg, b"h", prob["fixed"], rhs, 42 * 40)
        data.append((mg, samples))
    first = last = None
    for epoch in range(6):
        for mg, samples in data:
            for start in range(0, samples.shape[0], 42):
                loss = net.train_step(mg, b"h", samples[start:start + 42], lr=1e-3) / min(42, samples.shape[0] - start)
                first = loss if first is None else first
                last = loss
    print("energy loss per sample: first", first, "last", last, flush=True)
    assert last < first
    prob, ke, mod, mg = small_frame(0, (32, 32, 8), 3, 42)
    rhs = torch.as_tensor(np.ascontiguousarray(prob["forces"].T), device="cuda")
    _, plain = mg.solve_device(b"p", prob["fixed"], rhs)
    net.attach(mg)
    x, learned = mg.solve_device(b"p", prob["fixed"], rhs)
    it_plain, it_learned = max(r[0] for r in plain), max(r[0] for r in learned)
    print("iterations plain", it_plain, "learned", it_learned, flush=True)
    assert all(r[2] and r[1] < 1e-8 for r in learned)
    assert it_learned <= it_plain
    for mg_, _ in data:
        mg_.close()
    mg.close()
    net.close()
