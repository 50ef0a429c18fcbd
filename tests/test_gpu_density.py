"""Density map on the GPU (deep_frame/topology_optimization.py:1110-1195; TopologyProblem.evaluate
pullbacks, deep_frame/topology_problem.py:366-379) against the oracle (oracle/filters.py: brute-force cone
sums and the projection, pinned by closed forms and central differences), and a solve on the optimised
half-frame density the reference exports."""
from pathlib import Path

import numpy as np
import pytest

from deep_frame_amd import formats, problems
from oracle import filters, hex8

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("robust", [False, True])
@pytest.mark.parametrize("beta", [0.0, 1.0, 8.0, 64.0])
def test_projection_matches_oracle(torch_gpu, robust, beta):
    from deep_frame_amd import _native
    torch = torch_gpu
    rng = np.random.default_rng(int(beta) + 7 * robust)
    n = 4099
    filtered = rng.uniform(0, 1, n)
    filtered[:3] = [0.0, 1.0, 0.5]
    preserve, forbidden = rng.random(n) < 0.05, rng.random(n) < 0.05
    forbidden &= ~preserve
    free = ~preserve & ~forbidden
    flags = preserve.astype(np.uint8) | (forbidden.astype(np.uint8) << 1) | (free.astype(np.uint8) << 2)
    etas = [0.75, 0.5, 0.25]
    fd, fl = torch.as_tensor(filtered, device="cuda"), torch.as_tensor(flags, device="cuda")
    phys, der = torch.empty((3, n), dtype=torch.float64, device="cuda"), torch.empty((3, n), dtype=torch.float64, device="cuda")
    from ctypes import c_double
    _native.call("dfmi_project", _native.ptr(fd), _native.ptr(fl), n, beta, (c_double * 3)(*etas), 3, int(robust),
                 _native.ptr(phys), _native.ptr(der), _native.stream_handle(torch))
    torch.cuda.synchronize()
    for q, eta in enumerate(etas):
        p, d = filters.project_masked(filtered, beta, eta, robust, preserve, forbidden, free)
        assert np.max(np.abs(phys[q].cpu().numpy() - p)) < 1e-14
        assert np.max(np.abs(der[q].cpu().numpy() - d)) < 1e-13 * max(1.0, beta)


def _domain(shape=(10, 8, 6), seed=0):
    rng = np.random.default_rng(seed)
    allowed = rng.random(shape) < 0.85
    allowed[:, :, 0] = True
    preserve = allowed & (rng.random(shape) < 0.08)
    return allowed, preserve, ~allowed


def test_density_map_chain_matches_oracle(torch_gpu):
    """Filter, robust fields, dilated volume, pullbacks and the evaluate gradient assembly vs the oracle,
    and the volume gradient against central differences of the volume."""
    from deep_frame_amd.density import DensityMap, evaluate_pullbacks
    torch = torch_gpu
    shape, spacing, radius = (10, 8, 6), (1.0, 0.8, 1.2), 2.5
    allowed, preserve, forbidden = _domain(shape)
    settings = {"filter_radius_mm": radius, "projection": "robust", "projection_beta": 4.0}
    dm = DensityMap(shape, spacing, allowed, preserve, forbidden, settings)
    rng = np.random.default_rng(1)
    design = rng.uniform(0, 1, dm.n)
    cone = filters.Cone(allowed, spacing, radius, preserve)
    filtered = cone.filtered(design)
    fields, f_dev = dm.fields(design)
    assert np.max(np.abs(f_dev.cpu().numpy() - filtered)) < 1e-14
    for name, eta in dm.thresholds.items():
        p, d = filters.project_masked(filtered, 4.0, eta, True, preserve.ravel(), forbidden.ravel(), dm.free)
        assert np.max(np.abs(fields[name][0].cpu().numpy() - p)) < 1e-13
        assert np.max(np.abs(fields[name][1].cpu().numpy() - d)) < 1e-12
    dil = filters.project_masked(filtered, 4.0, 0.25, True, preserve.ravel(), forbidden.ravel(), dm.free)
    assert dm.volume(design) == pytest.approx(dil[0].sum(), rel=1e-13)
    grad = dm.pullback(dm.weights, fields["dilated"][1]).cpu().numpy()
    assert np.allclose(grad, cone.pullback(np.ones(dm.n), dil[1]), rtol=1e-12, atol=1e-15)
    for e in np.flatnonzero(dm.free)[::97][:5]:
        step = np.zeros(dm.n)
        step[e] = 1e-6
        fd = (dm.volume(np.clip(design + step, 0, 1)) - dm.volume(np.clip(design - step, 0, 1))) / 2e-6
        assert fd == pytest.approx(grad[e], rel=1e-6, abs=1e-10)
    rows = [torch.as_tensor(rng.normal(size=dm.n), device="cuda") for _ in range(3)]
    geo = [torch.as_tensor(rng.normal(size=dm.n), device="cuda")]
    objective, og, grads = evaluate_pullbacks(dm, fields, rows, geo, 2.0, 50.0)
    inter = filters.project_masked(filtered, 4.0, 0.5, True, preserve.ravel(), forbidden.ravel(), dm.free)
    ero = filters.project_masked(filtered, 4.0, 0.75, True, preserve.ravel(), forbidden.ravel(), dm.free)
    assert objective == pytest.approx(2.0 * inter[0].sum() / 50.0, rel=1e-13)
    assert np.allclose(og.cpu().numpy(), cone.pullback(np.full(dm.n, 2.0 / 50.0), inter[1]), rtol=1e-12, atol=1e-16)
    for i, r in enumerate(rows):
        assert np.allclose(grads[i].cpu().numpy(), cone.pullback(r.cpu().numpy(), ero[1]), rtol=1e-12, atol=1e-14)
    assert np.allclose(grads[3].cpu().numpy(), cone.pullback(geo[0].cpu().numpy(), inter[1]), rtol=1e-12, atol=1e-14)
    target = 0.3 * dm.n
    initial = dm.initial(target)
    assert dm.volume(initial) <= target and dm.volume(initial) == pytest.approx(target, rel=1e-9)
    dm.close()


def test_solve_on_exported_half_frame_density(torch_gpu):
    """The optimised half-frame density the reference exports (51 x 96 x 24, 3 % material, min stiffness
    ratio 1e-6): six unit-wrench solves under clamped pads converge and the residual of the assembled FP64
    operator meets the tolerance."""
    from deep_frame_amd.multigrid import GeometricMultigrid
    torch = torch_gpu
    density = formats.load_density_half(GOLDEN / "density_half_neural_r3_v05.npz")
    shape = density.shape
    prob = problems.frame(shape, rhs=6)
    ke, _, _ = hex8.hex8((1.0, 1.0, 1.0), hex8.isotropic_d(0.3))
    mod = 1000.0 * (1e-6 + (1 - 1e-6) * density.ravel() ** 3)
    mg = GeometricMultigrid(shape, (1.0, 1.0, 1.0), ke, {"precision": "float32", "refine": True})
    mg.update(mod)
    f = torch.as_tensor(np.ascontiguousarray(prob["forces"].T), device="cuda")
    u, reps = mg.solve_device(b"real", prob["fixed"], f)
    print("real density", [r[:2] for r in reps], flush=True)
    assert all(r[2] for r in reps)
    pts, conn, dofs = hex8.grid(shape, (1.0, 1.0, 1.0))
    k = hex8.assemble(dofs, ke, mod, 3 * len(pts))
    free = np.ones(3 * len(pts), bool)
    free[prob["fixed"]] = False
    uh, fh = u.cpu().numpy().T, prob["forces"]
    res = np.linalg.norm((k @ uh - fh)[free], axis=0) / np.linalg.norm(fh[free], axis=0)
    assert np.all(res < 2e-8)
    mg.close()


@pytest.mark.parametrize("support", [False, True])
def test_solve_on_exported_075mm_half_frame_density(torch_gpu, support):
    """The optimised 0.75 mm half-frame density the reference exports (91 x 170 x 48, 2,312,604 DOF; the
    grid on which the reference's multigrid probe stalled at rel-res 4.2e-8): six unit-wrench solves under
    clamped pads, and under a floating 3-2-1 support, converge to 1e-8; the FP64 operator confirms the
    residual."""
    from deep_frame_amd import element
    from deep_frame_amd.multigrid import GeometricMultigrid
    torch = torch_gpu
    density = formats.load_density_half(GOLDEN / "density_half_075mm_run3.npz")
    shape = density.shape
    prob = problems.frame(shape, rhs=6)
    ke = element.stiffness((1.0, 1.0, 1.0), element.isotropic(0.3))
    mod = 1000.0 * (1e-6 + (1 - 1e-6) * density.ravel() ** 3)
    mg = GeometricMultigrid(shape, (1.0, 1.0, 1.0), ke, {"precision": "float32", "refine": True})
    mg.update(mod)
    if support:
        points = element.node_points(shape)
        pins = problems.relief_support(points, np.flatnonzero(points[:, 2] == 0))
        u, report = mg.solve(b"real75", pins, prob["forces"], pins)
        print("0.75 mm floating", max(report["iterations"]), max(report["relative_residual"]), flush=True)
        assert report["converged"] and report["floating"] and report["kernel"] == 6
        assert np.abs(u[pins]).max() < 1e-9 * np.abs(u).max()
        return
    f = torch.as_tensor(np.ascontiguousarray(prob["forces"].T), device="cuda")
    u, reps = mg.solve_device(b"real75", prob["fixed"], f)
    print("0.75 mm clamped", [r[:2] for r in reps], flush=True)
    assert all(r[2] for r in reps)
    res = f - mg.operator_device(u, b"real75")
    rel = (torch.linalg.vector_norm(res, dim=1) / torch.linalg.vector_norm(f, dim=1)).cpu().numpy()
    assert np.all(rel < 1.5e-8)
    mg.close()


@pytest.mark.parametrize("radii", [(1.5, 3.0), (2.5, 2.0)])
def test_corridor_blend_matches_oracle(torch_gpu, radii):
    """DensityMap's corridor (topology_optimization.py:1125-1129, :1142, :1176): a wider (or narrower) cone
    blended in with per-cell weights; forward, pullback and the OC volume go through the blended filter."""
    from deep_frame_amd.density import DensityMap
    torch = torch_gpu
    shape, spacing = (10, 8, 6), (1.0, 0.8, 1.2)
    allowed, preserve, forbidden = _domain(shape, seed=3)
    rng = np.random.default_rng(5)
    weights = np.where(rng.random(shape) < 0.4, rng.uniform(-0.2, 1.3, shape), 0.0)
    corridor = {"radius_mm": radii[1], "weights": weights.ravel()}
    dm = DensityMap(shape, spacing, allowed, preserve, forbidden,
                    {"filter_radius_mm": radii[0], "projection_beta": 0.0, "corridor": corridor})
    ref = filters.Cone(allowed, spacing, radii[0], preserve, corridor=(radii[1], weights))
    x = rng.uniform(0, 1, allowed.size)
    got = dm.filtered(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert np.max(np.abs(got - ref.filtered(x))) < 1e-13
    s, d = rng.normal(size=allowed.size), rng.uniform(0.5, 1.5, allowed.size)
    back = dm.pullback(torch.as_tensor(s, device="cuda"), torch.as_tensor(d, device="cuda")).cpu().numpy()
    assert np.max(np.abs(back - ref.pullback(s, d))) < 1e-12 * np.abs(ref.pullback(s, d)).max()
    assert dm.filter.volume(torch.as_tensor(x, device="cuda")) == pytest.approx(ref.filtered(x).sum(), rel=1e-13)
    dm.close()
