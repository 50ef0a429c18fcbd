"""SIMP step kernels against the oracle: cone filter and pullback (reference
tests/test_topology_multigrid.py:140-156 pattern), SIMP interpolation, OC update,
MMA step, and one full design iteration of the 48 x 24 x 24 cantilever."""
import numpy as np
import pytest

from deep_frame_amd import element, problems
from oracle import filters, hex8, mma, oc, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,spacing,radius", [((10, 8, 6), (1.0, 1.0, 1.0), 1.5),
                                                  ((14, 11, 9), (0.747, 0.753, 0.667), 2.84)])
def test_filter_and_pullback_match_oracle(shape, spacing, radius, torch_gpu):
    from deep_frame_amd.simp import DensityFilter
    print(f"filter shape={shape}", flush=True)
    torch = torch_gpu
    rng = np.random.default_rng(3)
    allowed = rng.uniform(size=shape) > 0.2
    allowed[0] = allowed[-1] = True
    preserve = np.zeros(shape, bool)
    preserve[0, :2, :2] = True
    ref = filters.Cone(allowed, spacing, radius, preserve)
    f = DensityFilter(shape, spacing, radius, allowed, preserve)
    x = rng.uniform(size=allowed.size)
    s = rng.normal(size=allowed.size)
    d = rng.uniform(0.5, 1.5, size=allowed.size)
    out = f.filtered(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert np.max(np.abs(out - ref.filtered(x))) < 1e-12
    back = f.pullback(torch.as_tensor(s, device="cuda"), torch.as_tensor(d, device="cuda")).cpu().numpy()
    exp = ref.pullback(s, d)
    assert np.max(np.abs(back - exp)) < 1e-12 * np.max(np.abs(exp))
    assert f.volume(torch.as_tensor(x, device="cuda")) == pytest.approx(ref.filtered(x).sum(), rel=1e-12)
    f.close()


def test_simp_interpolation(torch_gpu):
    from deep_frame_amd.simp import simp
    rho = np.random.default_rng(0).uniform(1e-3, 1, 1000)
    e, de = simp(torch_gpu.as_tensor(rho, device="cuda"), 2.0, 3.0, 1e-9, active=rho > 0.01)
    ref_e, ref_de = hex8.simp(rho, 2.0, 3.0, 1e-9)
    ref_e[rho <= 0.01] = 0
    ref_de[rho <= 0.01] = 0
    assert np.allclose(e.cpu().numpy(), ref_e, rtol=1e-14, atol=0)
    assert np.allclose(de.cpu().numpy(), ref_de, rtol=1e-14, atol=0)


def test_oc_update_matches_oracle(torch_gpu):
    from deep_frame_amd.simp import DensityFilter
    shape = (12, 8, 6)
    rng = np.random.default_rng(5)
    allowed = np.ones(shape, bool)
    ref = filters.Cone(allowed, (1, 1, 1), 1.5)
    f = DensityFilter(shape, (1, 1, 1), 1.5)
    x = rng.uniform(0.2, 0.7, allowed.size)
    dc = -rng.uniform(0.1, 2.0, allowed.size)
    dv = ref.pullback(np.ones(allowed.size))
    target = 0.4 * allowed.size
    exp = oc.oc_update(x, dc, dv, ref.free.ravel(), lambda z: ref.filtered(z).sum(), target, 1e-3, 0.2)
    t = torch_gpu
    out = f.oc_update(t.as_tensor(x, device="cuda"), t.as_tensor(dc, device="cuda"), t.as_tensor(dv, device="cuda"), target, 1e-3, 0.2).cpu().numpy()
    assert np.max(np.abs(out - exp)) < 1e-6
    assert f.volume(t.as_tensor(out, device="cuda")) <= target + 1e-9
    f.close()


@pytest.mark.parametrize("iteration", [1, 4])
@pytest.mark.parametrize("m", [1, 3])
def test_mma_step_matches_oracle(iteration, m, torch_gpu):
    """One device MMA step with m constraints (MMAOptimizer.step passes m = len(g) rows,
    topology_optimizers.py:55-67) against the oracle, whose steps are pinned by hand-derived cases."""
    from deep_frame_amd.simp import MMAState
    print(f"mma iteration={iteration} m={m}", flush=True)
    rng = np.random.default_rng(6 + iteration + 10 * m)
    n = 500
    x = rng.uniform(0.2, 0.8, n)
    old1 = np.clip(x + rng.normal(0, 0.05, n), 0, 1)
    old2 = np.clip(old1 + rng.normal(0, 0.05, n), 0, 1)
    low, upp = x - 0.3, x + 0.3
    df0 = rng.uniform(-1.0, 0.2, n)
    dg = rng.uniform(0.5, 1.5, (m, n)) / n
    dg[1:] *= rng.choice([-1.0, 1.0], size=(m - 1, 1))
    g = dg @ x - rng.uniform(0.3, 0.5, m)
    exp, elow, eupp, _, _ = mma.mma_step(iteration, x, old1, old2, low, upp, 0.0, 1.0, 1.0, df0, g, dg,
                                         move=0.1, scale=100.0)
    t = torch_gpu
    dev = lambda a: t.as_tensor(np.ascontiguousarray(a), device="cuda")
    st = MMAState(dev(x))
    st.iteration = iteration
    st.old1, st.old2 = dev(old1), dev(old2)
    st.low, st.upp = dev(low), dev(upp)
    out = st.step(dev(x), 1.0, dev(df0), g if m > 1 else float(g[0]), dev(dg if m > 1 else dg[0]), move=0.1,
                  scale=100.0).cpu().numpy()
    assert np.max(np.abs(out - np.clip(exp, 0, 1))) < 1e-6
    assert np.allclose(st.low.cpu().numpy(), elow, atol=1e-14) and np.allclose(st.upp.cpu().numpy(), eupp, atol=1e-14)
    assert np.allclose(st.old1.cpu().numpy(), x) and np.allclose(st.old2.cpu().numpy(), old1)


def test_cantilever_design_iteration_matches_oracle(torch_gpu):
    """Config-1 step: filter r = 1.5, SIMP p = 3, FP64-accurate solve, sensitivity pullback, MMA update."""
    from deep_frame_amd.simp import MMAState, SimpProblem
    prob = problems.cantilever(seed=0)
    shape = prob["shape"]
    sp = SimpProblem(shape, prob["fixed"], prob["forces"], volume_fraction=0.3, settings={"coarsest_dofs": 5000})
    x = torch_gpu.as_tensor(prob["density"], device="cuda")
    res = sp.evaluate(x)
    # oracle chain
    cone = filters.Cone(np.ones(shape, bool), (1, 1, 1), 1.5)
    rho = cone.filtered(prob["density"])
    mod, dmod = hex8.simp(rho, 1.0, 3.0, 1e-9)
    pts, conn, dofs = hex8.grid(shape)
    ke, _, _ = hex8.hex8((1, 1, 1), hex8.isotropic_d(0.3))
    u, _ = system.solve(dofs, conn, ke, mod, prob["fixed"], prob["forces"], len(pts))
    c, s = system.compliance_sensitivity(dofs, ke, u, prob["forces"], dmod)
    dc = cone.pullback(s[:, 0])
    assert res["compliance"] == pytest.approx(c[0], rel=1e-6)
    got = res["dc"].cpu().numpy()
    assert np.linalg.norm(got - dc) <= 1e-5 * np.linalg.norm(dc)
    assert np.max(np.abs(res["physical"].cpu().numpy() - rho)) < 1e-13
    # one MMA update of the volume-constrained problem, against the oracle MMA on the same sensitivities
    st = MMAState(x)
    new = st.step(x, res["compliance"], res["dc"], res["volume"] / sp.target - 1, sp.dv / sp.target).cpu().numpy()
    dv = cone.pullback(np.ones(rho.size))
    exp, _, _, _, _ = mma.mma_step(1, prob["density"], prob["density"], prob["density"], np.zeros(rho.size),
                                   np.ones(rho.size), 0.0, 1.0, res["compliance"], dc, rho.sum() / sp.target - 1,
                                   (dv / sp.target)[None, :], move=0.1)
    assert np.max(np.abs(new - np.clip(exp, 0, 1))) < 1e-5
    sp.close()
