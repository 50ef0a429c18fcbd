"""Floating fixtures on the GPU (deep_frame/topology_multigrid.py:29-38, 151-195, 224-246, 262-283,
340-370): rigid-kernel projection at the fine and coarsest levels, pinned coarsest DOFs, support
reactions and rigid shift, against the oracle (oracle/floating.py, pinned to the clamped solve for a
3-2-1 support and to equilibrium / kernel orthogonality for a line fixture)."""
import numpy as np
import pytest

from deep_frame_amd import problems
from oracle import floating, hex8, system

pytestmark = pytest.mark.gpu


def _case(shape, seed=0):
    h = (1.0, 1.0, 1.0)
    ke, _, _ = hex8.hex8(h, hex8.isotropic_d(0.3))
    pts, conn, dofs = hex8.grid(shape, h)
    density = np.clip(np.random.default_rng(seed).uniform(0.3, 1.2, int(np.prod(shape))), 0, 1)
    mod = 1000.0 * (1e-6 + (1 - 1e-6) * density ** 3)
    return h, ke, pts, conn, dofs, mod, np.asarray(shape) * np.asarray(h) / 2


@pytest.mark.parametrize("projection", [True, False])
@pytest.mark.parametrize("shape", [(8, 3, 5), (16, 8, 12)])
def test_relief_support_matches_clamped_solve(torch_gpu, shape, projection):
    """As tests/test_topology_multigrid.py:77-92 (inertia-relief parts): the floating solve with a 3-2-1
    support equals the direct solve with the support clamped, for unbalanced forces too."""
    from deep_frame_amd.multigrid import GeometricMultigrid
    h, ke, pts, conn, dofs, mod, center = _case(shape)
    support = problems.relief_support(pts, np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    rng = np.random.default_rng(1)
    f = np.zeros((3 * len(pts), 3))
    tip = np.flatnonzero(np.abs(pts[:, 0] - shape[0]) < 1e-9)
    f[problems.clamp_dofs(tip)] = rng.normal(size=(3 * len(tip), 3))
    mg = GeometricMultigrid(shape, h, ke, {"coarsest_dofs": 200, "projection": projection})
    mg.update(mod)
    u, report = mg.solve(b"relief", support, f, support)
    print("floating", shape, projection, report["iterations"], report["kernel"], report["rhs_kernel_component"],
          flush=True)
    expected, _ = system.solve(dofs, conn, ke, mod, support, f, len(pts))
    assert report["converged"] and report["levels"] > 1
    assert report["floating"] == projection and report["kernel"] == (6 if projection else 0)
    if projection:
        assert max(report["rhs_kernel_component"]) < 1e-12
        # hierarchy["kernel_check"] (topology_multigrid.py:325-335): every level's operator annihilates its
        # masked rigid modes up to FP32 rounding
        assert len(report["kernel_check"]) == report["levels"] and max(report["kernel_check"]) < 1e-4
    else:
        assert report["kernel_check"] == []
    assert np.linalg.norm(u - expected) < 1e-7 * np.linalg.norm(expected)
    mg.close()


def test_line_fixture_kernel_projection(torch_gpu):
    """Dirichlet DOFs on one edge line: the rotation about it stays free (one kernel mode, projected
    without a support); forces without a moment about the line give the kernel-orthogonal solution."""
    from deep_frame_amd.multigrid import GeometricMultigrid
    shape = (12, 6, 8)
    h, ke, pts, conn, dofs, mod, center = _case(shape, seed=4)
    line = np.flatnonzero((np.abs(pts[:, 1]) < 1e-9) & (np.abs(pts[:, 2]) < 1e-9))
    fixed = problems.clamp_dofs(line)
    coef = floating.kernel_coefficients(pts, conn, mod, fixed, center)
    mode = (floating.rigid_modes(pts, center) @ coef)[:, 0]
    free = system.free_dofs(conn, mod, fixed, len(pts))
    f = np.random.default_rng(5).normal(size=(3 * len(pts), 2))
    f[np.setdiff1d(np.arange(3 * len(pts)), free)] = 0
    f -= np.outer(mode, mode @ f) / (mode @ mode)
    mg = GeometricMultigrid(shape, h, ke, {"coarsest_dofs": 200})
    mg.update(mod)
    u, report = mg.solve(b"line", fixed, f)
    print("line", report["iterations"], report["rhs_kernel_component"], flush=True)
    expected, nk, _ = floating.solve(dofs, conn, pts, ke, mod, fixed, np.zeros(0, dtype=np.int64), f, center)
    assert nk == 1 and mg.hierarchies[(b"line", False)]["kernel"] == 1
    assert report["converged"] and not report["floating"]
    assert max(report["rhs_kernel_component"]) < 1e-12
    assert np.linalg.norm(u - expected) < 1e-7 * np.linalg.norm(expected)
    mg.close()
