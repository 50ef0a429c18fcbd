"""V-cycle pieces against the oracle: transfer operators (P, P^T) and one full
V-cycle (deep_frame/topology_multigrid.py:284-324) with the GPU's lambda_max."""
import numpy as np
import pytest

from deep_frame_amd import element, problems
from oracle import gmg

pytestmark = pytest.mark.gpu


def case(shape, spacing, seed, precision, torch, fine_smoother="vector", coarse_krylov=0):
    from deep_frame_amd.multigrid import GeometricMultigrid
    ke = element.stiffness(spacing, element.isotropic(0.3))
    rho = np.clip(np.random.default_rng(seed).uniform(0.3, 1.2, int(np.prod(shape))), 0, 1)
    mod = 1e-6 + (1 - 1e-6) * rho ** 3
    pts = element.node_points(shape, spacing)
    fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    mg = GeometricMultigrid(shape, spacing, ke, {"coarsest_dofs": 200, "precision": precision,
                                                 "fine_smoother": fine_smoother, "coarse_krylov": coarse_krylov})
    mg.update(mod)
    entry = mg.hierarchy("v", fixed)
    levels = gmg.hierarchy(shape, ke, mod, fixed, gmg.level_count(shape, 200))
    return mg, entry, levels, fixed


@pytest.mark.parametrize("shape", [(24, 4, 4), (16, 8, 8)])
def test_transfer_operators_match_prolongation(shape, torch_gpu):
    print(f"transfer shape={shape}", flush=True)
    mg, entry, levels, _ = case(shape, (1.0, 1.0, 1.0), 0, "float64", torch_gpu)
    for index in range(len(levels) - 1):
        fine, coarse = levels[index], levels[index + 1]
        p = coarse["P"]
        vf = np.random.default_rng(index).normal(size=(p.shape[0], 2)) * fine["mask"][:, None]
        vc = np.random.default_rng(10 + index).normal(size=(p.shape[1], 2)) * coarse["mask"][:, None]
        down = mg.transfer("v", index, 0, torch_gpu.as_tensor(vf.T.copy(), device="cuda"), p.shape[1]).cpu().numpy().T
        up = mg.transfer("v", index, 1, torch_gpu.as_tensor(vc.T.copy(), device="cuda"), p.shape[0]).cpu().numpy().T
        assert np.allclose(down, (p.T @ vf) * coarse["mask"][:, None], rtol=0, atol=1e-12 * np.abs(vf).max() * 8)
        assert np.allclose(up, (p @ vc) * fine["mask"][:, None], rtol=0, atol=1e-12 * np.abs(vc).max())
    mg.close()


@pytest.mark.parametrize("shape,spacing", [((24, 4, 4), (2.0, 1.5, 1.5)), ((16, 8, 8), (1.0, 1.0, 1.0)),
                                           ((16, 8, 40), (1.0, 1.0, 0.5))])
@pytest.mark.parametrize("precision", ["float64", "float32", "float32-matrix"])
def test_vcycle_matches_oracle(shape, spacing, precision, torch_gpu):
    """FP64 V-cycle to 1e-9; FP32 V-cycle to 1e-4, with the fine level on the FP32 vector units (the
    default) or on the matrix cores (split bf16)."""
    print(f"vcycle shape={shape} precision={precision}", flush=True)
    smoother = "matrix" if precision.endswith("matrix") else "vector"
    mg, entry, levels, fixed = case(shape, spacing, 1, precision.split("-")[0], torch_gpu, smoother)
    # 2 columns: one column per stencil thread; 7 columns: column pairs per stencil thread and an odd last one;
    # 20 columns: the LDS-tiled stencil form in two column chunks
    for cols in (2, 7, 20):
        r = np.random.default_rng(2).normal(size=(levels[0]["A"].shape[0], cols)) * levels[0]["mask"][:, None]
        ref = np.column_stack([gmg.vcycle(levels, entry["lambda_max"], r[:, c]) for c in range(cols)])
        z = mg.precondition_device("v", torch_gpu.as_tensor(r.T.copy(), device="cuda")).cpu().numpy().T
        tol = 1e-9 if precision == "float64" else 1e-4
        for c in range(cols):
            assert np.linalg.norm(z[:, c] - ref[:, c]) < tol * np.linalg.norm(ref[:, c]), (cols, c)
    mg.close()


@pytest.mark.parametrize("shape,spacing", [((16, 8, 8), (1.0, 1.0, 1.0)), ((16, 8, 40), (1.0, 1.0, 0.5))])
@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("steps", [1, 2])
def test_kcycle_matches_oracle(shape, spacing, precision, steps, torch_gpu):
    """Fine V-cycle whose level-1 correction is solved by one or two flexible-CG steps (coarse_krylov)
    against oracle.gmg.kcycle: FP64 to 1e-9, FP32 to 1e-4; 1, 2, 7 and 20 columns (single-column layout,
    column pairs, an odd last column, the LDS-tiled stencil form)."""
    print(f"kcycle shape={shape} precision={precision} steps={steps}", flush=True)
    mg, entry, levels, fixed = case(shape, spacing, 1, precision, torch_gpu, coarse_krylov=steps)
    assert len(levels) >= 3
    for cols in (1, 2, 7, 20):
        r = np.random.default_rng(3).normal(size=(levels[0]["A"].shape[0], cols)) * levels[0]["mask"][:, None]
        ref = np.column_stack([gmg.kcycle(levels, entry["lambda_max"], r[:, c], steps) for c in range(cols)])
        z = mg.precondition_device("v", torch_gpu.as_tensor(r.T.copy(), device="cuda")).cpu().numpy().T
        tol = 1e-9 if precision == "float64" else 1e-4
        for c in range(cols):
            assert np.linalg.norm(z[:, c] - ref[:, c]) < tol * np.linalg.norm(ref[:, c]), (cols, c)
    mg.close()


@pytest.mark.parametrize("settings", [{"smoother": "jacobi"}, {"smoother": "jacobi", "sweeps": 3, "damping": 0.8},
                                      {"cycle": "W"}, {"smoother": "jacobi", "cycle": "W"},
                                      {"coarse": "rediscretize"}, {"coarse": "rediscretize", "cycle": "W"},
                                      {"chebyshev_kind": 4}, {"chebyshev_kind": 4, "cycle": "W"}])
@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_reference_cycle_settings_match_oracle(settings, precision, torch_gpu):
    """The reference's other MULTIGRID choices (topology_multigrid.py:51, :160-166, :284-290, :312): damped Jacobi
    smoothing, the W-cycle and rediscretised coarse levels, and the fourth-kind Chebyshev smoother, against the
    oracle's cycle on the same hierarchy (FP64 to 1e-9, FP32 to 1e-4) for 1, 2 and 7 columns."""
    from deep_frame_amd.multigrid import GeometricMultigrid
    print(f"cycle settings={settings} precision={precision}", flush=True)
    shape, spacing = (16, 8, 8), (1.0, 1.0, 1.0)
    ke = element.stiffness(spacing, element.isotropic(0.3))
    rho = np.clip(np.random.default_rng(1).uniform(0.3, 1.2, int(np.prod(shape))), 0, 1)
    mod = 1e-6 + (1 - 1e-6) * rho ** 3
    pts = element.node_points(shape, spacing)
    fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    mg = GeometricMultigrid(shape, spacing, ke, {"coarsest_dofs": 200, "precision": precision, **settings})
    mg.update(mod)
    entry = mg.hierarchy("v", fixed)
    levels = gmg.hierarchy(shape, ke, mod, fixed, gmg.level_count(shape, 200), coarse=settings.get("coarse", "galerkin"))
    assert len(levels) == mg.level_count >= 3
    for cols in (1, 2, 7):
        r = np.random.default_rng(2).normal(size=(levels[0]["A"].shape[0], cols)) * levels[0]["mask"][:, None]
        ref = np.column_stack([gmg.vcycle(levels, entry["lambda_max"], r[:, c], smoother=settings.get("smoother", "chebyshev"),
                                          sweeps=settings.get("sweeps", 2), damping=settings.get("damping", 1.0),
                                          cycle=settings.get("cycle", "V"), kind=settings.get("chebyshev_kind", 1))
                               for c in range(cols)])
        z = mg.precondition_device("v", torch_gpu.as_tensor(r.T.copy(), device="cuda")).cpu().numpy().T
        tol = 1e-9 if precision == "float64" else 1e-4
        for c in range(cols):
            assert np.linalg.norm(z[:, c] - ref[:, c]) < tol * np.linalg.norm(ref[:, c]), (cols, c)
    mg.close()


@pytest.mark.parametrize("settings", [{"smoother": "jacobi"}, {"cycle": "W"}, {"coarse": "rediscretize"}])
def test_reference_cycle_settings_solve_with_history(settings, torch_gpu):
    """A solve with each setting meets 1e-8 against the assembled FP64 operator, and the report carries the
    reference's per-batch ``history`` (one entry per PCG iteration, ending below the tolerance) and an empty
    ``kernel_check`` for a clamped fixture."""
    from deep_frame_amd.multigrid import GeometricMultigrid
    from oracle import hex8
    print(f"solve settings={settings}", flush=True)
    prob = problems.cantilever(seed=0, shape=(16, 8, 8))
    ke = element.stiffness((1, 1, 1), element.isotropic(0.3))
    mod, _ = problems.moduli(prob["density"])
    mg = GeometricMultigrid(prob["shape"], (1, 1, 1), ke, {"coarsest_dofs": 500, **settings})
    mg.update(mod)
    x, report = mg.solve(b"s", prob["fixed"], prob["forces"])
    pts, conn, dofs = hex8.grid(prob["shape"])
    k = hex8.assemble(dofs, ke, mod, 3 * len(pts))
    free = np.setdiff1d(np.arange(k.shape[0]), prob["fixed"])
    res = prob["forces"][free] - (k @ x)[free]
    assert np.linalg.norm(res) < 1e-8 * np.linalg.norm(prob["forces"][free])
    assert report["converged"] and len(report["history"]) == 1
    assert len(report["history"][0]) == max(report["iterations"])
    assert report["history"][0][-1] < 1e-8 and report["kernel_check"] == []
    mg.close()


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("krylov_levels,kind", [(2, 1), (3, 1), (2, 4)])
def test_recursive_kcycle_matches_oracle(precision, krylov_levels, kind, torch_gpu):
    """The Krylov coarse solve applied recursively (krylov_levels): the V-cycle inside each flexible-CG step
    of level l solves its own coarse correction by flexible CG on level l + 1, on a 5-level hierarchy, against
    oracle.gmg.kcycle (FP64 to 1e-9, FP32 to 1e-4) for 1, 2 and 7 columns; first- and fourth-kind Chebyshev
    smoothing."""
    from deep_frame_amd.multigrid import GeometricMultigrid
    print(f"recursive kcycle levels={krylov_levels} kind={kind} precision={precision}", flush=True)
    shape, spacing = (32, 16, 16), (1.0, 1.0, 1.0)
    ke = element.stiffness(spacing, element.isotropic(0.3))
    rho = np.clip(np.random.default_rng(5).uniform(0.3, 1.2, int(np.prod(shape))), 0, 1)
    mod = 1e-6 + (1 - 1e-6) * rho ** 3
    pts = element.node_points(shape, spacing)
    fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    mg = GeometricMultigrid(shape, spacing, ke, {"levels": 5, "precision": precision, "coarse_krylov": 2,
                                                 "krylov_levels": krylov_levels, "growth": 1.0,
                                                 "chebyshev_kind": kind})
    mg.update(mod)
    entry = mg.hierarchy("k", fixed)
    levels = gmg.hierarchy(shape, ke, mod, fixed, (5, shape))
    for cols in (1, 2, 7):
        r = np.random.default_rng(6).normal(size=(levels[0]["A"].shape[0], cols)) * levels[0]["mask"][:, None]
        ref = np.column_stack([gmg.kcycle(levels, entry["lambda_max"], r[:, c], 2, 3, 1.0, krylov_levels, kind)
                               for c in range(cols)])
        z = mg.precondition_device("k", torch_gpu.as_tensor(r.T.copy(), device="cuda")).cpu().numpy().T
        tol = 1e-9 if precision == "float64" else 1e-4
        for c in range(cols):
            assert np.linalg.norm(z[:, c] - ref[:, c]) < tol * np.linalg.norm(ref[:, c]), (cols, c)
    mg.close()
