"""Mass operator, LOBPCG eigenpairs and the modal constraint on the GPU (deep_frame/topology_multigrid.py:
460-540, :583-601; deep_frame/topology_optimization.py:904-967) against the oracle (oracle/modal.py:
shift-invert eigsh of the assembled pencil, pinned to the documented 867.609 Hz first frequency and to
central differences of the aggregate)."""
import numpy as np
import pytest

from deep_frame_amd import problems
from oracle import hex8
from oracle import modal as omodal
from oracle import system

pytestmark = pytest.mark.gpu


def _cantilever(shape, h, nu=0.3):
    ke, me, _ = hex8.hex8(h, hex8.isotropic_d(nu))
    pts, conn, dofs = hex8.grid(shape, h)
    fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    return ke, me, pts, conn, dofs, fixed


def test_mass_operator_and_mask_match_assembled(torch_gpu):
    from deep_frame_amd.multigrid import GeometricMultigrid
    torch = torch_gpu
    shape, h = (6, 4, 5), (1.0, 0.8, 1.2)
    ke, me, pts, conn, dofs, fixed = _cantilever(shape, h)
    rng = np.random.default_rng(1)
    ne, ndof = int(np.prod(shape)), 3 * len(pts)
    moduli = rng.uniform(0.5, 2.0, ne)
    moduli[rng.random(ne) < 0.2] = 0.0
    mass = rng.uniform(0.5, 1.5, ne) * (moduli > 0)
    lumped = np.where(rng.random(ndof) < 0.1, rng.uniform(0.1, 1.0, ndof), 0.0)
    mg = GeometricMultigrid(shape, h, ke, {"coarsest_dofs": 100}, me)
    mg.update(moduli)
    mg.hierarchy(b"m", fixed)
    free = np.zeros(ndof)
    free[system.free_dofs(conn, moduli, fixed, len(pts))] = 1.0
    x = rng.normal(size=(3, ndof))
    xd = torch.as_tensor(x, device="cuda")
    for lump in (None, lumped):
        m = hex8.assemble(dofs, me, mass, ndof)
        expect = free[:, None] * ((m @ (free[:, None] * x.T)) + (0 if lump is None else lump[:, None] * x.T))
        got = mg.mass_device(b"m", torch.as_tensor(mass, device="cuda"),
                             None if lump is None else torch.as_tensor(lump, device="cuda"), xd).cpu().numpy().T
        assert np.max(np.abs(got - expect)) < 1e-13 * np.max(np.abs(expect))
    from deep_frame_amd import _native
    y = xd.clone()
    _native.call("dfmi_mask", mg.handle, mg.hierarchies[b"m"]["slot"], _native.ptr(y), 3)
    assert np.array_equal(y.cpu().numpy(), x * free[None, :])
    mg.close()


def test_first_frequency_known_answer(torch_gpu):
    """Clamped 48 x 6 x 6 mm beam, 24 x 4 x 4 cells, E = 4430 MPa, nu = 0.3, 1.09 g/cm3: first eigenfrequency
    867.609 Hz (docs/topology_optimization.md; tests/test_topology_optimization.py:56-68)."""
    from deep_frame_amd.multigrid import GeometricMultigrid
    shape, h = (24, 4, 4), (2.0, 1.5, 1.5)
    ke, me, pts, conn, dofs, fixed = _cantilever(shape, h)
    ne = int(np.prod(shape))
    moduli, mass = np.full(ne, 4430.0), np.full(ne, 1.09e-9)
    mg = GeometricMultigrid(shape, h, ke, {"coarsest_dofs": 200}, me)
    mg.update(moduli)
    # the square section pairs the bending modes (1-2, 3-4): ask for whole pairs, so that no tracked mode
    # shares its eigenvalue with a guard vector
    values, vectors, report = mg.eigenpairs(b"beam", fixed, mass, None, 4)
    print("lobpcg", report["iterations"], report["relative_residual"], flush=True)
    assert report["converged"] and report["levels"] > 1
    assert omodal.frequencies(values)[0] == pytest.approx(867.609, abs=6e-4)
    exact, _ = omodal.eigenpairs(dofs, conn, ke, me, moduli, mass, fixed, len(pts), 4)
    assert np.max(np.abs(values - exact) / exact) < 1e-9
    v = vectors.cpu().numpy()
    m = hex8.assemble(dofs, me, mass, 3 * len(pts))
    assert np.allclose(v @ (m @ v.T), np.eye(v.shape[0]), atol=1e-9)
    assert np.all(v[:, fixed] == 0)
    mg.close()


def _modal_cantilever(shape, spacing, tip_mass_g, seed=3):
    """Cantilever clamped at x = 0 with a tip point mass spread over the last cell layer of nodes
    (tools/multigrid_study.py:303-308), SIMP density uniform(0.2, 1.3) clipped to [0, 1]."""
    h = (spacing,) * 3
    ke, me, pts, conn, dofs, fixed = _cantilever(shape, h)
    length = shape[0] * spacing
    tip = np.flatnonzero(pts[:, 0] >= length - spacing - 1e-6)
    lumped = np.zeros(len(pts))
    lumped[tip] = tip_mass_g * 1e-6 / len(tip)
    density = np.clip(np.random.default_rng(seed).uniform(0.2, 1.3, int(np.prod(shape))), 0, 1)
    return ke, me, pts, conn, dofs, fixed, np.repeat(lumped, 3), density


def test_lobpcg_modes_match_shift_invert(torch_gpu):
    """As tests/test_topology_multigrid.py:106-137: four modes of a SIMP cantilever with a tip mass, two
    tracked; eigenvalues to 1e-9, f1 to 1e-9 and the aggregate gradient to 1e-6 of the shift-invert
    reference; a warm start from the converged block takes fewer iterations."""
    from deep_frame_amd.modal import MultigridModal
    from deep_frame_amd.multigrid import GeometricMultigrid
    shape, young, rho_g = (16, 3, 6), 2000.0, 1.2
    ke, me, pts, conn, dofs, fixed, lumped, density = _modal_cantilever(shape, 1.0, 0.5)
    settings = {"f1_min_hz": 300.0, "modes": 4, "tracked": 2, "ks": 40.0, "mass_cutoff": 0.1}
    mg = GeometricMultigrid(shape, (1.0, 1.0, 1.0), ke, {"coarsest_dofs": 200}, me)
    modal = MultigridModal(mg, [{"key": b"modal+1", "fixed": fixed}], young, rho_g, lumped, settings=settings)
    violation, gradient, info = modal.measure(density, 3.0, 1e-6)
    report = modal.reports[0]
    print("lobpcg", report["iterations"], report["relative_residual"], flush=True)
    assert report["converged"] and report["levels"] > 1
    moduli = young * (1e-6 + (1 - 1e-6) * density ** 3)
    mass = rho_g * 1e-9 * omodal.interpolation(density, 0.1)[0]
    exact, vectors = omodal.eigenpairs(dofs, conn, ke, me, moduli, mass, fixed, len(pts), 4, lumped)
    assert np.max(np.abs(modal.parts[0]["values"] - exact) / exact) < 1e-9
    target = (2 * np.pi * 300.0) ** 2
    ev, eg, f1, _ = omodal.aggregate(dofs, ke, me, [(exact[i], vectors[:, i]) for i in range(2)], density, young,
                                     rho_g, target, 40.0, 0.1)
    assert abs(info["f1_hz"] / f1 - 1) < 1e-9
    assert violation == pytest.approx(ev, rel=1e-9, abs=1e-12)
    assert np.linalg.norm(gradient - eg) < 1e-6 * np.linalg.norm(eg)
    first = report["iterations"]
    modal.measure(density, 3.0, 1e-6)
    warm = modal.reports[0]
    assert warm["converged"] and warm["iterations"] < first
    assert np.max(np.abs(modal.parts[0]["values"] - exact) / exact) < 1e-9
    mg.close()
