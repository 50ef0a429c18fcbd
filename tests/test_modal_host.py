"""Host logic of the LOBPCG eigensolver and the modal constraint (deep_frame_amd/modal.py) on the CPU:
the device passes (K V, M V, V-cycle, mask, Gram matrices, block combinations, modal sensitivity) are
replaced by assembled sparse matrices of the oracle, so that the block algebra, the M-orthonormalisation and
Rayleigh-Ritz bookkeeping, the refresh and convergence logic and the aggregate are checked against
shift-invert eigsh without a GPU.  The device passes themselves are checked in test_gpu_modal.py."""
import types

import numpy as np
import pytest
import torch
from scipy.sparse import diags
from scipy.sparse.linalg import splu

from deep_frame_amd import modal, problems
from oracle import hex8, system
from oracle import modal as omodal


class FakeLibrary:
    """dfmi_gram / dfmi_combine / dfmi_mask / dfmi_modal_sensitivity on CPU tensors."""

    def __init__(self, mg):
        self.mg = mg

    def call(self, name, *args):
        if name == "dfmi_gram":
            a, b, n, ka, kb, out = args[:6]
            out.view(-1)[:ka * kb] = (a @ b.T).reshape(-1)
        elif name == "dfmi_combine":
            u, n, k, mat, mcols, y, accumulate = args[:7]
            value = mat.T @ u
            y.copy_(y + value if accumulate else value)
        elif name == "dfmi_mask":
            _, _, v, k = args
            v.mul_(torch.as_tensor(self.mg.free))
        elif name == "dfmi_modal_sensitivity":
            _, u, k, weight, value, ss, ms, target, out = args
            u = u.numpy()
            acc = np.zeros(self.mg.nelem)
            for r in range(k):
                e = u[r][self.mg.dofs]
                strain = np.einsum("ei,ij,ej->e", e, self.mg.ke64, e)
                kinetic = np.einsum("ei,ij,ej->e", e, self.mg.me, e)
                acc += weight[r] * (ss.numpy() * strain - value[r] * ms.numpy() * kinetic)
            out.copy_(torch.as_tensor(acc / target))
        else:
            raise AssertionError(name)

    @staticmethod
    def ptr(t):
        assert t.is_contiguous()
        return t

    @staticmethod
    def stream_handle(torch, device=None):
        return None


class FakeMultigrid:
    """Assembled stand-ins for the device operators; the preconditioner is an exact solve of the masked
    stiffness (the V-cycle approximates it)."""

    def __init__(self, shape, h, ke, me, conn, dofs, settings=None):
        self.torch, self.device = torch, torch.device("cpu")
        self.settings = {**modal.EIGEN, **(settings or {})}
        self.ke64, self.me, self.conn, self.dofs = ke, me, conn, dofs
        self.nelem, self.ndof = int(np.prod(shape)), 3 * int(np.prod([n + 1 for n in shape]))
        self.level_count, self.statistics, self.handle = 3, [], None

    def update(self, moduli):
        self.moduli = np.asarray(moduli, float)

    def hierarchy(self, key, fixed):
        self.free = np.zeros(self.ndof)
        self.free[system.free_dofs(self.conn, self.moduli, fixed, self.ndof // 3)] = 1.0
        k = hex8.assemble(self.dofs, self.ke64, self.moduli, self.ndof)
        self.k = diags(self.free) @ k @ diags(self.free)
        self.lu = splu((self.k + diags(1.0 - self.free)).tocsc())
        return {"slot": 0, "coarsest_dofs": 0, "lambda_max": [], "setup_s": 0.0}

    def _stream(self):
        pass

    def _tensor(self, array):
        return torch.as_tensor(np.ascontiguousarray(array), dtype=torch.float64)

    def operator_device(self, v, key, precision):
        return torch.as_tensor((self.k @ v.numpy().T).T.copy())

    def mass_device(self, key, mass, lumped, v):
        m = hex8.assemble(self.dofs, self.me, mass.numpy(), self.ndof)
        if lumped is not None:
            m = m + diags(lumped.numpy())
        m = diags(self.free) @ m @ diags(self.free)
        return torch.as_tensor((m @ v.numpy().T).T.copy())

    def precondition_device(self, key, r):
        return torch.as_tensor(self.lu.solve(r.numpy().T * self.free[:, None]).T.copy())


@pytest.fixture
def fake(monkeypatch):
    def make(shape, h):
        ke, me, _ = hex8.hex8(h, hex8.isotropic_d(0.3))
        pts, conn, dofs = hex8.grid(shape, h)
        mg = FakeMultigrid(shape, h, ke, me, conn, dofs)
        monkeypatch.setattr(modal, "_native", FakeLibrary(mg))
        fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
        return mg, ke, me, pts, conn, dofs, fixed
    return make


def test_lobpcg_host_logic_matches_shift_invert(fake):
    mg, ke, me, pts, conn, dofs, fixed = fake((10, 2, 3), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(2)
    ne = mg.nelem
    moduli = 2000.0 * (1e-6 + (1 - 1e-6) * rng.uniform(0.3, 1.0, ne) ** 3)
    mass = 1.2e-9 * rng.uniform(0.3, 1.0, ne)
    lumped = np.where(pts[:, 0] > 9.5, 1e-7, 0.0).repeat(3)
    mg.update(moduli)
    values, vectors, report = modal.eigenpairs(mg, "k", fixed, mass, lumped, 3)
    assert report["converged"] and report["iterations"] > 0
    exact, _ = omodal.eigenpairs(dofs, conn, ke, me, moduli, mass, fixed, len(pts), 3, lumped)
    assert np.max(np.abs(values - exact) / exact) < 1e-9
    v = vectors.numpy()
    m = hex8.assemble(dofs, me, mass, mg.ndof) + diags(lumped)
    assert v.shape == (5, mg.ndof) and np.allclose(v @ (m @ v.T), np.eye(5), atol=1e-9)
    again, _, warm = modal.eigenpairs(mg, "k", fixed, mass, lumped, 3, vectors)
    assert warm["iterations"] == 0 and np.max(np.abs(again - exact) / exact) < 1e-9


def test_modal_constraint_host_logic_matches_oracle(fake):
    mg, ke, me, pts, conn, dofs, fixed = fake((8, 2, 3), (1.0, 1.0, 1.0))
    density = np.random.default_rng(4).uniform(0.05, 1.0, mg.nelem)
    lumped = np.where(pts[:, 0] > 7.5, 1e-7, 0.0).repeat(3)
    settings = {"f1_min_hz": 2000.0, "modes": 3, "tracked": 2, "ks": 40.0, "mass_cutoff": 0.1}
    constraint = modal.MultigridModal(mg, [{"key": "m", "fixed": fixed}], 2000.0, 1.2, lumped, settings=settings)
    violation, gradient, info = constraint.measure(density)
    moduli = 2000.0 * (1e-6 + (1 - 1e-6) * density ** 3)
    mass = 1.2e-9 * omodal.interpolation(density, 0.1)[0]
    exact, vectors = omodal.eigenpairs(dofs, conn, ke, me, moduli, mass, fixed, len(pts), 3, lumped)
    target = (2 * np.pi * 2000.0) ** 2
    ev, eg, f1, _ = omodal.aggregate(dofs, ke, me, [(exact[i], vectors[:, i]) for i in range(2)], density, 2000.0,
                                     1.2, target, 40.0, 0.1)
    assert violation == pytest.approx(ev, rel=1e-9) and info["f1_hz"] == pytest.approx(f1, rel=1e-9)
    assert np.linalg.norm(gradient - eg) < 1e-7 * np.linalg.norm(eg)


def test_modal_constraint_lumped_slope_and_augmented_call(fake):
    """``lumped_slope`` subtracts sum_i w_i lambda_i slope(u_i) / target from the sensitivity of KS, and the
    call form returns the augmented-Lagrangian penalty of the violation with the multiplier it used."""
    mg, ke, me, pts, conn, dofs, fixed = fake((8, 2, 3), (1.0, 1.0, 1.0))
    density = np.random.default_rng(5).uniform(0.05, 1.0, mg.nelem)
    settings = {"f1_min_hz": 2000.0, "modes": 3, "tracked": 2, "ks": 40.0, "mass_cutoff": 0.1, "penalty": 10.0,
                "multiplier_interval": 1}
    plain = modal.MultigridModal(mg, [{"key": "m", "fixed": fixed}], 2000.0, 1.2, settings=settings)
    violation, gradient, info = plain.measure(density)
    field = np.random.default_rng(6).uniform(0.0, 1.0, mg.nelem)
    hooked = modal.MultigridModal(mg, [{"key": "m", "fixed": fixed}], 2000.0, 1.2, settings=settings)
    seen = []
    hooked.lumped_slope = lambda mode: (seen.append(mode.shape), field)[1]
    _, hooked_gradient, _ = hooked.measure(density)
    values = np.array([f for _, f in info["frequencies_hz"]])
    values = (2 * np.pi * values) ** 2
    target = (2 * np.pi * 2000.0) ** 2
    ratios = values / target
    weights = np.exp(-40.0 * (ratios - ratios.min()))
    weights /= weights.sum()
    assert seen == [(mg.ndof,)] * 2
    assert np.allclose(hooked_gradient - gradient, np.sum(weights * values) * field / target, rtol=1e-7, atol=0)
    penalty, augmented, details = plain(density)
    shifted = max(violation, 0.0)
    assert penalty == pytest.approx(5.0 * shifted ** 2, rel=1e-9) and details["multiplier"] == 0.0
    assert np.allclose(augmented, 10.0 * shifted * gradient, rtol=1e-7, atol=1e-300)
