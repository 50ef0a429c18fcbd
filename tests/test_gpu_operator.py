"""HIP operator parity against the oracle's assembled CSR (mirrors the reference's
tests/test_topology_multigrid.py:24-38: operator equals the assembled matrix to 1e-12,
full and Dirichlet-reduced)."""
import numpy as np
import pytest

from deep_frame_amd import element, problems
from oracle import gmg, hex8, system

pytestmark = pytest.mark.gpu


def setup_case(shape, seed, torch, levels=0, settings=None):
    from deep_frame_amd.multigrid import GeometricMultigrid
    ke, _, _ = hex8.hex8((1.0, 1.0, 1.0), hex8.isotropic_d(0.3))
    rng = np.random.default_rng(seed)
    rho = np.clip(rng.uniform(0.3, 1.2, int(np.prod(shape))), 0, 1)
    mod = 1e-6 + (1 - 1e-6) * rho ** 3
    mg = GeometricMultigrid(shape, (1.0, 1.0, 1.0), ke, {"coarsest_dofs": 200, "levels": levels, **(settings or {})})
    mg.update(mod)
    return ke, mod, mg


@pytest.mark.parametrize("shape", [(8, 3, 5), (4, 8, 4), (16, 8, 8), (13, 7, 9)])
def test_operator_matches_assembled_matrix(shape, torch_gpu):
    print(f"operator shape={shape}", flush=True)
    ke, mod, mg = setup_case(shape, 0, torch_gpu)
    pts, conn, dofs = hex8.grid(shape)
    k = hex8.assemble(dofs, ke, mod, 3 * len(pts))
    vec = np.random.default_rng(1).normal(size=(3 * len(pts), 3))
    ref = k @ vec
    out = mg.product(vec)
    assert np.linalg.norm(out - ref) < 1e-12 * np.linalg.norm(ref)
    out32 = mg.operator_device(torch_gpu.as_tensor(vec.T.copy(), device="cuda"), precision="float32").cpu().numpy().T
    assert np.linalg.norm(out32 - ref) < 2e-6 * np.linalg.norm(ref)
    # Dirichlet-reduced operator with the x = 0 face clamped
    fixn = np.flatnonzero(pts[:, 0] == 0)
    fixed = (3 * fixn[:, None] + np.arange(3)).ravel()
    free = np.setdiff1d(np.arange(3 * len(pts)), fixed)
    mg.hierarchy("clamp", fixed)
    masked = np.zeros_like(vec)
    masked[free] = vec[free]
    red = mg.operator_device(torch_gpu.as_tensor(masked.T.copy(), device="cuda"), key="clamp").cpu().numpy().T
    exp = k[free][:, free] @ vec[free]
    assert np.linalg.norm(red[free] - exp) < 1e-12 * np.linalg.norm(exp)
    assert not np.any(red[fixed])
    mg.close()


def test_element_energy_matches_oracle(torch_gpu):
    shape = (9, 6, 5)
    ke, mod, mg = setup_case(shape, 2, torch_gpu)
    _, _, dofs = hex8.grid(shape)
    u = np.random.default_rng(3).normal(size=(3 * int(np.prod(np.asarray(shape) + 1)), 2))
    ref = system.element_energy(dofs, ke, u).T
    out = mg.element_energy(u)
    assert np.max(np.abs(out - ref)) < 1e-12 * np.max(np.abs(ref))
    mg.close()


def test_product_of_rigid_translation_is_zero(torch_gpu):
    shape = (6, 5, 4)
    ke, mod, mg = setup_case(shape, 4, torch_gpu)
    n = int(np.prod(np.asarray(shape) + 1))
    for c in range(3):
        v = np.zeros((3 * n, 1))
        v[c::3] = 1.0
        assert np.max(np.abs(mg.product(v))) < 1e-13
    mg.close()


def test_galerkin_levels_equal_rap(torch_gpu):
    """Coarse operators equal P^T (M A M) P of the oracle (tests/test_topology_multigrid.py:40-53)."""
    shape = (8, 4, 8)
    ke, mod, mg = setup_case(shape, 0, torch_gpu, levels=3)
    pts, _, _ = hex8.grid(shape)
    fixn = np.flatnonzero(pts[:, 0] == 0)
    fixed = (3 * fixn[:, None] + np.arange(3)).ravel()
    entry = mg.hierarchy("rap", fixed)
    levels = gmg.hierarchy(shape, ke, mod, fixed, gmg.level_count(shape, 200))
    for index, lv in enumerate(levels):
        for cols in (1, 2):  # one column: the thread-per-node stencil form of the power iterations
            vec = np.random.default_rng(index).normal(size=(lv["A"].shape[0], cols)) * lv["mask"][:, None]
            ref = lv["A"] @ vec
            out = mg.level_operator("rap", index, torch_gpu.as_tensor(vec.T.copy(), device="cuda")).cpu().numpy().T
            assert np.linalg.norm(out - ref) < 1e-12 * np.linalg.norm(ref), (index, cols)
    ref_lam = [gmg.spectrum(lv) for lv in levels[:-1]]
    # 20 power steps from a different start: a lower bound within a few per cent
    assert all(0.9 * r <= g <= r * (1 + 1e-9) for g, r in zip(entry["lambda_max"], ref_lam))
    live = int(np.count_nonzero(levels[-1]["mask"]))
    assert entry["coarsest_dofs"] == live
    mg.close()


@pytest.mark.parametrize("shape,tile", [((13, 7, 9), (16, 4)), ((16, 8, 8), (8, 16)), ((6, 9, 37), (16, 2)),
                                        ((5, 20, 50), (8, 3)), ((7, 15, 64), (16, 7)), ((4, 3, 16), (8, 1))])
def test_matrix_core_operator_matches_assembled_matrix(shape, tile, torch_gpu):
    """Split-bf16 MFMA fine kernel (bf16x3): K x within 1e-4 of the FP64 CSR product (about 2^-16 per
    product), unmasked and Dirichlet-reduced, for 1 to 4 z blocks of 16 elements and both tile heights."""
    print(f"matrix-core operator shape={shape} tile={tile}", flush=True)
    ke, mod, mg = setup_case(shape, 5, torch_gpu)
    pts, conn, dofs = hex8.grid(shape)
    k = hex8.assemble(dofs, ke, mod, 3 * len(pts))
    fixn = np.flatnonzero(pts[:, 0] == 0)
    fixed = (3 * fixn[:, None] + np.arange(3)).ravel()
    free = np.setdiff1d(np.arange(3 * len(pts)), fixed)
    mg.hierarchy("clamp", fixed)
    vec = np.random.default_rng(6).normal(size=(3 * len(pts), 3))
    assert mg.tune_fine("clamp", 3, tile)[:2] == tile
    ref = k @ vec
    out = mg.operator_device(torch_gpu.as_tensor(vec.T.copy(), device="cuda"), precision="bf16x3").cpu().numpy().T
    for c in range(3):
        assert np.linalg.norm(out[:, c] - ref[:, c]) < 1e-4 * np.linalg.norm(ref[:, c]), c
    masked = np.zeros_like(vec)
    masked[free] = vec[free]
    red = mg.operator_device(torch_gpu.as_tensor(masked.T.copy(), device="cuda"), key="clamp",
                             precision="bf16x3").cpu().numpy().T
    exp = k[free][:, free] @ vec[free]
    assert np.linalg.norm(red[free] - exp) < 1e-4 * np.linalg.norm(exp)
    assert not np.any(red[fixed])
    # rigid translations are annihilated up to the split rounding of Ke (about 2^-17 per entry, summed
    # over the 24 entries of a row and the 8 elements of a node)
    n = len(pts)
    for c in range(3):
        v = np.zeros((1, 3 * n))
        v[0, c::3] = 1.0
        rig = mg.operator_device(torch_gpu.as_tensor(v, device="cuda"), precision="bf16x3").cpu().numpy()
        assert np.max(np.abs(rig)) < 2e-3 * np.max(np.abs(ke))
    mg.close()


def test_matrix_core_autotuner_picks_a_candidate(torch_gpu):
    shape = (32, 16, 24)
    ke, mod, mg = setup_case(shape, 7, torch_gpu)
    mg.hierarchy("free", np.zeros(0, dtype=np.int64))
    rows, chunk, ms = mg.tune_fine("free", 4)
    assert rows in (8, 16) and 1 <= chunk <= shape[0] + 1 and ms > 0
    mg.close()
