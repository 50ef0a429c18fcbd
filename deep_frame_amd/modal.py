"""Lowest eigenpairs K v = lambda M v of a clamped structure, and the modal frequency constraint.

Mirrors ``deep_frame/topology_multigrid.py`` ``GeometricMultigrid.eigenpairs`` (:494-540) with its
helpers ``_norms`` / ``_svqb`` / ``_combine`` / ``_project`` / ``_rayleigh`` / ``_residual`` (:460-493),
and ``MultigridModal`` (:583-601) with ``ModalConstraint.interpolation`` / ``aggregate``
(deep_frame/topology_optimization.py:904-907, :949-967).

The method is block LOBPCG preconditioned by one V-cycle of the stiffness hierarchy: blocks [V, K V, M V]
of FP64 device columns; M-orthonormalisation by SVQB; Rayleigh-Ritz on [current, search, previous];
residuals K v - lambda M v; K V and M V recomputed from V every ``eigen_refresh`` steps and once converged;
search directions = V-cycle of the unconverged residuals, masked, M-orthogonalised (twice) against the
current block; the previous-direction block is the non-current part of the Ritz rotation.

Every pass over the DOFs runs in libdfmi: K V (``dfmi_apply``, FP64), M V (``dfmi_mass_apply``), the
V-cycle (``dfmi_precondition``), the mask (``dfmi_mask``), the Gram matrices V^T W (``dfmi_gram``) and
the block combinations sum_r V_r C_rm (``dfmi_combine``).  The Gram matrices are at most a few dozen
square; their eigen-decompositions are host arithmetic, as the reference's torch.linalg.eigh on them.
Random starting blocks come from numpy's generator with ``eigen_seed`` (the reference draws them with
torch's device generator; the converged pairs do not depend on the start).
"""
from ctypes import POINTER, c_double
from time import perf_counter

import numpy as np

from deep_frame_amd import _native

EIGEN = {"eigen_tolerance": 1e-8, "eigen_iterations": 300, "eigen_guard": 2, "eigen_refresh": 10,
         "eigen_drop": 1e-12, "eigen_seed": 0}
MODAL = {"f1_min_hz": 300.0, "modes": 6, "tracked": 3, "ks": 40.0, "mass_cutoff": 0.1}
GRAM_MAX = 64      # columns per operand of one dfmi_gram / dfmi_combine pass
GRAM_BLOCKS = 512  # workgroups of the Gram partial sums


class _Blocks:
    """Device primitives over row blocks (m, ndof) of FP64 columns."""

    def __init__(self, mg):
        self.mg, self.torch = mg, mg.torch
        self.work = None

    def _stream(self):
        return _native.stream_handle(self.torch, self.mg.device)

    def gram(self, a, b):
        """a @ b.T as a host (ka, kb) array."""
        torch = self.torch
        ka, kb, n = a.shape[0], b.shape[0], a.shape[1]
        out = np.zeros((ka, kb))
        if ka == 0 or kb == 0:
            return out
        if self.work is None:
            self.work = torch.empty(GRAM_BLOCKS * GRAM_MAX * GRAM_MAX, dtype=torch.float64, device=self.mg.device)
        part = torch.empty((GRAM_MAX, GRAM_MAX), dtype=torch.float64, device=self.mg.device)
        for i in range(0, ka, GRAM_MAX):
            for j in range(0, kb, GRAM_MAX):
                ai, bj = a[i:i + GRAM_MAX], b[j:j + GRAM_MAX]
                _native.call("dfmi_gram", _native.ptr(ai), _native.ptr(bj), n, ai.shape[0], bj.shape[0],
                             _native.ptr(part), _native.ptr(self.work), GRAM_BLOCKS, self._stream())
                out[i:i + ai.shape[0], j:j + bj.shape[0]] = (
                    part.reshape(-1)[:ai.shape[0] * bj.shape[0]].reshape(ai.shape[0], bj.shape[0]).cpu().numpy())
        return out

    def combine(self, u, mat, out=None):
        """mat.T @ u (rows sum_r mat[r, m] u_r) -> (m, ndof); added into ``out`` when given."""
        torch = self.torch
        mat = np.ascontiguousarray(mat, dtype=float)
        k, m = mat.shape
        accumulate = out is not None
        if out is None:
            out = torch.empty((m, u.shape[1]), dtype=torch.float64, device=self.mg.device)
            if k == 0:
                out.zero_()
        if k == 0 or m == 0:
            return out
        for r in range(0, k, GRAM_MAX):
            chunk = torch.as_tensor(np.ascontiguousarray(mat[r:r + GRAM_MAX]), device=self.mg.device)
            _native.call("dfmi_combine", _native.ptr(u[r:r + GRAM_MAX]), u.shape[1], chunk.shape[0],
                         _native.ptr(chunk), m, _native.ptr(out), int(accumulate or r > 0), self._stream())
        return out


def _settings(mg):
    return {**EIGEN, **{k: v for k, v in mg.settings.items() if k in EIGEN}}


def eigenpairs(mg, key, fixed, mass, lumped, count, initial=None):
    """Lowest ``count`` eigenpairs of the masked pencil (K, M) for the fixture (key, fixed).

    mass: element field m_e (Ne,), lumped: flat DOF diagonal (ndof,) or None (host arrays or device
    tensors), initial: a previous (count + eigen_guard, ndof) block to start from.  Returns (eigenvalues
    (count,) host, M-orthonormal vectors (count + eigen_guard, ndof) FP64 device tensor, report)
    (topology_multigrid.py:494-540)."""
    torch = mg.torch
    settings = _settings(mg)
    tolerance, drop = settings["eigen_tolerance"], settings["eigen_drop"]
    if mg.me is None:
        raise RuntimeError("GeometricMultigrid was built without the element mass matrix (me)")
    if mg.device.type == "cuda":
        torch.cuda.synchronize(mg.device)
    started = perf_counter()
    entry = mg.hierarchy(key, np.asarray(fixed, dtype=np.int64))
    slot = entry["slot"]
    ops = _Blocks(mg)
    mass = torch.as_tensor(np.asarray(mass, float) if not isinstance(mass, torch.Tensor) else mass,
                           dtype=torch.float64, device=mg.device).reshape(-1).contiguous()
    if mass.numel() != mg.nelem:
        raise ValueError("mass must have one value per element")
    if lumped is not None:
        lumped = torch.as_tensor(np.asarray(lumped, float) if not isinstance(lumped, torch.Tensor) else lumped,
                                 dtype=torch.float64, device=mg.device).reshape(-1).contiguous()
        if lumped.numel() != mg.ndof:
            raise ValueError("lumped must have one value per DOF")

    def masked(v):
        v = v.contiguous()
        mg._stream()
        _native.call("dfmi_mask", mg.handle, slot, _native.ptr(v), v.shape[0])
        return v

    def apply(v):
        v = v.contiguous()
        if v.shape[0] == 0:
            return [v, v.clone(), v.clone()]
        return [v, mg.operator_device(v, key, "float64"), mg.mass_device(key, mass, lumped, v)]

    def norms(blocks):
        return np.diagonal(ops.gram(blocks[0], blocks[2])).copy() if blocks[0].shape[0] else np.zeros(0)

    def combine(blocks, coefficients):
        return [ops.combine(b, coefficients) for b in blocks]

    def svqb(blocks, reference=None):
        gram = ops.gram(blocks[0], blocks[2])
        gram = (gram + gram.T) / 2
        diagonal = np.diagonal(gram)
        live = diagonal > drop * (diagonal.max(initial=0.0) if reference is None else reference)
        coefficients = np.zeros((len(diagonal), 0))
        if live.any():
            scale = 1 / np.sqrt(diagonal[live])
            values, rotation = np.linalg.eigh(scale[:, None] * gram[np.ix_(live, live)] * scale[None, :])
            keep = values > drop * values.max()
            coefficients = np.zeros((len(diagonal), int(keep.sum())))
            coefficients[live] = scale[:, None] * rotation[:, keep] / np.sqrt(values[keep])
        return combine(blocks, coefficients)

    def project(blocks, basis):
        if blocks[0].shape[0] == 0 or basis[0].shape[0] == 0:
            return blocks
        coefficients = ops.gram(basis[2], blocks[0])
        return [ops.combine(base, -coefficients, out=block.clone()) for block, base in zip(blocks, basis)]

    def rayleigh(blocks, size):
        stiffness, gram = ops.gram(blocks[0], blocks[1]), ops.gram(blocks[0], blocks[2])
        values, rotation = np.linalg.eigh((gram + gram.T) / 2)
        basis = rotation[:, values > drop * values.max()]
        basis = basis / np.sqrt(np.diagonal(basis.T @ gram @ basis))
        reduced = basis.T @ stiffness @ basis
        values, inner = np.linalg.eigh((reduced + reduced.T) / 2)
        return values[:size], basis @ inner[:, :size]

    def residual(blocks, values):
        res = ops.combine(blocks[2], -np.diag(values), out=blocks[1].clone())
        top = np.sqrt(np.maximum(np.diagonal(ops.gram(res, res)), 0.0))
        bottom = np.sqrt(np.diagonal(ops.gram(blocks[1], blocks[1])))
        return res, top / bottom

    size = count + settings["eigen_guard"]
    if initial is None or initial.shape[0] != size:
        rng = np.random.default_rng(settings["eigen_seed"])
        initial = torch.as_tensor(rng.random((size, mg.ndof)) - 0.5, device=mg.device)
    else:
        initial = initial.to(device=mg.device, dtype=torch.float64).clone()
    blocks = svqb(apply(masked(initial)))
    values, rotation = rayleigh(blocks, size)
    current, previous, step, history = combine(blocks, rotation), None, 0, []
    while True:
        res, relative = residual(current, values)
        if step % settings["eigen_refresh"] == 0 or bool((relative[:count] < tolerance).all()):
            current = apply(current[0])
            res, relative = residual(current, values)
        history.append(float(relative[:count].max()))
        if bool((relative[:count] < tolerance).all()) or step >= settings["eigen_iterations"]:
            break
        rows = torch.as_tensor(np.flatnonzero(relative >= tolerance), device=mg.device)
        search = apply(masked(mg.precondition_device(key, res.index_select(0, rows))))
        search = svqb(project(project(search, current), current), norms(search))
        parts = [current, search]
        if previous is not None:
            reference = norms(previous)
            previous = svqb(project(project(project(previous, current), search), search), reference)
            if previous[0].shape[0]:
                parts.append(previous)
        blocks = [torch.cat(group) for group in zip(*parts)]
        values, rotation = rayleigh(blocks, size)
        current = combine(blocks, rotation)
        rotation[:size] = 0
        previous = combine(blocks, rotation)
        step += 1
    blocks = svqb(apply(current[0]))
    values, rotation = rayleigh(blocks, size)
    current = combine(blocks, rotation)
    relative = residual(current, values)[1]
    if mg.device.type == "cuda":
        torch.cuda.synchronize(mg.device)
    report = {"iterations": step, "relative_residual": relative[:count].tolist(),
              "converged": bool((relative[:count] < tolerance).all()), "history": history,
              "seconds": perf_counter() - started, "setup_s": entry.pop("setup_s", 0.0), "levels": mg.level_count,
              "coarsest_dofs": entry["coarsest_dofs"], "kernel": entry.get("kernel") or None,
              "lambda_max": entry["lambda_max"]}
    mg.statistics.append(report)
    return values[:count], current[0], report


def frequencies(values):
    """Eigenfrequencies in Hz from eigenvalues in (rad/s)^2 (N/mm, tonne units)."""
    return np.sqrt(np.maximum(np.asarray(values, float), 0.0)) / (2 * np.pi)


class MultigridModal:
    """Lowest-frequency constraint on the GPU: eigenpairs per fixture part, KS aggregate of the tracked
    eigenvalue ratios lambda / (2 pi f1_min)^2 and its element sensitivity
    (topology_multigrid.py:583-601, topology_optimization.py:885-967).

    parts: list of {"key", "fixed", "sign"} (one per symmetry sign), lumped: flat DOF point-mass diagonal
    (ndof,) or None, young (MPa) and density (g/cm^3) of the material, active: element mask or None."""

    def __init__(self, mg, parts, young, density, lumped=None, active=None, settings=None):
        self.mg = mg
        self.settings = {**MODAL, **(settings or {})}
        if self.settings["tracked"] > self.settings["modes"]:
            raise ValueError("tracked modes exceed computed modes")
        self.target = (2 * np.pi * self.settings["f1_min_hz"]) ** 2
        self.young, self.density = float(young), float(density)
        self.parts = [{"sign": 1.0, "vectors": None, "values": None, **p} for p in parts]
        self.lumped = None if lumped is None else np.asarray(lumped, float).ravel()
        self.active = np.ones(mg.nelem, bool) if active is None else np.asarray(active, bool).ravel()
        self.reports = []

    def interpolation(self, density):
        """Mass interpolation and its slope: rho^6 / c^5 below the cutoff c, rho above
        (topology_optimization.py:904-907)."""
        cutoff = self.settings["mass_cutoff"]
        low = density < cutoff
        return np.where(low, density ** 6 / cutoff ** 5, density), np.where(low, 6 * density ** 5 / cutoff ** 5, 1.0)

    def measure(self, physical_density, penalization=3.0, min_stiffness_ratio=1e-6):
        """(violation, gradient, info) as MultigridModal.measure (topology_multigrid.py:587-601)."""
        mg, s = self.mg, self.settings
        density = np.asarray(physical_density, dtype=float).ravel()
        moduli = np.where(self.active, self.young * (min_stiffness_ratio + (1 - min_stiffness_ratio) * density ** penalization), 0)
        mass = np.where(self.active, self.density * 1e-9 * self.interpolation(density)[0], 0)
        mg.update(moduli)
        modes, self.reports = [], []
        for part in self.parts:
            values, vectors, report = eigenpairs(mg, part["key"], part["fixed"], mass, self.lumped, s["modes"],
                                                 part["vectors"])
            part["vectors"], part["values"] = vectors, values
            self.reports.append(report)
            modes += [(part, float(values[i]), vectors[i]) for i in range(s["tracked"])]
        return self.aggregate(density, modes, penalization, min_stiffness_ratio)

    def aggregate(self, density, modes, penalization, min_stiffness_ratio):
        """KS of the eigenvalue ratios, its sensitivity and the frequency report
        (topology_optimization.py:949-967); mode vectors are flat FP64 device rows."""
        mg, torch = self.mg, self.mg.torch
        ratios = np.array([value / self.target for _, value, _ in modes])
        lowest = ratios.min()
        exponents = np.exp(-self.settings["ks"] * (ratios - lowest))
        aggregate = lowest - np.log(exponents.sum()) / self.settings["ks"]
        weights = exponents / exponents.sum()
        stiffness_slope = np.where(self.active, self.young * (1 - min_stiffness_ratio) * penalization
                                   * density ** (penalization - 1), 0.0)
        mass_slope = np.where(self.active, self.density * 1e-9 * self.interpolation(density)[1], 0.0)
        vectors = torch.stack([v for _, _, v in modes]).contiguous()
        values = np.array([value for _, value, _ in modes])
        ss, ms = mg._tensor(stiffness_slope), mg._tensor(mass_slope)
        sensitivity = torch.empty(mg.nelem, dtype=torch.float64, device=mg.device)
        mg._stream()
        weights, values = np.ascontiguousarray(weights, float), np.ascontiguousarray(values, float)
        _native.call("dfmi_modal_sensitivity", mg.handle, _native.ptr(vectors), len(modes),
                     weights.ctypes.data_as(POINTER(c_double)), values.ctypes.data_as(POINTER(c_double)),
                     _native.ptr(ss), _native.ptr(ms), float(self.target), _native.ptr(sensitivity))
        violation = 1 - aggregate
        frequencies_hz = [[part["sign"], float(np.sqrt(max(value, 0)) / (2 * np.pi))] for part, value, _ in modes]
        info = {"f1_hz": float(np.sqrt(max(lowest * self.target, 0)) / (2 * np.pi)), "frequencies_hz": frequencies_hz,
                "aggregate_ratio": float(aggregate), "violation": float(violation)}
        return violation, -sensitivity.cpu().numpy(), info
