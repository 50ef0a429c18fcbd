"""Lowest eigenpairs K v = lambda M v of a clamped structure, and the modal frequency constraint.

The roles of ``GeometricMultigrid.eigenpairs`` (deep_frame/topology_multigrid.py:494-540, its settings
``eigen_*`` and report fields), ``MultigridModal`` (:583-601) and ``ModalConstraint``
(deep_frame/topology_optimization.py:885-967: mass interpolation, KS aggregate of the tracked eigenvalue
ratios, its sensitivity with the optional ``lumped_slope`` term, and the augmented-Lagrangian call).

The eigen-solver is this build's block LOBPCG (Knyazev, SIAM J. Sci. Comput. 23 (2001)) with soft locking,
Cholesky-QR M-orthonormalisation with an eigen-decomposition fallback, classical Gram-Schmidt twice for the
cross-block M-orthogonality, Knyazev's conjugate block, and a generalised Rayleigh-Ritz on the Gram matrices
(``_Lobpcg``).  Every pass over the DOFs runs in libdfmi: K V (``dfmi_apply``, FP64), M V
(``dfmi_mass_apply``), the V-cycle (``dfmi_precondition``), the mask (``dfmi_mask``), the Gram matrices
V^T W (``dfmi_gram``) and the block combinations sum_r V_r C_rm (``dfmi_combine``).  Random starting blocks
come from numpy's generator with ``eigen_seed``; the converged pairs do not depend on the start.
"""
from ctypes import POINTER, c_double
from time import perf_counter

import numpy as np
from scipy.linalg import LinAlgError, cholesky, eigh, solve_triangular

from deep_frame_amd import _native
from deep_frame_amd.constraints import AugmentedLagrangian, ks_minimum

EIGEN = {"eigen_tolerance": 1e-8, "eigen_iterations": 300, "eigen_guard": 2, "eigen_refresh": 10,
         "eigen_drop": 1e-12, "eigen_seed": 0}
MODAL = {"f1_min_hz": 300.0, "modes": 6, "tracked": 3, "ks": 40.0, "mass_cutoff": 0.1, "penalty": 1.0,
         "multiplier_interval": 1}
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


class _Span:
    """A block of m FP64 device rows v together with K v and M v (each (m, ndof))."""

    __slots__ = ("v", "kv", "mv")

    def __init__(self, v, kv, mv):
        self.v, self.kv, self.mv = v, kv, mv

    @property
    def size(self):
        return int(self.v.shape[0])


class _Lobpcg:
    """Block LOBPCG (Knyazev 2001) for the lowest eigenpairs of the masked pencil (K, M), preconditioned by
    one V-cycle of the stiffness hierarchy, with soft locking (converged columns stay in the basis but get no
    new search direction).  Each step:

    * residuals R = K X - M X diag(theta), relative |R_i| / |K x_i|;
    * search block W = V-cycle(R) of the unconverged columns, masked, made M-orthogonal to X by classical
      Gram-Schmidt twice, then M-orthonormal by Cholesky QR (an eigen-decomposition of the Gram matrix with
      relative drop ``eigen_drop`` when Cholesky fails or is ill-conditioned);
    * the conjugate block P (Knyazev's P = W C_W + P C_P of the previous step) made M-orthogonal to [X, W]
      the same way;
    * Rayleigh-Ritz on span[X, W, P]: the generalised eigenproblem (S^T K S, S^T M S) of the small Gram
      matrices, the lowest ``size`` Ritz pairs kept;
    * K X and M X recomputed from X every ``eigen_refresh`` steps and once converged, so the recurrences do
      not drift.

    Every pass over the DOFs is a libdfmi call (dfmi_apply, dfmi_mass_apply, dfmi_precondition, dfmi_mask,
    dfmi_gram, dfmi_combine); only the Gram matrices (a few dozen square) are host arithmetic."""

    def __init__(self, mg, key, slot, mass, lumped, settings):
        self.mg, self.key, self.slot = mg, key, slot
        self.mass, self.lumped = mass, lumped
        self.ops = _Blocks(mg)
        self.drop = settings["eigen_drop"]

    # ---- device passes -------------------------------------------------------------------------
    def masked(self, v):
        v = v.contiguous()
        self.mg._stream()
        _native.call("dfmi_mask", self.mg.handle, self.slot, _native.ptr(v), v.shape[0])
        return v

    def span(self, v):
        """v with K v and M v recomputed."""
        v = v.contiguous()
        if v.shape[0] == 0:
            return _Span(v, v.clone(), v.clone())
        return _Span(v, self.mg.operator_device(v, self.key, "float64"),
                     self.mg.mass_device(self.key, self.mass, self.lumped, v))

    def mapped(self, span, coefficients):
        """The rows sum_r c[r, j] row_r of v, K v and M v."""
        return _Span(*(self.ops.combine(part, coefficients) for part in (span.v, span.kv, span.mv)))

    def joined(self, spans):
        torch = self.mg.torch
        spans = [s for s in spans if s is not None and s.size]
        return _Span(*(torch.cat([getattr(s, name) for s in spans]) for name in ("v", "kv", "mv")))

    def energies(self, span):
        """x_i^T M x_i per row."""
        return np.diagonal(self.ops.gram(span.v, span.mv)).copy() if span.size else np.zeros(0)

    # ---- small dense steps -------------------------------------------------------------------------
    def m_orthonormal(self, span, reference=None):
        """Coefficients C with (span C)^T M (span C) = I over the rows whose M-energy exceeds ``eigen_drop``
        times ``reference`` (the largest energy when None); rows below it are dropped."""
        if not span.size:
            return span
        gram = self.ops.gram(span.v, span.mv)
        gram = 0.5 * (gram + gram.T)
        energy = np.diagonal(gram)
        # a row that kept less than eigen_drop of its energy through the projections is numerically dependent
        floor = self.drop * (energy.max(initial=0.0) if reference is None else np.asarray(reference))
        rows = np.flatnonzero(energy > floor)
        coefficients = np.zeros((span.size, 0))
        if rows.size:
            inv_root = 1.0 / np.sqrt(energy[rows])
            scaled = inv_root[:, None] * gram[np.ix_(rows, rows)] * inv_root[None, :]
            try:
                factor = cholesky(scaled, lower=True)
                if np.min(np.diagonal(factor)) ** 2 <= self.drop:
                    raise LinAlgError("ill-conditioned block")
                local = solve_triangular(factor, np.eye(rows.size), lower=True).T
            except LinAlgError:
                lam, vec = np.linalg.eigh(scaled)
                good = lam > self.drop * lam.max()
                local = vec[:, good] / np.sqrt(lam[good])
            coefficients = np.zeros((span.size, local.shape[1]))
            coefficients[rows] = inv_root[:, None] * local
        return self.mapped(span, coefficients)

    def m_orthogonal_to(self, span, basis):
        """span minus its M-projection on the M-orthonormal basis, classical Gram-Schmidt applied twice."""
        if not span.size or basis is None or not basis.size:
            return span
        for _ in range(2):
            coefficients = -self.ops.gram(basis.mv, span.v)
            span = _Span(*(self.ops.combine(b, coefficients, out=s.clone())
                           for s, b in zip((span.v, span.kv, span.mv), (basis.v, basis.kv, basis.mv))))
        return span

    def ritz(self, span, size):
        """The lowest ``size`` Ritz values of (S^T K S, S^T M S) and their coefficient columns."""
        stiffness = self.ops.gram(span.v, span.kv)
        gram = self.ops.gram(span.v, span.mv)
        stiffness, gram = 0.5 * (stiffness + stiffness.T), 0.5 * (gram + gram.T)
        take = min(size, span.size)
        try:
            values, vectors = eigh(stiffness, gram, subset_by_index=[0, take - 1])
        except LinAlgError:
            lam, vec = np.linalg.eigh(gram)  # drop the numerically dependent directions first
            keep = vec[:, lam > self.drop * lam.max()] / np.sqrt(lam[lam > self.drop * lam.max()])
            values, inner = np.linalg.eigh(keep.T @ stiffness @ keep)
            values, vectors = values[:take], keep @ inner[:, :take]
        return values, vectors

    def residuals(self, span, values):
        res = self.ops.combine(span.mv, -np.diag(values), out=span.kv.clone())
        top = np.sqrt(np.maximum(np.diagonal(self.ops.gram(res, res)), 0.0))
        bottom = np.sqrt(np.maximum(np.diagonal(self.ops.gram(span.kv, span.kv)), 1e-300))
        return res, top / bottom

    # ---- iteration -----------------------------------------------------------------------------
    def run(self, start, count, size, settings):
        tol, refresh, limit = settings["eigen_tolerance"], settings["eigen_refresh"], settings["eigen_iterations"]
        x = self.m_orthonormal(self.span(self.masked(start)))
        values, c = self.ritz(x, size)
        x = self.mapped(x, c)
        conjugate, step, history = None, 0, []
        while True:
            res, rel = self.residuals(x, values)
            done = bool((rel[:count] < tol).all())
            if step % refresh == 0 or done:
                x = self.span(x.v)
                res, rel = self.residuals(x, values)
                done = bool((rel[:count] < tol).all())
            history.append(float(rel[:count].max()))
            if done or step >= limit:
                break
            open_rows = self.mg.torch.as_tensor(np.flatnonzero(rel >= tol), device=self.mg.device)
            raw = self.span(self.masked(self.mg.precondition_device(self.key, res.index_select(0, open_rows))))
            w = self.m_orthonormal(self.m_orthogonal_to(raw, x), self.energies(raw))
            if conjugate is not None:
                reference = self.energies(conjugate)
                conjugate = self.m_orthonormal(self.m_orthogonal_to(conjugate, self.joined([x, w])), reference)
                if not conjugate.size:
                    conjugate = None
            basis = self.joined([x, w, conjugate])
            values, c = self.ritz(basis, size)
            x_new = self.mapped(basis, c)
            # Knyazev's conjugate block: the W and P parts of the new Ritz vectors
            nx, nw = x.size, w.size
            new_conjugate = self.mapped(w, c[nx:nx + nw]) if nw else None
            if conjugate is not None:
                tail = self.mapped(conjugate, c[nx + nw:])
                new_conjugate = tail if new_conjugate is None else _Span(
                    *(a + b for a, b in zip((new_conjugate.v, new_conjugate.kv, new_conjugate.mv),
                                            (tail.v, tail.kv, tail.mv))))
            x, conjugate = x_new, new_conjugate
            step += 1
        # final Rayleigh-Ritz on freshly applied operators
        x = self.m_orthonormal(self.span(x.v))
        values, c = self.ritz(x, size)
        x = self.mapped(x, c)
        return values, x, self.residuals(x, values)[1], step, history


def eigenpairs(mg, key, fixed, mass, lumped, count, initial=None):
    """Lowest ``count`` eigenpairs of the masked pencil (K, M) for the fixture (key, fixed) by V-cycle
    preconditioned block LOBPCG (the role of topology_multigrid.py:494-540).

    mass: element field m_e (Ne,), lumped: flat DOF diagonal (ndof,) or None (host arrays or device
    tensors), initial: a previous (count + eigen_guard, ndof) block to start from.  Returns (eigenvalues
    (count,) host, M-orthonormal vectors (count + eigen_guard, ndof) FP64 device tensor, report)."""
    torch = mg.torch
    settings = _settings(mg)
    if mg.me is None:
        raise RuntimeError("GeometricMultigrid was built without the element mass matrix (me)")
    if mg.device.type == "cuda":
        torch.cuda.synchronize(mg.device)
    started = perf_counter()
    entry = mg.hierarchy(key, np.asarray(fixed, dtype=np.int64))

    def device(values, length, what):
        if values is None:
            return None
        out = torch.as_tensor(values if isinstance(values, torch.Tensor) else np.asarray(values, float),
                              dtype=torch.float64, device=mg.device).reshape(-1).contiguous()
        if out.numel() != length:
            raise ValueError(f"{what} must have one value per {'element' if length == mg.nelem else 'DOF'}")
        return out

    mass = device(mass, mg.nelem, "mass")
    lumped = device(lumped, mg.ndof, "lumped")
    size = count + settings["eigen_guard"]
    if initial is None or initial.shape[0] != size:
        start = torch.as_tensor(np.random.default_rng(settings["eigen_seed"]).random((size, mg.ndof)) - 0.5,
                                device=mg.device)
    else:
        start = initial.to(device=mg.device, dtype=torch.float64).clone()
    solver = _Lobpcg(mg, key, entry["slot"], mass, lumped, settings)
    values, x, relative, step, history = solver.run(start, count, size, settings)
    if mg.device.type == "cuda":
        torch.cuda.synchronize(mg.device)
    tolerance = settings["eigen_tolerance"]
    report = {"iterations": step, "relative_residual": relative[:count].tolist(),
              "converged": bool((relative[:count] < tolerance).all()), "history": history,
              "seconds": perf_counter() - started, "setup_s": entry.pop("setup_s", 0.0), "levels": mg.level_count,
              "coarsest_dofs": entry["coarsest_dofs"], "kernel": entry.get("kernel") or None,
              "lambda_max": entry["lambda_max"]}
    mg.statistics.append(report)
    return values[:count], x.v, report


def frequencies(values):
    """Eigenfrequencies in Hz from eigenvalues in (rad/s)^2 (N/mm, tonne units)."""
    return np.sqrt(np.maximum(np.asarray(values, float), 0.0)) / (2 * np.pi)


def mass_interpolation(density, cutoff):
    """Mass of a cell of density rho and its slope: rho above ``cutoff``, the C0 continuation
    rho^6 / cutoff^5 below it, which makes near-void cells light enough not to create spurious local
    modes (the interpolation of topology_optimization.py:904-907)."""
    rho = np.asarray(density, dtype=float)
    below = rho < cutoff
    scale = cutoff ** -5.0
    return (np.where(below, scale * rho ** 6, rho), np.where(below, 6.0 * scale * rho ** 5, 1.0))


class MultigridModal(AugmentedLagrangian):
    """Lowest-frequency constraint on the GPU: eigenpairs per fixture part, KS aggregate of the tracked
    eigenvalue ratios lambda / (2 pi f1_min)^2 and its element sensitivity, and the augmented-Lagrangian
    penalty of the violation 1 - KS (the roles of topology_multigrid.py:583-601 and
    topology_optimization.py:885-967).

    parts: list of {"key", "fixed", "sign"} (one per symmetry sign), lumped: flat DOF point-mass diagonal
    (ndof,) or None, young (MPa) and density (g/cm^3) of the material, active: element mask or None.
    ``lumped_slope`` (None, or a callable mode (ndof,) -> (Ne,)) adds the design dependence of the point
    masses: the sensitivity then subtracts lambda * lumped_slope(mode) per tracked mode."""

    def __init__(self, mg, parts, young, density, lumped=None, active=None, settings=None):
        self.mg = mg
        self.settings = {**MODAL, **(settings or {})}
        super().__init__(self.settings)
        if self.settings["tracked"] > self.settings["modes"]:
            raise ValueError("tracked modes exceed computed modes")
        self.target = (2 * np.pi * self.settings["f1_min_hz"]) ** 2
        self.young, self.density = float(young), float(density)
        self.parts = [{"sign": 1.0, "vectors": None, "values": None, **p} for p in parts]
        self.lumped = None if lumped is None else np.asarray(lumped, float).ravel()
        self.active = np.ones(mg.nelem, bool) if active is None else np.asarray(active, bool).ravel()
        self.lumped_slope = None
        self.reports = []

    def interpolation(self, density):
        return mass_interpolation(density, self.settings["mass_cutoff"])

    def _stiffness(self, density, penalization, min_stiffness_ratio):
        """SIMP moduli and their slope on active cells (0 elsewhere)."""
        solid = 1.0 - min_stiffness_ratio
        moduli = self.young * (min_stiffness_ratio + solid * density ** penalization)
        slope = self.young * solid * penalization * density ** (penalization - 1)
        return np.where(self.active, moduli, 0.0), np.where(self.active, slope, 0.0)

    def measure(self, physical_density, penalization=3.0, min_stiffness_ratio=1e-6):
        """(violation, gradient, info) of the tracked modes of every part."""
        mg, s = self.mg, self.settings
        density = np.asarray(physical_density, dtype=float).ravel()
        moduli, _ = self._stiffness(density, penalization, min_stiffness_ratio)
        mass = np.where(self.active, self.density * 1e-9 * self.interpolation(density)[0], 0.0)
        mg.update(moduli)
        modes, self.reports = [], []
        for part in self.parts:
            values, vectors, report = eigenpairs(mg, part["key"], part["fixed"], mass, self.lumped, s["modes"],
                                                 part["vectors"])
            part["vectors"], part["values"] = vectors, values
            self.reports.append(report)
            modes.extend((part, float(values[i]), vectors[i]) for i in range(s["tracked"]))
        return self.aggregate(density, modes, penalization, min_stiffness_ratio)

    def __call__(self, physical_density, penalization=3.0, min_stiffness_ratio=1e-6):
        """(penalty, gradient, info with the multiplier used) of ModalConstraint.__call__."""
        violation, slope, info = self.measure(physical_density, penalization, min_stiffness_ratio)
        penalty, gradient, multiplier = self.augment(violation, slope)
        return penalty, gradient, {**info, "multiplier": multiplier}

    def aggregate(self, density, modes, penalization, min_stiffness_ratio):
        """KS of the eigenvalue ratios, the element sensitivity of 1 - KS and the frequency report; mode
        vectors are flat FP64 device rows."""
        mg, torch = self.mg, self.mg.torch
        values = np.array([value for _, value, _ in modes])
        ratio, weights = ks_minimum(values / self.target, self.settings["ks"])
        _, stiffness_slope = self._stiffness(density, penalization, min_stiffness_ratio)
        mass_slope = np.where(self.active, self.density * 1e-9 * self.interpolation(density)[1], 0.0)
        vectors = torch.stack([v for _, _, v in modes]).contiguous()
        sensitivity = torch.empty(mg.nelem, dtype=torch.float64, device=mg.device)
        weights, values = np.ascontiguousarray(weights, float), np.ascontiguousarray(values, float)
        # the slope tensors must stay referenced until the call returns (their memory is the kernel's input)
        stiffness_dev, mass_dev = mg._tensor(stiffness_slope), mg._tensor(mass_slope)
        mg._stream()
        _native.call("dfmi_modal_sensitivity", mg.handle, _native.ptr(vectors), len(modes),
                     weights.ctypes.data_as(POINTER(c_double)), values.ctypes.data_as(POINTER(c_double)),
                     _native.ptr(stiffness_dev), _native.ptr(mass_dev), float(self.target), _native.ptr(sensitivity))
        gradient = sensitivity.cpu().numpy()
        if self.lumped_slope is not None:
            for w, (_, value, vector) in zip(weights, modes):
                gradient -= w * value * np.asarray(self.lumped_slope(vector.cpu().numpy()), float) / self.target
            gradient[~self.active] = 0.0
        hz = lambda value: float(np.sqrt(max(value, 0.0)) / (2 * np.pi))
        info = {"f1_hz": hz(values.min()), "frequencies_hz": [[part["sign"], hz(value)] for part, value, _ in modes],
                "aggregate_ratio": ratio, "violation": 1.0 - ratio}
        return 1.0 - ratio, -gradient, info
