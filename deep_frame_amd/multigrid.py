"""GeometricMultigrid with the reference's call surface, backed by libdfmi (HIP).

Mirrors ``deep_frame/topology_multigrid.py`` ``GeometricMultigrid``: constructor
(:47-79), ``update`` (:97-100), ``hierarchy`` (:145-208), ``solve`` (:336-375),
``product`` (:455-457), ``element_energy`` (:458-459), ``eigenpairs`` (:494-540, in ``modal``),
``diagnostics`` (:541-542),
``close`` (:543-548) and the ``statistics`` list.  PyTorch only carries device
memory and the current stream; all arithmetic runs in libdfmi.
"""
import ctypes
from ctypes import byref, c_double, c_int32, c_void_p
from time import perf_counter

import numpy as np

from deep_frame_amd import _native

MULTIGRID = {"coarse": "galerkin", "coarsest_dofs": 80000, "max_levels": 8, "smoother": "chebyshev", "sweeps": 2,
             "damping": 1.0, "chebyshev_degree": 3, "chebyshev_ratio": 20.0, "growth": 2.0, "cycle": "V",
             "power_iterations": 20, "precision": "float32", "tolerance": 1e-8, "max_iterations": 2000,
             "batch": 64, "projection": True, "coarsest_projection": True, "dead_ratio": 1e-12, "zero_rhs": 1e-12, "levels": 0, "seed": 0,
             "refine": False, "inner_tolerance": 1e-2, "device": "cuda", "fine_smoother": "vector", "retries": 2,
             "coarse_krylov": 0, "krylov_levels": 1, "block_krylov": False, "chebyshev_kind": 1,
             "eigen_tolerance": 1e-8, "eigen_iterations": 300, "eigen_guard": 2, "eigen_refresh": 10,
             "eigen_drop": 1e-12, "eigen_seed": 0}
PRECISIONS = {"float64": 0, "float32": 1}
# MULTIGRID["smoother"], ["cycle"] and ["coarse"] as the library's codes (topology_multigrid.py:51)
SMOOTHERS = {"chebyshev": 0, "jacobi": 1}
CYCLES = {"V": 0, "W": 1}
COARSE = {"galerkin": 0, "rediscretize": 1}
# operator precisions of dfmi_apply: "bf16x3" is the V-cycle's fine-level matrix-core kernel (FP32 vectors,
# three split-bf16 products per element block)
APPLY_PRECISIONS = {**PRECISIONS, "bf16x3": 2}
# fine-level smoother and residual of the fp32 V-cycle: on the matrix cores or on the vector units
FINE_SMOOTHERS = {"vector": 0, "matrix": 1}
SLOTS = 8


def _torch():
    import torch
    return torch


def rigid_modes(points, center):
    """The six rigid modes at node positions (n, 3) about ``center`` -> (6, 3n): translations x, y, z and
    rotations (0, -z, y), (z, 0, -x), (-y, x, 0) (topology_multigrid.py:29-38)."""
    d = np.asarray(points, float) - np.asarray(center, float)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    modes = np.zeros((6, len(d), 3))
    for axis in range(3):
        modes[axis, :, axis] = 1.0
    modes[3, :, 1], modes[3, :, 2] = -z, y
    modes[4, :, 0], modes[4, :, 2] = z, -x
    modes[5, :, 0], modes[5, :, 1] = -y, x
    return modes.reshape(6, -1)


class GeometricMultigrid:
    def __init__(self, shape, spacing, ke, settings=None, me=None):
        torch = self.torch = _torch()
        self.settings = {**MULTIGRID, **(settings or {})}
        s = self.settings
        if (s["precision"] not in PRECISIONS or s["coarse"] not in COARSE or s["smoother"] not in SMOOTHERS
                or s["cycle"] not in CYCLES or s["fine_smoother"] not in FINE_SMOOTHERS):
            raise ValueError("Invalid multigrid settings")
        if not torch.cuda.is_available():
            raise RuntimeError("GeometricMultigrid needs a ROCm GPU (libdfmi has no CPU path)")
        self.device = torch.device(s["device"])
        self.shape = tuple(int(n) for n in shape)
        self.spacing = np.asarray(spacing, dtype=float)
        self.ke64 = np.ascontiguousarray(np.asarray(ke, dtype=float).reshape(24, 24))
        self.nodes = tuple(n + 1 for n in self.shape)
        self.center = np.asarray(self.shape) * self.spacing / 2
        self.ndof = 3 * int(np.prod(self.nodes))
        self.nelem = int(np.prod(self.shape))
        self.native = self._settings()
        handle = c_void_p()
        _native.call("dfmi_create", (c_int32 * 3)(*self.shape), (c_double * 3)(*self.spacing),
                     self.ke64.ctypes.data_as(ctypes.POINTER(c_double)), byref(self.native), byref(handle))
        self.handle = handle
        levels, padded = c_int32(), (c_int32 * 3)()
        _native.call("dfmi_levels", self.handle, byref(levels), padded)
        self.level_count, self.padded = levels.value, tuple(padded)
        self.moduli = None
        self.hierarchies, self.statistics = {}, []
        self._fixed_cache = {}
        self.me = None
        if me is not None:
            self.me = np.ascontiguousarray(np.asarray(me, dtype=float).reshape(24, 24))
            _native.call("dfmi_set_mass", self.handle, self.me.ctypes.data_as(ctypes.POINTER(c_double)))

    def _settings(self):
        s = self.settings
        n = _native.default_settings()
        n.coarsest_dofs, n.max_levels, n.levels = int(s["coarsest_dofs"]), int(s["max_levels"]), int(s["levels"])
        n.chebyshev_degree, n.chebyshev_ratio, n.growth = int(s["chebyshev_degree"]), float(s["chebyshev_ratio"]), float(s["growth"])
        n.power_iterations, n.precision = int(s["power_iterations"]), PRECISIONS[s["precision"]]
        n.tolerance, n.max_iterations, n.dead_ratio = float(s["tolerance"]), int(s["max_iterations"]), float(s["dead_ratio"])
        n.seed, n.refine, n.inner_tolerance = int(s["seed"]), int(bool(s["refine"])), float(s["inner_tolerance"])
        n.fine_smoother = FINE_SMOOTHERS[s["fine_smoother"]]
        n.retries = int(s["retries"])
        n.coarse_krylov = int(s["coarse_krylov"])
        n.block_krylov = int(bool(s["block_krylov"]))
        n.smoother, n.sweeps, n.damping = SMOOTHERS[s["smoother"]], int(s["sweeps"]), float(s["damping"])
        n.cycle, n.coarse = CYCLES[s["cycle"]], COARSE[s["coarse"]]
        n.krylov_levels = int(s["krylov_levels"])
        n.chebyshev_kind = int(s["chebyshev_kind"])
        return n

    def _stream(self):
        _native.call("dfmi_set_stream", self.handle, _native.stream_handle(self.torch, self.device))

    def _tensor(self, array, dtype=None):
        torch = self.torch
        return torch.as_tensor(np.ascontiguousarray(array), dtype=dtype or torch.float64, device=self.device)

    def update(self, moduli):
        """E per element (FP64, 0 = inactive); drops cached hierarchies (topology_multigrid.py:97-100)."""
        torch = self.torch
        m = moduli if isinstance(moduli, torch.Tensor) else self._tensor(np.asarray(moduli, float).ravel())
        m = m.to(device=self.device, dtype=torch.float64).contiguous().reshape(-1)
        if m.numel() != self.nelem:
            raise ValueError("moduli must have one value per element")
        self._stream()
        _native.call("dfmi_update", self.handle, _native.ptr(m))
        self.moduli = m
        self.hierarchies.clear()

    def hierarchy(self, key, fixed):
        """Build (or reuse) the Galerkin hierarchy for a fixture; returns its slot."""
        if self.moduli is None:
            raise RuntimeError("update(moduli) must precede solve")
        if key in self.hierarchies:
            return self.hierarchies[key]
        if len(self.hierarchies) >= SLOTS:
            self.hierarchies.pop(next(iter(self.hierarchies)))
        used = set(h["slot"] for h in self.hierarchies.values())
        slot = next(i for i in range(SLOTS) if i not in used)
        fixed = np.asarray(fixed, dtype=np.int64)
        if fixed.size and (fixed.min() < 0 or fixed.max() >= self.ndof):
            raise ValueError("fixed DOF index out of range")
        # the DOF flags are scattered on the device: only the fixed indices cross PCIe
        dev = self.torch.zeros(self.ndof, dtype=self.torch.uint8, device=self.device)
        if fixed.size:
            dev[self._tensor(fixed, self.torch.int64)] = 1
        coefficients = self.rigid_coefficients(fixed)
        nk = coefficients.shape[1] if self.settings["projection"] else 0
        self._stream()
        seconds = c_double()
        if nk:
            if not self.settings["coarsest_projection"]:
                raise ValueError("a floating hierarchy needs coarsest_projection (the coarsest operator is singular)")
            coef = np.ascontiguousarray(coefficients)
            _native.call("dfmi_build_hierarchy_kernel", self.handle, slot, _native.ptr(dev),
                         coef.ctypes.data_as(ctypes.POINTER(c_double)), nk, byref(seconds))
        else:
            _native.call("dfmi_build_hierarchy", self.handle, slot, _native.ptr(dev), byref(seconds))
        lam = (c_double * max(1, self.level_count - 1))()
        coarsest = c_int32()
        _native.call("dfmi_diagnostics", self.handle, slot, lam, byref(coarsest))
        check, checked = (c_double * self.level_count)(), c_int32()
        _native.call("dfmi_kernel_check", self.handle, slot, check, byref(checked))
        entry = {"slot": slot, "setup_s": seconds.value, "lambda_max": list(lam)[:self.level_count - 1],
                 "coarsest_dofs": coarsest.value, "fixed": fixed, "coefficients": coefficients, "kernel": nk,
                 "kernel_check": list(check)[:checked.value]}
        self.hierarchies[key] = entry
        return entry

    def _points(self, nodes=None):
        """Coordinates (n, 3) of all nodes, or of the given flat node indices."""
        if nodes is not None:
            return np.stack(np.unravel_index(np.asarray(nodes, dtype=np.int64), self.nodes), axis=1) * \
                np.asarray(self.spacing, float)
        axes = [np.arange(n) * h for n, h in zip(self.nodes, self.spacing)]
        return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)

    def rigid_coefficients(self, fixed):
        """Combinations (6, nk) of the rigid modes that vanish on the fixed DOFs of active nodes
        (topology_multigrid.py:151-159): null space of their Gram matrix over those DOFs."""
        fixed = np.asarray(fixed, dtype=np.int64)
        nodes = np.unique(fixed // 3)
        # a node is active when one of its (up to eight) elements has a positive modulus
        ijk = np.stack(np.unravel_index(nodes, self.nodes), axis=1)
        corners = np.array([(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)])
        elems = ijk[:, None, :] - corners[None]
        inside = np.all((elems >= 0) & (elems < np.asarray(self.shape)), axis=2)
        flat = np.ravel_multi_index(tuple(np.clip(elems, 0, np.asarray(self.shape) - 1).transpose(2, 0, 1)),
                                    self.shape)
        gathered = self.moduli[self._tensor(flat.ravel(), self.torch.int64)].cpu().numpy().reshape(flat.shape)
        active_nodes = nodes[np.any(inside & (gathered > 0), axis=1)]
        constrained = fixed[np.isin(fixed // 3, active_nodes)]
        if not constrained.size:
            return np.eye(6)
        nodes = np.unique(constrained // 3)
        modes = rigid_modes(self._points(nodes), self.center)
        local = np.searchsorted(nodes, constrained // 3) * 3 + constrained % 3
        fixed_modes = modes[:, local]
        values, vectors = np.linalg.eigh(fixed_modes @ fixed_modes.T)
        return vectors[:, values <= 1e-10 * max(values.max(), 1.0)]

    def rigid_flat(self, key):
        """Kernel modes of a hierarchy as masked flat DOF columns (ndof, nk), host."""
        entry = self.hierarchies[key]
        rigid = (rigid_modes(self._points(), self.center).T @ entry["coefficients"])
        ones = self.torch.ones((1, self.ndof), dtype=self.torch.float64, device=self.device)
        self._stream()
        _native.call("dfmi_mask", self.handle, entry["slot"], _native.ptr(ones), 1)
        return rigid * ones.cpu().numpy()[0][:, None]

    def solve_device(self, key, fixed, rhs):
        """rhs: (k, ndof) FP64 device tensor -> (x (k, ndof) device tensor, per-column reports
        (iterations, relative residual, converged, FP64 retry rounds))."""
        torch = self.torch
        entry = self.hierarchy(key, fixed)
        rhs = rhs.to(device=self.device, dtype=torch.float64).contiguous()
        k = rhs.shape[0]
        x = torch.empty_like(rhs)
        reports = (_native.Report * k)()
        self._stream()
        _native.call("dfmi_solve", self.handle, entry["slot"], _native.ptr(rhs), _native.ptr(x), k, reports)
        return x, [(r.iterations, r.relative_residual, bool(r.converged), r.retried) for r in reports]

    def last_history(self):
        """Largest relative recursive residual over the active columns after each PCG iteration of the last
        device solve, FP64 retries appended (the per-batch ``history`` of topology_multigrid.py:421-426)."""
        count = c_int32()
        _native.call("dfmi_history", self.handle, None, 0, byref(count))
        values = (c_double * max(1, count.value))()
        _native.call("dfmi_history", self.handle, values, count.value, byref(count))
        return list(values)[:count.value]

    def solve(self, key, fixed, forces, support=None):
        """Reference signature (topology_multigrid.py:336-375); forces (ndof, k) host FP64.

        With ``support`` (and ``projection``) the fixture floats: only fixed DOFs outside the support are
        Dirichlet, the support DOFs take the least-squares reactions that remove the forces' rigid-mode
        components, and the solution is shifted by the rigid motion that best fits zero displacement on the
        support (topology_multigrid.py:340-349, 368-370)."""
        started = perf_counter()
        floating = support is not None and bool(self.settings["projection"])
        dirichlet = np.setdiff1d(fixed, support) if floating else np.asarray(fixed)
        key = (key, floating)
        entry = self.hierarchy(key, dirichlet)
        fixed = dirichlet
        forces = np.array(forces, dtype=float).reshape(self.ndof, -1)
        if floating:
            rigid = self.rigid_flat(key)
            support = np.asarray(support)
            forces[support] += np.linalg.lstsq(rigid[support].T, -(rigid.T @ forces), rcond=None)[0]
        norms = np.linalg.norm(forces, axis=0)
        live = np.flatnonzero(norms > self.settings["zero_rhs"] * max(norms.max(initial=0.0), 1e-300))
        flat = np.zeros_like(forces)
        iterations, relative = np.zeros(forces.shape[1], dtype=int), np.zeros(forces.shape[1])
        batches = retried = 0
        consistency, history = [], []
        for start in range(0, len(live), self.settings["batch"]):
            cols = live[start:start + self.settings["batch"]]
            rhs = self._tensor(forces[:, cols].T)
            if entry["kernel"]:
                component = (c_double * len(cols))()
                self._stream()
                _native.call("dfmi_kernel_component", self.handle, entry["slot"], _native.ptr(rhs), len(cols),
                             None, component)
                consistency += list(component)
            x, reps = self.solve_device(key, fixed, rhs)
            history.append(self.last_history())
            flat[:, cols] = x.cpu().numpy().T
            iterations[cols] = [r[0] for r in reps]
            relative[cols] = [r[1] for r in reps]
            retried += sum(r[3] for r in reps)
            batches += 1
        converged = bool(np.all(relative[live] < self.settings["tolerance"])) if len(live) else True
        if floating:
            shift = np.linalg.lstsq(rigid[support], flat[support], rcond=None)[0]
            flat = flat - rigid @ shift
        report = {"iterations": iterations.tolist(), "relative_residual": relative.tolist(), "converged": converged,
                  "history": history, "batches": batches, "retried_columns": int(retried), "zero_rhs_columns": int(forces.shape[1] - len(live)),
                  "seconds": perf_counter() - started, "setup_s": entry.pop("setup_s", 0.0), "floating": floating,
                  "levels": self.level_count, "coarsest_dofs": entry["coarsest_dofs"],
                  "kernel": int(entry["kernel"]) if floating else 0, "rhs_kernel_component": consistency or None,
                  "kernel_check": entry["kernel_check"], "lambda_max": entry["lambda_max"]}
        self.statistics.append(report)
        return flat, report

    def operator_device(self, x, key=None, precision="float64"):
        """K x (key None) or M K M x with a hierarchy's fine mask; x (k, ndof) FP64 device tensor."""
        x = x.to(device=self.device, dtype=self.torch.float64).contiguous()
        y = self.torch.empty_like(x)
        slot = -1 if key is None else self.hierarchies[key]["slot"]
        self._stream()
        _native.call("dfmi_apply", self.handle, slot, APPLY_PRECISIONS[precision], _native.ptr(x), _native.ptr(y),
                     x.shape[0])
        return y

    def mass_device(self, key, mass, lumped, x):
        """M x with M = sum_e m_e Me + diag(lumped), masked by a hierarchy's fine mask (the mass operator of
        topology_multigrid.py:491); mass (Ne,), lumped (ndof,) or None, x (k, ndof) FP64 device tensors."""
        torch = self.torch
        if self.me is None:
            raise RuntimeError("GeometricMultigrid was built without the element mass matrix (me)")
        x = x.to(device=self.device, dtype=torch.float64).contiguous()
        mass = mass.to(device=self.device, dtype=torch.float64).contiguous()
        lump = None if lumped is None else lumped.to(device=self.device, dtype=torch.float64).contiguous()
        y = torch.empty_like(x)
        self._stream()
        _native.call("dfmi_mass_apply", self.handle, self.hierarchies[key]["slot"], _native.ptr(mass),
                     None if lump is None else _native.ptr(lump), _native.ptr(x), _native.ptr(y), x.shape[0])
        return y

    def eigenpairs(self, key, fixed, mass, lumped, count, initial=None):
        """Lowest eigenpairs of (K, M) by V-cycle preconditioned LOBPCG (topology_multigrid.py:494-540):
        (eigenvalues (count,) host, vectors (count + eigen_guard, ndof) device tensor, report)."""
        from deep_frame_amd import modal
        return modal.eigenpairs(self, key, fixed, mass, lumped, count, initial)

    def tune_fine(self, key, k, tile=None):
        """Run the matrix-core fine-kernel autotuner for k columns, or set ``tile`` = (rows, x chunk) for k;
        returns (tile rows, x chunk, ms per launch)."""
        tile, ms = (c_int32 * 2)(*(tile or (0, 0))), c_double()
        self._stream()
        _native.call("dfmi_tune_fine", self.handle, self.hierarchies[key]["slot"], int(k), tile, byref(ms))
        return int(tile[0]), int(tile[1]), ms.value

    def level_operator(self, key, level, x):
        """Level operator on the padded node grid of ``level``; x (k, 3 * nodes_l) FP64 device tensor."""
        x = x.to(device=self.device, dtype=self.torch.float64).contiguous()
        y = self.torch.empty_like(x)
        self._stream()
        _native.call("dfmi_level_apply", self.handle, self.hierarchies[key]["slot"], level, _native.ptr(x),
                     _native.ptr(y), x.shape[0])
        return y

    def precondition_device(self, key, r):
        """One V-cycle z = M^-1 r on (k, ndof) FP64 device columns."""
        r = r.to(device=self.device, dtype=self.torch.float64).contiguous()
        z = self.torch.empty_like(r)
        self._stream()
        _native.call("dfmi_precondition", self.handle, self.hierarchies[key]["slot"], _native.ptr(r), _native.ptr(z),
                     r.shape[0])
        return z

    def transfer(self, key, level, direction, x, out_size):
        """Restriction (direction 0) or prolongation (1) between padded levels ``level`` and ``level + 1``."""
        x = x.to(device=self.device, dtype=self.torch.float64).contiguous()
        y = self.torch.empty((x.shape[0], out_size), dtype=self.torch.float64, device=self.device)
        self._stream()
        _native.call("dfmi_transfer", self.handle, self.hierarchies[key]["slot"], level, direction, _native.ptr(x),
                     _native.ptr(y), x.shape[0])
        return y

    def product(self, flat):
        """FP64 K u for host columns (ndof, k) (topology_multigrid.py:455-457)."""
        flat = np.asarray(flat, dtype=float).reshape(self.ndof, -1)
        return self.operator_device(self._tensor(flat.T)).cpu().numpy().T

    def element_energy_device(self, u):
        u = u.to(device=self.device, dtype=self.torch.float64).contiguous()
        out = self.torch.empty((u.shape[0], self.nelem), dtype=self.torch.float64, device=self.device)
        self._stream()
        _native.call("dfmi_element_energy", self.handle, _native.ptr(u), u.shape[0], _native.ptr(out))
        return out

    def element_energy(self, flat):
        """u_e^T Ke u_e per column and element, (k, Ne) host (topology_multigrid.py:458-459)."""
        flat = np.asarray(flat, dtype=float).reshape(self.ndof, -1)
        return self.element_energy_device(self._tensor(flat.T)).cpu().numpy()

    def compliance_sensitivity_device(self, u, f, dmoduli, factor=1.0):
        """(compliance (k,), sensitivity (Ne,)) device tensors (topology_multigrid.py:568-579)."""
        torch = self.torch
        u = u.to(device=self.device, dtype=torch.float64).contiguous()
        f = f.to(device=self.device, dtype=torch.float64).contiguous()
        d = dmoduli.to(device=self.device, dtype=torch.float64).contiguous()
        comp = torch.empty(u.shape[0], dtype=torch.float64, device=self.device)
        sens = torch.empty(self.nelem, dtype=torch.float64, device=self.device)
        self._stream()
        _native.call("dfmi_compliance_sensitivity", self.handle, _native.ptr(u), _native.ptr(f), u.shape[0],
                     _native.ptr(d), float(factor), _native.ptr(comp), _native.ptr(sens))
        return comp, sens

    def diagnostics(self):
        return {"settings": self.settings, "levels": [{"shape": [n >> l for n in self.padded],
                                                      "spacing_mm": (self.spacing * 2 ** l).tolist()}
                                                     for l in range(self.level_count)],
                "padded_shape": list(self.padded), "solves": self.statistics[-64:]}

    def close(self):
        errors = []
        if getattr(self, "handle", None) is not None and self.handle.value:
            try:
                _native.call("dfmi_destroy", self.handle)
            except RuntimeError as error:
                errors.append(str(error))
            self.handle = None
        return errors

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
