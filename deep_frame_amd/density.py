"""Design-to-physical density map on the GPU: cone filter, single or robust Heaviside projection, volume,
pullback and the initial design.

Mirrors ``deep_frame/topology_optimization.py`` ``DensityMap`` (:1110-1195): ``filtered`` (:1135-1142),
``physical`` / ``fields`` (:1143-1147; robust thresholds eroded eta + delta, intermediate eta, dilated
eta - delta, :1114-1116), ``volume`` (:1148-1151; the dilated field when robust), ``project``
(:1152-1170), ``pullback`` (:1171-1178) and ``initial`` (:1179-1195, volume bisection); and the row
pullbacks of ``TopologyProblem.evaluate`` (deep_frame/topology_problem.py:366-379).  The filter is
libdfmi's cone kernel (``dfmi_filter_*``), the projection ``dfmi_project`` (all thresholds in one pass),
the weighted volume one ``dfmi_gram`` dot; the corridor blend (``settings["corridor"]``, :1125-1129, :1142,
:1176) is the filter's second, wider cone (``dfmi_filter_set_corridor``).  Volume weights of soft propeller
discs (``volume_weights``, :827-837) are computed on the host once.
"""
from ctypes import POINTER, c_double

import numpy as np

from deep_frame_amd import _native
from deep_frame_amd.simp import DensityFilter

DENSITY = {"filter_radius_mm": 6.0, "projection_beta": 1.0, "projection_eta": 0.5, "projection": "single",
           "robust_delta": 0.25, "beta_schedule": None, "minimum_design_density": 1e-3, "prop_discs": None,
           "corridor": None}


def volume_weights(points, discs):
    """Volume weight of each cell centre (n, 3): 1 outside the propeller discs; with soft discs
    (``discs["mode"] == "soft"``) 1 + weight * s * exp(-|z - plane| / length), where s is the largest disc
    profile over the motors: 1 inside a disc of ``radius_mm`` about a motor axis (x, y), or with
    ``radial_power`` the ramp ((r - hub) / (radius - hub))^power clipped to [0, 1], and 0 outside every disc
    (the soft-disc weighting of topology_optimization.py:827-837)."""
    p = np.asarray(points, dtype=float).reshape(-1, 3)
    if not discs or discs.get("mode") != "soft":
        return np.ones(len(p))
    motors = np.asarray(discs["motors_mm"], dtype=float).reshape(-1, 2)
    radius, hub = float(discs["radius_mm"]), float(discs.get("hub_radius_mm", 0.0))
    power = discs.get("radial_power")
    # distance of every cell centre to every motor axis: (motors, cells)
    r = np.hypot(p[None, :, 0] - motors[:, :1], p[None, :, 1] - motors[:, 1:])
    profile = np.ones_like(r) if power is None else np.clip((r - hub) / (radius - hub), 0.0, 1.0) ** power
    share = np.max(np.where(r <= radius, profile, 0.0), axis=0, initial=0.0)
    decay = np.exp(-np.abs(p[:, 2] - discs["plane_mm"]) / discs["length_mm"])
    return 1.0 + discs["weight"] * share * decay


class DensityMap:
    """Filter + projection of a design on one grid; design, filtered and physical fields are FP64 device
    tensors of the element count (C order)."""

    def __init__(self, shape, spacing, allowed, preserve, forbidden, settings=None, origin=(0.0, 0.0, 0.0)):
        import torch
        self.torch = torch
        self.settings = s = {**DENSITY, **(settings or {})}
        if s["projection"] not in ("single", "robust"):
            raise ValueError("Invalid topology projection")
        self.shape = tuple(int(n) for n in shape)
        self.n = int(np.prod(self.shape))
        masks = [np.asarray(m, dtype=bool).reshape(-1) for m in (allowed, preserve, forbidden)]
        if any(m.size != self.n for m in masks):
            raise ValueError("masks must have the grid shape")
        self.allowed, self.preserve, self.forbidden = masks
        if not np.array_equal(self.forbidden, ~self.allowed) or np.any(self.preserve & self.forbidden):
            raise ValueError("Forbidden must complement allowed; preserve must be allowed")
        self.free = self.allowed & ~self.preserve
        if not self.free.any():
            raise ValueError("The design domain has no free material cells")
        self.robust = s["projection"] == "robust"
        eta, delta = s["projection_eta"], s["robust_delta"]
        if self.robust and not 0 < eta - delta < eta + delta < 1:
            raise ValueError("Robust projection thresholds must lie within (0, 1)")
        self.thresholds = {"eroded": eta + delta, "intermediate": eta, "dilated": eta - delta} if self.robust \
            else {"intermediate": eta}
        self.beta = s["beta_schedule"][0] if s["beta_schedule"] else s["projection_beta"]
        self.filter = DensityFilter(self.shape, spacing, s["filter_radius_mm"], self.allowed, self.preserve)
        corridor = s["corridor"]
        if corridor is not None:  # blended with a wider cone where the corridor weights are positive
            self.filter.set_corridor(corridor["radius_mm"], corridor["weights"])
        flags = (self.preserve.astype(np.uint8) | (self.forbidden.astype(np.uint8) << 1)
                 | (self.free.astype(np.uint8) << 2))
        self.flags = torch.as_tensor(flags, device="cuda")
        coordinates = np.indices(self.shape).reshape(3, -1).T * np.asarray(spacing, float)
        weights = np.ones(self.n)
        weights[self.free] = volume_weights((coordinates + 0.5 * np.asarray(spacing, float)
                                             + np.asarray(origin, float))[self.free], s["prop_discs"])
        self.weights = torch.as_tensor(weights, device="cuda")
        self.ones = torch.ones(self.n, dtype=torch.float64, device="cuda")
        self._work = torch.empty(512, dtype=torch.float64, device="cuda")
        self._dot = torch.empty(1, dtype=torch.float64, device="cuda")

    def _stream(self):
        return _native.stream_handle(self.torch)

    def _design(self, design):
        torch = self.torch
        if isinstance(design, torch.Tensor):
            design = design.to(device="cuda", dtype=torch.float64).reshape(-1).contiguous()
            bad = not bool(torch.isfinite(design).all()) or bool((design < 0).any()) or bool((design > 1).any())
        else:
            host = np.asarray(design, dtype=float).ravel()
            bad = not np.all(np.isfinite(host)) or np.any(host < 0) or np.any(host > 1)
            design = torch.as_tensor(host, device="cuda")
        if design.numel() != self.n or bad:
            raise ValueError("Design densities must be finite and within [0, 1]")
        return design

    def filtered(self, design):
        return self.filter.filtered(self._design(design))

    def _project(self, filtered, etas):
        torch = self.torch
        count = len(etas)
        physical = torch.empty((count, self.n), dtype=torch.float64, device="cuda")
        derivative = torch.empty_like(physical)
        eta = (c_double * count)(*etas)
        _native.call("dfmi_project", _native.ptr(filtered.contiguous()), _native.ptr(self.flags), self.n,
                     float(self.beta), eta, count, int(self.robust), _native.ptr(physical), _native.ptr(derivative),
                     self._stream())
        return physical, derivative

    def project(self, filtered, eta):
        """(physical, derivative) at one threshold (topology_optimization.py:1152-1170)."""
        physical, derivative = self._project(filtered, [eta])
        return physical[0], derivative[0]

    def physical(self, design):
        return self.project(self.filtered(design), self.settings["projection_eta"])

    def fields(self, design):
        """({name: (physical, derivative)}, filtered) for every threshold in one projection pass."""
        filtered = self.filtered(design)
        names = list(self.thresholds)
        physical, derivative = self._project(filtered, [self.thresholds[n] for n in names])
        return {name: (physical[i], derivative[i]) for i, name in enumerate(names)}, filtered

    def _dot_with(self, a, field):
        _native.call("dfmi_gram", _native.ptr(a), _native.ptr(field.contiguous()), self.n, 1, 1,
                     _native.ptr(self._dot), _native.ptr(self._work), 512, self._stream())
        return float(self._dot.item())

    def weighted_sum(self, field):
        """sum_e w_e field_e with the volume weights."""
        return self._dot_with(self.weights, field)

    def total(self, field):
        """sum_e field_e."""
        return self._dot_with(self.ones, field)

    def volume(self, design):
        name = "dilated" if self.robust else "intermediate"
        return self.weighted_sum(self.project(self.filtered(design), self.thresholds[name])[0])

    def pullback(self, sensitivity, projection_derivative):
        """d/d design of a sensitivity w.r.t. a physical field (topology_optimization.py:1171-1178)."""
        return self.filter.pullback(sensitivity, projection_derivative)

    def initial(self, target):
        """Uniform start design: preserved cells 1, every free cell the largest common density (at least
        ``minimum_design_density``) whose volume stays within ``target``, found by bisection down to adjacent
        doubles (the start of topology_optimization.py:1179-1195)."""
        base = np.where(self.preserve, 1.0, 0.0)

        def design(value):
            out = base.copy()
            out[self.free] = value
            return out

        low, high = float(self.settings["minimum_design_density"]), 1.0
        if self.volume(design(low)) > target + 1e-8:
            raise ValueError(f"volume target {target} lies below the volume of the preserved cells with every "
                             f"free cell at the minimum density {low}")
        while True:
            middle = 0.5 * (low + high)
            if not low < middle < high:
                break
            if self.volume(design(middle)) > target:
                high = middle
            else:
                low = middle
        return design(low)

    def close(self):
        self.filter.close()


def evaluate_pullbacks(density_map, fields, stiffness_rows, geometric_rows, mass_per_cell, mass_scale,
                       stiffness="eroded", mass="intermediate"):
    """Gradient assembly of TopologyProblem.evaluate (deep_frame/topology_problem.py:366-379).

    fields: from ``DensityMap.fields``; stiffness_rows / geometric_rows: per-element gradients (device
    tensors) of the physics rows w.r.t. the stiffness field and of the geometric rows w.r.t. the mass field;
    mass_per_cell: grams of one full cell, mass_scale: grams of the volume budget.  Returns (objective,
    objective gradient, constraint gradients (rows, n)) with every gradient pulled back to the design."""
    torch = density_map.torch
    eroded, intermediate = fields[stiffness], fields[mass]
    objective = mass_per_cell * density_map.total(intermediate[0]) / mass_scale
    grads = [density_map.pullback(g, eroded[1]) for g in stiffness_rows] + \
            [density_map.pullback(g, intermediate[1]) for g in geometric_rows]
    ones = torch.full((density_map.n,), mass_per_cell / mass_scale, dtype=torch.float64, device="cuda")
    objective_gradient = density_map.pullback(ones, intermediate[1])
    stacked = torch.stack(grads) if grads else torch.empty((0, density_map.n), dtype=torch.float64, device="cuda")
    return objective, objective_gradient, stacked
