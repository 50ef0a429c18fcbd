"""One SIMP design iteration on the GPU: density filter, SIMP interpolation, the
multigrid solve, compliance and sensitivity, pullback, and the OC or MMA update.

Mirrors the reference's loop pieces: ``DensityMap`` (deep_frame/topology_optimization.py:1110-1178,
identity projection), ``HexElasticity.solve`` (:651-734), ``_oc_update`` (:1197-1233) and
``MMAOptimizer.step`` (deep_frame/topology_optimizers.py:55-67).  All arithmetic runs in
libdfmi; torch tensors only hold device memory.
"""
import ctypes
from ctypes import byref, c_double, c_int32, c_void_p

import numpy as np

from deep_frame_amd import _native, element
from deep_frame_amd.multigrid import GeometricMultigrid

MMA = {"move": 0.1, "scale": 100.0}


def _torch():
    import torch
    return torch


class DensityFilter:
    """Cone filter x~ = (W (x a)) a / sums and its pullback (topology_optimization.py:1082-1178)."""

    def __init__(self, shape, spacing, radius, allowed=None, preserve=None):
        self.torch = _torch()
        self.shape = tuple(int(n) for n in shape)
        n = int(np.prod(self.shape))
        allowed = np.ones(n, np.uint8) if allowed is None else np.asarray(allowed, bool).ravel().astype(np.uint8)
        preserve = np.zeros(n, np.uint8) if preserve is None else np.asarray(preserve, bool).ravel().astype(np.uint8)
        self.free = (allowed > 0) & ~(preserve > 0)
        handle = c_void_p()
        _native.call("dfmi_filter_create", (c_int32 * 3)(*self.shape), (c_double * 3)(*spacing), float(radius),
                     allowed.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                     preserve.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), byref(handle))
        self.handle = handle
        self.n = n

    def _stream(self):
        _native.call("dfmi_filter_set_stream", self.handle, _native.stream_handle(self.torch))

    def _dev(self, x):
        return x.to(device="cuda", dtype=self.torch.float64).contiguous()

    def filtered(self, x):
        x = self._dev(x)
        out = self.torch.empty_like(x)
        self._stream()
        _native.call("dfmi_filter_forward", self.handle, _native.ptr(x), _native.ptr(out))
        return out

    def pullback(self, sensitivity, dproj=None):
        s = self._dev(sensitivity)
        d = self.torch.ones_like(s) if dproj is None else self._dev(dproj)
        out = self.torch.empty_like(s)
        self._stream()
        _native.call("dfmi_filter_pullback", self.handle, _native.ptr(s), _native.ptr(d), _native.ptr(out))
        return out

    def set_corridor(self, radius, weights):
        """Blend a cone of radius ``radius`` into the filter with per-cell weights clip(weights, 0, 1) on
        allowed cells (DensityMap's corridor, topology_optimization.py:1125-1129, :1142, :1176)."""
        w = np.ascontiguousarray(np.asarray(weights, dtype=float).ravel())
        if w.size != self.n:
            raise ValueError("corridor weights need one value per cell")
        _native.call("dfmi_filter_set_corridor", self.handle, float(radius),
                     w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def volume(self, x):
        x = self._dev(x)
        v = c_double()
        self._stream()
        _native.call("dfmi_filter_volume", self.handle, _native.ptr(x), byref(v))
        return v.value

    def oc_update(self, x, dc, dv, target, xmin=1e-3, move=0.2):
        x, dc, dv = self._dev(x), self._dev(dc), self._dev(dv)
        out = self.torch.empty_like(x)
        self._stream()
        _native.call("dfmi_oc_update", self.handle, _native.ptr(x), _native.ptr(dc), _native.ptr(dv), float(target),
                     float(xmin), float(move), _native.ptr(out))
        return out

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            _native.call("dfmi_filter_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simp(rho, young=1.0, penal=3.0, eps=1e-9, active=None):
    """(E, dE/drho) device tensors: E = E0 (eps + (1 - eps) rho^p) (topology_optimization.py:657-661)."""
    torch = _torch()
    rho = rho.to(device="cuda", dtype=torch.float64).contiguous()
    e, de = torch.empty_like(rho), torch.empty_like(rho)
    act = None if active is None else torch.as_tensor(np.asarray(active, np.uint8), device="cuda")
    _native.call("dfmi_simp", _native.ptr(rho), None if act is None else _native.ptr(act), rho.numel(), float(young),
                 float(penal), float(eps), _native.ptr(e), _native.ptr(de), _native.stream_handle(torch))
    return e, de


class MMAState:
    """Design-variable history and asymptotes of MMAOptimizer.fresh (topology_optimizers.py:87-89)."""

    def __init__(self, x):
        torch = _torch()
        x = x.to(device="cuda", dtype=torch.float64).contiguous()
        self.iteration = 1
        self.old1, self.old2 = x.clone(), x.clone()
        self.low, self.upp = torch.zeros_like(x), torch.ones_like(x)

    def step(self, x, f0, df0, g, dg, move=MMA["move"], scale=MMA["scale"]):
        """One MMA update for the constraints g_j <= 0 (g a scalar or m values, dg (n,) or (m, n) device
        tensors), as MMAOptimizer.step passes every constraint row (topology_optimizers.py:55-67, m = len(g));
        returns the new design (bounds [0, 1])."""
        torch = _torch()
        x = x.to(device="cuda", dtype=torch.float64).contiguous().clone()
        df0 = df0.to(device="cuda", dtype=torch.float64).contiguous()
        g = np.atleast_1d(np.asarray(g, dtype=float)).ravel()
        dg = dg.to(device="cuda", dtype=torch.float64).reshape(len(g), -1).contiguous()
        if dg.shape[1] != x.numel():
            raise ValueError("dg needs one row of n values per constraint")
        _native.call("dfmi_mma_step", x.numel(), len(g), self.iteration, _native.ptr(x), _native.ptr(self.old1),
                     _native.ptr(self.old2), _native.ptr(self.low), _native.ptr(self.upp), float(f0), _native.ptr(df0),
                     np.ascontiguousarray(g).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _native.ptr(dg),
                     float(move), float(scale), _native.stream_handle(torch))
        self.iteration += 1
        return x


class SimpProblem:
    """Minimum compliance under a volume bound on a regular hex8 grid (one load group)."""

    def __init__(self, shape, fixed, forces, volume_fraction=0.3, radius=1.5, young=1.0, emin_ratio=1e-9, penal=3.0,
                 nu=0.3, settings=None):
        torch = self.torch = _torch()
        self.shape = tuple(int(n) for n in shape)
        self.n = int(np.prod(self.shape))
        self.ke = element.stiffness((1.0, 1.0, 1.0), element.isotropic(nu))
        self.mg = GeometricMultigrid(self.shape, (1.0, 1.0, 1.0), self.ke, settings)
        self.filter = DensityFilter(self.shape, (1.0, 1.0, 1.0), radius)
        self.fixed = np.asarray(fixed, np.int64)
        self.forces = torch.as_tensor(np.ascontiguousarray(np.asarray(forces, float).reshape(self.mg.ndof, -1).T),
                                      device="cuda")
        self.target = volume_fraction * self.n
        self.young, self.eps, self.penal = young, emin_ratio, penal
        self.ones = torch.ones(self.n, dtype=torch.float64, device="cuda")
        self.dv = self.filter.pullback(self.ones)

    def evaluate(self, x):
        """Physical density, compliance, d compliance / d x, volume and its gradient."""
        rho = self.filter.filtered(x)
        e, de = simp(rho, self.young, self.penal, self.eps)
        self.mg.update(e)
        u, reports = self.mg.solve_device(b"simp", self.fixed, self.forces)
        if not all(r[2] for r in reports):
            raise RuntimeError(f"Multigrid PCG did not converge: {reports}")
        comp, sens = self.mg.compliance_sensitivity_device(u, self.forces, de)
        dc = self.filter.pullback(sens)
        return {"physical": rho, "compliance": float(comp.sum()), "dc": dc, "volume": float(rho.sum()), "dv": self.dv,
                "reports": reports, "displacement": u}

    def mma_iteration(self, x, state):
        res = self.evaluate(x)
        g = res["volume"] / self.target - 1.0
        dg = self.dv / self.target
        return state.step(x, res["compliance"], res["dc"], g, dg), res

    def oc_iteration(self, x, move=0.2, xmin=1e-3):
        res = self.evaluate(x)
        return self.filter.oc_update(x, res["dc"], self.dv, self.target, xmin, move), res

    def close(self):
        self.filter.close()
        return self.mg.close()
