"""Cone density filter, pullback and Heaviside projection (FP64 oracle).

Restates ``deep_frame/topology_optimization.py``:
  * cone weights max(r - |o h|, 0) over |o_a| <= floor(r / h_a) (:1082-1092, :1094-1105),
  * ``DensityMap.filtered`` x~ = (W (x a)) a / sums, sums = (W a) a, 1 on
    forbidden cells (:1123, :1130, :1135-1142),
  * ``DensityMap.project`` single and robust tanh projections with the preserve / forbidden / free
    masks (:1152-1170) and ``fields`` at the eroded / intermediate / dilated thresholds (:1114-1116, :1145-1147),
  * ``DensityMap.pullback`` (:1171-1178).
Implemented as an explicit shifted-sum over the offset box (not a convolution
library call).  Test infrastructure only.
"""
import numpy as np


def cone_weights(spacing, radius):
    pad = [int(np.floor(radius / h)) for h in spacing]
    offs = np.stack(np.meshgrid(*[np.arange(-k, k + 1) for k in pad], indexing="ij"), -1).reshape(-1, 3)
    dist = np.sqrt(((offs * np.asarray(spacing, float)) ** 2).sum(1))
    w = np.maximum(radius - dist, 0.0)
    keep = w > 0
    return offs[keep], w[keep]


def _neighbour_sum(field, offs, w):
    shape = field.shape
    out = np.zeros(shape)
    for (dx, dy, dz), wt in zip(offs, w):
        src = [slice(max(0, d), n + min(0, d)) for d, n in zip((dx, dy, dz), shape)]
        dst = [slice(max(0, -d), n + min(0, -d)) for d, n in zip((dx, dy, dz), shape)]
        out[tuple(dst)] += wt * field[tuple(src)]
    return out


class Cone:
    """With ``corridor`` = (radius, weights) the filtered field is blended with a wider cone,
    (1 - b) x~ + b x~_wide with b = clip(weights, 0, 1) on allowed cells and the wide sums 1 where b = 0
    (DensityMap corridor, topology_optimization.py:1125-1129, :1142, :1176)."""

    def __init__(self, allowed, spacing, radius, preserve=None, corridor=None):
        self.allowed = np.asarray(allowed, bool)
        self.shape = self.allowed.shape
        self.preserve = np.zeros(self.shape, bool) if preserve is None else np.asarray(preserve, bool)
        self.free = self.allowed & ~self.preserve
        self.offs, self.w = cone_weights(spacing, radius)
        a = self.allowed.astype(float)
        self.sums = _neighbour_sum(a, self.offs, self.w) * a
        self.sums[~self.allowed] = 1.0
        self.blend = None
        if corridor is not None:
            blend = np.clip(np.asarray(corridor[1], float).reshape(self.shape), 0, 1) * a
            if np.any(blend > 0):
                self.blend = blend
                self.offs2, self.w2 = cone_weights(spacing, corridor[0])
                self.sums2 = _neighbour_sum(a, self.offs2, self.w2) * a
                self.sums2[blend <= 0] = 1.0

    def apply(self, x, wide=False):
        a = self.allowed.astype(float)
        offs, w = (self.offs2, self.w2) if wide else (self.offs, self.w)
        return _neighbour_sum(np.asarray(x, float).reshape(self.shape) * a, offs, w) * a

    def filtered(self, design):
        out = self.apply(design) / self.sums
        if self.blend is not None:
            out = (1 - self.blend) * out + self.blend * self.apply(design, True) / self.sums2
        return out.ravel()

    def pullback(self, sensitivity, dproj=None):
        s = np.asarray(sensitivity, float).reshape(self.shape)
        if dproj is not None:
            s = s * np.asarray(dproj, float).reshape(self.shape)
        if self.blend is None:
            out = self.apply(s / self.sums)
        else:
            out = self.apply((1 - self.blend) * s / self.sums) + self.apply(self.blend * s / self.sums2, True)
        out[~self.free] = 0.0
        return out.ravel()


def project(filtered, beta, eta=0.5):
    """(physical, d physical / d filtered) for the single tanh projection."""
    x = np.asarray(filtered, float)
    if not beta:
        return x.copy(), np.ones_like(x)
    den = np.tanh(beta * eta) + np.tanh(beta * (1 - eta))
    t = np.tanh(beta * (x - eta))
    return np.clip((np.tanh(beta * eta) + t) / den, 0, 1), beta * (1 - t ** 2) / den


def project_masked(filtered, beta, eta, robust, preserve=None, forbidden=None, free=None):
    """(physical, derivative) of DensityMap.project (topology_optimization.py:1152-1170): the robust form
    writes the derivative as beta 4 e^(-2|s|) / (1 + e^(-2|s|))^2 / den, the single form as
    beta (1 - tanh^2 s) / den; preserve -> 1, forbidden -> 0, derivative 0 off the free cells, clip."""
    x = np.asarray(filtered, float)
    n = x.size
    if beta and robust:
        den = np.tanh(beta * eta) + np.tanh(beta * (1 - eta))
        shifted = beta * (x - eta)
        decay = np.exp(-2 * np.abs(shifted))
        physical = (np.tanh(beta * eta) + np.tanh(shifted)) / den
        derivative = beta * 4 * decay / (1 + decay) ** 2 / den
    elif beta:
        den = np.tanh(beta * eta) + np.tanh(beta * (1 - eta))
        t = np.tanh(beta * (x - eta))
        physical, derivative = (np.tanh(beta * eta) + t) / den, beta * (1 - t ** 2) / den
    else:
        physical, derivative = x.copy(), np.ones(n)
    if preserve is not None:
        physical[np.asarray(preserve, bool)] = 1
    if forbidden is not None:
        physical[np.asarray(forbidden, bool)] = 0
    if free is not None:
        derivative[~np.asarray(free, bool)] = 0
    return np.clip(physical, 0, 1), derivative


def thresholds(eta, delta, robust):
    """Projection thresholds of DensityMap (topology_optimization.py:1114-1116)."""
    return {"eroded": eta + delta, "intermediate": eta, "dilated": eta - delta} if robust else {"intermediate": eta}
