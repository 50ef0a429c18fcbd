"""Optimality-criteria update (FP64 oracle).

Restates ``_oc_update`` (``deep_frame/topology_optimization.py:1197-1233``):
x_new = clip(x sqrt(max(1e-30, -dc / dv) / lambda), max(xmin, x - m), min(1, x + m))
with lambda bracketed by doubling and bisected (<= 70 steps, relative gap 1e-8)
so that volume(x_new) <= target.  ``volume`` is a callable on the full design.
Test infrastructure only.
"""
import numpy as np


def oc_update(design, dc, dv, free, volume, target, xmin, move):
    free = np.asarray(free, bool)
    dvf = np.asarray(dv, float)[free]
    if np.any(dvf < 0) or not np.max(dvf) > 0:
        raise RuntimeError("Invalid OC sensitivities")
    dvf = np.maximum(dvf, 1e-12 * np.max(dvf))
    ratios = np.maximum(1e-30, -np.asarray(dc, float)[free] / dvf)
    lo = np.maximum(xmin, design[free] - move)
    hi = np.minimum(1.0, design[free] + move)

    def proposal(lmb):
        out = design.copy()
        out[free] = np.clip(design[free] * np.sqrt(ratios / max(lmb, 1e-100)), lo, hi)
        return out

    lowest = design.copy()
    lowest[free] = lo
    if volume(lowest) > target:
        return lowest
    upper = max(float(np.max(ratios)), 1e-12)
    for _ in range(100):
        if volume(proposal(upper)) <= target + 1e-9:
            break
        upper *= 2
    else:
        raise RuntimeError("OC could not bracket the volume multiplier")
    lower, cand = 0.0, proposal(upper)
    for _ in range(70):
        mid = 0.5 * (lower + upper)
        prop = proposal(mid)
        if volume(prop) > target:
            lower = mid
        else:
            upper, cand = mid, prop
        if (upper - lower) / max(upper + lower, 1e-100) < 1e-8:
            break
    return cand
