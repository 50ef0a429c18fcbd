"""Constraint bookkeeping around the device measures: the augmented-Lagrangian penalty and the KS minimum.

* ``AugmentedLagrangian`` plays the role of ``deep_frame/topology_optimization.py:839-866`` (used by
  ``ModalConstraint`` and ``StiffnessConstraint``): the Powell-Hestenes-Rockafellar penalty of inequality
  constraints g <= 0, psi = sum_i [rho_i / 2 max(0, g_i + mu_i / rho_i)^2 - mu_i^2 / (2 rho_i)], its gradient
  rho_i max(0, g_i + mu_i / rho_i) dg_i, the first-order multiplier update mu <- max(0, mu + rho g) every
  ``multiplier_interval`` calls, and penalty growth rho <- min(growth rho, penalty_max) for constraints whose
  infeasibility |max(g, -mu / rho)| did not fall below ``penalty_progress`` times its previous value.
* ``ks_minimum`` is the Kreisselmeier-Steinhauser lower bound of ``ModalConstraint.aggregate`` (:949-967):
  KS = r_min - log(sum_i exp(-k (r_i - r_min))) / k and its weights dKS/dr_i.
"""
import numpy as np


class AugmentedLagrangian:
    """PHR augmented Lagrangian of one constraint (``count`` None: scalars) or of ``count`` constraints."""

    def __init__(self, settings, count=None):
        self.settings = settings
        self.count = count
        self.calls = 0
        self.progress = None
        shape = () if count is None else (int(count),)
        self.multiplier = np.zeros(shape)
        self.penalty = np.full(shape, float(settings["penalty"]))
        if count is None:
            self.multiplier, self.penalty = float(self.multiplier), float(self.penalty)

    def terms(self, violation, slope):
        """(psi, d psi, weights rho max(0, g + mu / rho)); slope is dg (scalar case) or the rows dg_i."""
        g = np.asarray(violation, dtype=float)
        shifted = np.maximum(g + np.asarray(self.multiplier) / np.asarray(self.penalty), 0.0)
        weights = np.asarray(self.penalty) * shifted
        value = float(np.sum(0.5 * np.asarray(self.penalty) * shifted ** 2
                             - np.asarray(self.multiplier) ** 2 / (2.0 * np.asarray(self.penalty))))
        gradient = weights * slope if self.count is None else np.tensordot(weights, np.asarray(slope), axes=1)
        return value, gradient, weights

    def update(self, violation):
        g = np.asarray(violation, dtype=float)
        mu, rho = np.asarray(self.multiplier, dtype=float), np.asarray(self.penalty, dtype=float)
        infeasibility = np.abs(np.maximum(g, -mu / rho))
        mu = np.maximum(mu + rho * g, 0.0)
        growth = float(self.settings.get("penalty_growth", 1.0))
        if growth > 1.0 and self.progress is not None:
            stalled = infeasibility > self.settings["penalty_progress"] * self.progress
            rho = np.where(stalled, np.minimum(growth * rho, self.settings["penalty_max"]), rho)
        self.progress = infeasibility
        if self.count is None:
            self.multiplier, self.penalty = float(mu), float(rho)
        else:
            self.multiplier, self.penalty = mu, rho

    def augment(self, violation, slope):
        """(psi, d psi, the multiplier psi was evaluated with); updates every ``multiplier_interval`` calls."""
        used = self.multiplier if self.count is None else np.array(self.multiplier, copy=True)
        value, gradient, _ = self.terms(violation, slope)
        self.calls += 1
        if self.calls % int(self.settings["multiplier_interval"]) == 0:
            self.update(violation)
        return value, gradient, used


def ks_minimum(values, k):
    """(KS lower bound of the values, its weights dKS / dvalue_i) with aggregation parameter k."""
    r = np.asarray(values, dtype=float)
    low = r.min()
    shifted = np.exp(-k * (r - low))
    total = shifted.sum()
    return float(low - np.log(total) / k), shifted / total
