"""Augmented-Lagrangian penalty and KS minimum (deep_frame_amd/constraints.py), CPU only.

The augmented-Lagrangian numbers are the reference's own known answers
(tests/test_topology_problem.py:35-50 of the reference: penalty 10 scalar case, the two-constraint vector
case with penalty growth 2, progress 0.5, cap 4); the KS minimum is pinned by its closed form for equal
values and by central differences of its weights."""
import numpy as np
import pytest

from deep_frame_amd.constraints import AugmentedLagrangian, ks_minimum


def test_augmented_lagrangian_known_answers():
    scalar = AugmentedLagrangian({"penalty": 10.0, "multiplier_interval": 1})
    assert scalar.augment(0.3, 2.0) == pytest.approx((10 / 2 * 0.3 ** 2, 10 * 0.3 * 2.0, 0.0))
    assert scalar.multiplier == pytest.approx(3.0)
    settings = {"penalty": 1.0, "multiplier_interval": 1, "penalty_growth": 2.0, "penalty_progress": 0.5,
                "penalty_max": 4.0}
    vector = AugmentedLagrangian(settings, 2)
    value, gradient, weights = vector.terms([0.3, -0.5], np.eye(2))
    assert value == pytest.approx(0.045) and gradient.tolist() == pytest.approx([0.3, 0.0])
    assert weights.tolist() == pytest.approx([0.3, 0.0])
    vector.update([0.3, -0.5])
    assert vector.multiplier.tolist() == pytest.approx([0.3, 0.0]) and vector.penalty.tolist() == [1.0, 1.0]
    for _ in range(3):
        vector.update([0.2, -0.5])
    assert vector.penalty.tolist() == [4.0, 1.0]
    vector.progress = None
    vector.update([0.2, -0.5])
    assert vector.penalty.tolist() == [4.0, 1.0]


def test_augmented_lagrangian_interval_and_inactive_constraint():
    al = AugmentedLagrangian({"penalty": 2.0, "multiplier_interval": 3})
    for call in range(1, 4):
        value, gradient, used = al.augment(-1.0, 5.0)
        assert used == 0.0 and gradient == 0.0 and value == 0.0  # inactive: max(0, g + mu / rho) = 0
    assert al.multiplier == 0.0 and al.calls == 3
    al.multiplier = 1.0  # with a multiplier the inactive value is -mu^2 / (2 rho)
    assert al.terms(-1.0, 5.0)[0] == pytest.approx(-0.25)


def test_ks_minimum_closed_form_and_weights():
    value, weights = ks_minimum([2.0, 2.0, 2.0, 2.0], 40.0)
    assert value == pytest.approx(2.0 - np.log(4.0) / 40.0) and np.allclose(weights, 0.25)
    r = np.array([1.3, 1.1, 1.25, 2.0])
    value, weights = ks_minimum(r, 40.0)
    assert value <= r.min() and weights.sum() == pytest.approx(1.0)
    for i in range(len(r)):
        step = np.zeros_like(r)
        step[i] = 1e-6
        fd = (ks_minimum(r + step, 40.0)[0] - ks_minimum(r - step, 40.0)[0]) / 2e-6
        assert fd == pytest.approx(weights[i], rel=1e-6, abs=1e-12)
