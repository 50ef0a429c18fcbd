"""Interface load-covariance measures (FP64 oracle).

Restates ``deep_frame/topology_problem.py`` ``InterfaceCovariance``:
  * load directions from the covariance Sigma: eigen-decomposition, descending, rank by a 1e-10
    relative cut, D = V sqrt(Lambda) (:583-591),
  * flexibility of the load cases F = factor Forces^T U, symmetrised (:639-646); with the cases built as
    unit-wrench combinations (Forces = W D, U = U_w D) this is factor D^T (W^T U_w) D,
  * element energy of a combined field (:647-652): -factor E' (U v)_e^T Ke (U v)_e,
  * measure (:657-684): mean = tr F, worst = lambda_max(F), the KS aggregate
    worst + scale log(sum exp(s (lambda - worst) / scale)) / s with its softmax weights, and the gradients
    sum_k of the case compliance sensitivities (mean) and sum_i w_i energy(v_i) over weights above
    ks_cutoff (worst).
Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np

from oracle.system import element_energy


def directions(sigma):
    """(D (n, rank), rank) with D D^T = Sigma on its range (topology_problem.py:583-591)."""
    sigma = np.asarray(sigma, float)
    values, vectors = np.linalg.eigh((sigma + sigma.T) / 2)
    values, vectors = values[::-1], vectors[:, ::-1]
    if values[-1] < -1e-9 * values[0]:
        raise ValueError("Load covariance is not positive semidefinite")
    rank = int(np.sum(values > 1e-10 * values[0]))
    return vectors[:, :rank] * np.sqrt(values[:rank]), rank


def matrix(unit_flexibility, d, factor=1.0):
    """Case flexibility factor D^T F_w D, symmetrised (topology_problem.py:639-646)."""
    f = factor * d.T @ np.asarray(unit_flexibility, float) @ d
    return (f + f.T) / 2


def aggregate(flex, ks=50.0, worst_limit=None):
    """mean, worst, KS value, softmax weights and eigenvectors (descending) (topology_problem.py:657-670)."""
    values, vectors = np.linalg.eigh(flex)
    values, vectors = values[::-1], vectors[:, ::-1]
    worst = float(values[0])
    scale = worst_limit or worst
    weights = np.exp(ks * (values - worst) / scale)
    value = worst + scale * np.log(weights.sum()) / ks
    return float(np.trace(flex)), worst, float(value), weights / weights.sum(), vectors


def gradients(dofs, ke, u_unit, d, vectors, weights, dmoduli, factor=1.0, cutoff=1e-9):
    """(d mean / d rho, d worst-KS / d rho) per element from the unit-wrench solutions u_unit (ndof, n).

    mean: sum over cases k of -factor E' (U_w d_k)_e^T Ke (U_w d_k)_e (the case compliance sensitivities);
    worst: sum_i w_i energy(U_w D v_i) over weights above the cutoff (topology_problem.py:675-684)."""
    cases = u_unit @ d
    mean = -factor * dmoduli * element_energy(dofs, ke, cases).sum(axis=1)
    worst = np.zeros_like(mean)
    for i, w in enumerate(weights):
        if w > cutoff:
            worst += w * (-factor * dmoduli * element_energy(dofs, ke, cases @ vectors[:, i : i + 1])[:, 0])
    return mean, worst
