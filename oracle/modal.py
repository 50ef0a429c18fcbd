"""Generalised eigenproblem K v = lambda M v of the clamped structure (FP64 oracle).

Restates ``deep_frame/topology_optimization.py`` ``elastic_frequencies`` (:796-816): stiffness with the
SIMP moduli, consistent mass sum_e m_e Me (m_e = density 1e-9 rho_e, :808-809) plus an optional lumped
diagonal (``ModalConstraint``, :895-913), both reduced to the free DOFs, and the lowest eigenpairs by
shift-invert Lanczos about 0 (``eigsh(..., sigma=0, which="LM")``, :813); frequencies sqrt(lambda) / 2 pi.
This is the quantity ``GeometricMultigrid.eigenpairs`` (deep_frame/topology_multigrid.py:481-540)
computes iteratively.
Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
from scipy.sparse import diags
from scipy.sparse.linalg import eigsh

from oracle.hex8 import assemble
from oracle.system import free_dofs


def eigenpairs(dofs, conn, ke, me, moduli, mass, fixed, nnode, count, lumped=None):
    """Lowest `count` eigenvalues (ascending) and M-orthonormal eigenvectors (ndof, count), zero off the free DOFs."""
    ndof = 3 * nnode
    k = assemble(dofs, ke, moduli, ndof)
    m = assemble(dofs, me, mass, ndof)
    if lumped is not None:
        m = (m + diags(np.asarray(lumped, float))).tocsr()
    free = free_dofs(conn, moduli, fixed, nnode)
    kff, mff = k[free][:, free].tocsc(), m[free][:, free].tocsc()
    values, vectors = eigsh(kff, k=count, M=mff, sigma=0.0, which="LM")
    order = np.argsort(values)
    out = np.zeros((ndof, count))
    out[free] = vectors[:, order]
    return values[order], out


def frequencies(values):
    return np.sqrt(np.asarray(values, float)) / (2 * np.pi)


def interpolation(density, cutoff):
    """Mass interpolation and slope (topology_optimization.py:904-907): rho^6 / c^5 below c, rho above."""
    density = np.asarray(density, float)
    low = density < cutoff
    return np.where(low, density ** 6 / cutoff ** 5, density), np.where(low, 6 * density ** 5 / cutoff ** 5, 1.0)


def aggregate(dofs, ke, me, modes, density, young, rho_g, target, ks, cutoff, penalization=3.0,
              min_stiffness_ratio=1e-6, active=None):
    """(violation, gradient, f1_hz, aggregate ratio) of ModalConstraint.aggregate
    (topology_optimization.py:949-967) for modes [(eigenvalue, M-normalised flat vector)]."""
    density = np.asarray(density, float)
    active = np.ones(len(density), bool) if active is None else np.asarray(active, bool)
    ratios = np.array([value / target for value, _ in modes])
    lowest = ratios.min()
    exponents = np.exp(-ks * (ratios - lowest))
    value = lowest - np.log(exponents.sum()) / ks
    weights = exponents / exponents.sum()
    stiffness_slope = young * (1 - min_stiffness_ratio) * penalization * density ** (penalization - 1)
    mass_slope = rho_g * 1e-9 * interpolation(density, cutoff)[1]
    sensitivity = np.zeros(len(density))
    for weight, (eig, vector) in zip(weights, modes):
        element = vector[dofs]
        strain = np.einsum("ei,ij,ej->e", element, ke, element)
        kinetic = np.einsum("ei,ij,ej->e", element, me, element)
        sensitivity += weight * (stiffness_slope * strain - eig * mass_slope * kinetic) / target
    sensitivity[~active] = 0
    return 1 - value, -sensitivity, float(np.sqrt(max(lowest * target, 0)) / (2 * np.pi)), float(value)
