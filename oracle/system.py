"""Reduced linear solve, compliance and sensitivity (FP64 oracle).

Restates ``deep_frame/topology_optimization.py``:
  * active nodes / DOFs and free = active minus fixed (:365-366, :433),
  * SuperLU solve of K_ff with MMD_AT_PLUS_A ordering (:757-759),
  * the residual gate (:716-720),
  * compliance f.u and sensitivity -E'(rho) u_e^T Ke u_e (:721-734, :707),
and ``deep_frame/topology_multigrid.py`` ``solve_elasticity`` (:553-581).
Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
from scipy.sparse.linalg import splu

from oracle.hex8 import assemble


def free_dofs(conn, moduli, fixed, nnode):
    active_nodes = np.unique(conn[np.asarray(moduli).ravel() > 0])
    active = (3 * active_nodes[:, None] + np.arange(3)).ravel()
    return np.setdiff1d(active, np.asarray(fixed, dtype=np.int64))


def solve(dofs, conn, ke, moduli, fixed, forces, nnode):
    """Displacements (ndof, k) with u = 0 on fixed / inactive DOFs, and rel. residuals."""
    ndof = 3 * nnode
    forces = np.asarray(forces, float).reshape(ndof, -1)
    k = assemble(dofs, ke, moduli, ndof)
    free = free_dofs(conn, moduli, fixed, nnode)
    kff = k[free][:, free].tocsc()
    lu = splu(kff, permc_spec="MMD_AT_PLUS_A", options={"SymmetricMode": True})
    u = np.zeros_like(forces)
    u[free] = lu.solve(forces[free])
    res = np.linalg.norm(kff @ u[free] - forces[free], axis=0)
    nrm = np.linalg.norm(forces[free], axis=0)
    return u, res / np.where(nrm > 0, nrm, 1.0)


def element_energy(dofs, ke, u):
    """u_e^T Ke u_e per element and column -> (Ne, k) (:732, topology_multigrid.py:458-459)."""
    u = np.asarray(u, float).reshape(-1, 1) if np.ndim(u) == 1 else np.asarray(u, float)
    ue = u[dofs]                                  # (Ne, 24, k)
    return np.einsum("eik,ij,ejk->ek", ue, ke, ue, optimize=True)


def compliance_sensitivity(dofs, ke, u, forces, dmoduli, factor=1.0):
    """Per-column compliance f.u and the summed sensitivity -E' sum_k u_e^T Ke u_e."""
    forces = np.asarray(forces, float).reshape(u.shape)
    comp = factor * np.einsum("ik,ik->k", forces, u)
    energy = factor * element_energy(dofs, ke, u)
    return comp, -np.asarray(dmoduli, float)[:, None] * energy
