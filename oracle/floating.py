"""Floating (inertia-relief) fixtures: rigid-mode kernel, support reactions and rigid shift (FP64 oracle).

Restates ``deep_frame/topology_multigrid.py``:
  * the six rigid modes about the centre shape * spacing / 2 (``_rigid``, :29-38; centre :64),
  * the free kernel: null space of the rigid modes' Gram matrix over the Dirichlet DOFs of active nodes,
    eigenvalues <= 1e-10 max(lambda_max, 1) (:151-159),
  * ``solve`` with a support (:340-349, :368-370): Dirichlet DOFs = fixed minus support; the support DOFs
    take the least-squares forces that cancel the forces' components along the masked kernel modes; the
    singular reduced system is solved for the displacement orthogonal to the kernel (the limit of the
    projected PCG from zero); the solution is shifted by the kernel motion that best fits zero on the
    support.
The kernel-orthogonal solve uses the saddle-point system [K_ff B; B^T 0].
Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
from scipy.sparse import bmat, csc_matrix
from scipy.sparse.linalg import spsolve

from oracle.hex8 import assemble
from oracle.system import free_dofs


def rigid_modes(points, center):
    """(ndof_of_points, 6): translations x, y, z, rotations (0, -z, y), (z, 0, -x), (-y, x, 0)."""
    d = np.asarray(points, float) - np.asarray(center, float)
    modes = np.zeros((len(d), 3, 6))
    for axis in range(3):
        modes[:, axis, axis] = 1.0
    modes[:, 1, 3], modes[:, 2, 3] = -d[:, 2], d[:, 1]
    modes[:, 0, 4], modes[:, 2, 4] = d[:, 2], -d[:, 0]
    modes[:, 0, 5], modes[:, 1, 5] = -d[:, 1], d[:, 0]
    return modes.reshape(-1, 6)


def kernel_coefficients(points, conn, moduli, dirichlet, center):
    """(6, nk) combinations of the rigid modes that vanish on the Dirichlet DOFs of active nodes."""
    active_nodes = np.unique(conn[np.asarray(moduli).ravel() > 0])
    dirichlet = np.asarray(dirichlet, dtype=np.int64)
    constrained = dirichlet[np.isin(dirichlet // 3, active_nodes)]
    if not constrained.size:
        return np.eye(6)
    modes = rigid_modes(points, center)[constrained]
    values, vectors = np.linalg.eigh(modes.T @ modes)
    return vectors[:, values <= 1e-10 * max(values.max(), 1.0)]


def solve(dofs, conn, points, ke, moduli, fixed, support, forces, center):
    """Displacements (ndof, k) of the floating solve, the kernel size and the adjusted forces."""
    nnode = len(points)
    ndof = 3 * nnode
    forces = np.array(forces, float).reshape(ndof, -1)
    support = np.asarray(support, dtype=np.int64)
    dirichlet = np.setdiff1d(fixed, support)
    free = free_dofs(conn, moduli, dirichlet, nnode)
    mask = np.zeros(ndof)
    mask[free] = 1.0
    coefficients = kernel_coefficients(points, conn, moduli, dirichlet, center)
    rigid = (rigid_modes(points, center) @ coefficients) * mask[:, None]
    forces[support] += np.linalg.lstsq(rigid[support].T, -(rigid.T @ forces), rcond=None)[0]
    k = assemble(dofs, ke, moduli, ndof)
    kff = k[free][:, free]
    b = csc_matrix(rigid[free])
    nk = rigid.shape[1]
    system = bmat([[kff, b], [b.T, None]], format="csc")
    rhs = np.vstack([forces[free], np.zeros((nk, forces.shape[1]))])
    sol = spsolve(system, rhs).reshape(len(free) + nk, -1)
    u = np.zeros_like(forces)
    u[free] = sol[:len(free)]
    shift = np.linalg.lstsq(rigid[support], u[support], rcond=None)[0]
    return u - rigid @ shift, nk, forces
