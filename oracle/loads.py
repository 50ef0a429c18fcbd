"""Unit-wrench nodal loads (FP64 oracle).

Restates ``InterfaceCovariance.wrenches`` (``deep_frame/topology_problem.py:624-638``):
for patch nodes with arms r_i = x_i - x_ref, B = [B_i] with B_i = [I; skew(r_i)]
(6 x 3), the minimum-norm nodal force producing unit wrench e_j is
f = B^T (B B^T)^-1 e_j.  Test infrastructure only.
"""
import numpy as np


def unit_wrenches(points, nodes, reference):
    """(6, 3*len(nodes)) nodal force blocks: row j is the force for wrench e_j."""
    arms = np.asarray(points, float)[np.asarray(nodes)] - np.asarray(reference, float)
    if np.linalg.matrix_rank(arms - arms[0], tol=1e-8) < 2:
        raise ValueError("An interface wrench needs at least three non-collinear nodes")
    b = np.zeros((6, 3 * len(arms)))
    for i, (x, y, z) in enumerate(arms):
        b[:3, 3 * i:3 * i + 3] = np.eye(3)
        b[3:, 3 * i:3 * i + 3] = [[0.0, -z, y], [z, 0.0, -x], [-y, x, 0.0]]
    return np.linalg.solve(b @ b.T, b)          # = ((B^T (B B^T)^-1))^T


def wrench_columns(points, nodes, reference, ndof):
    """Six dense force columns (ndof, 6) for Fx, Fy, Fz, Mx, My, Mz about ``reference``."""
    blocks = unit_wrenches(points, nodes, reference)
    out = np.zeros((ndof, 6))
    dofs = (3 * np.asarray(nodes)[:, None] + np.arange(3)).ravel()
    out[dofs] = blocks.T
    return out
