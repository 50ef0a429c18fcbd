"""Hex8 element matrices and regular-grid numbering (host set-up, FP64 NumPy).

Same definitions as deep_frame/topology_optimization.py: corner order (:246-249),
isotropic / orthotropic constitutive matrices (:251-267), 2x2x2 Gauss stiffness and the
consistent mass (:269-299) and C-order node/DOF numbering (:301-319).  These run once per problem
on the host; the per-iteration work is in libdfmi.
"""
import numpy as np

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])

# Voigt strain rows (xx, yy, zz, xy, yz, xz) as (row, displacement component, derivative axis)
_VOIGT = [(0, 0, 0), (1, 1, 1), (2, 2, 2), (3, 0, 1), (3, 1, 0), (4, 1, 2), (4, 2, 1), (5, 0, 2), (5, 2, 0)]


def isotropic(nu):
    if not (np.isfinite(nu) and -1 < nu < 0.5):
        raise ValueError("Poisson ratio must be between -1 and 0.5")
    e = np.eye(6)
    d = np.zeros((6, 6))
    d[:3, :3] = nu
    d[np.arange(3), np.arange(3)] = 1 - nu
    d[3:, 3:] = e[3:, 3:] * (1 - 2 * nu) / 2
    return d / ((1 + nu) * (1 - 2 * nu))


def orthotropic(c):
    ex, ez = c["e_xy_mpa"], c["e_z_mpa"]
    s = np.diag([1 / ex, 1 / ex, 1 / ez, 1 / c["g_xy_mpa"], 1 / c["g_z_mpa"], 1 / c["g_z_mpa"]])
    s[0, 1] = s[1, 0] = -c["nu_xy"] / ex
    s[0, 2] = s[2, 0] = s[1, 2] = s[2, 1] = -c["nu_xz"] / ex
    if np.any(np.linalg.eigvalsh(s) <= 0):
        raise ValueError("Orthotropic engineering constants are not positive definite")
    return np.linalg.inv(s)


def stiffness(spacing, d):
    """24 x 24 unit-modulus stiffness, DOF = 3 * corner + component."""
    h = np.asarray(spacing, dtype=float)
    sign = 2.0 * CORNERS - 1.0                                   # (8, 3)
    gauss = (2.0 * CORNERS - 1.0) / np.sqrt(3.0)                 # (8 points, 3)
    f = 0.5 * (1.0 + sign[None, :, :] * gauss[:, None, :])       # (g, corner, axis)
    grad = np.empty((8, 8, 3))
    for a in range(3):
        other = np.prod(np.delete(f, a, axis=2), axis=2)
        grad[:, :, a] = sign[None, :, a] * other / h[a]          # dN/dx_a
    b = np.zeros((8, 6, 8, 3))
    for row, comp, axis in _VOIGT:
        b[:, row, :, comp] = grad[:, :, axis]
    b = b.reshape(8, 6, 24)
    return np.einsum("gsi,st,gtj->ij", b, d, b) * np.prod(h) / 8.0


def mass(spacing):
    """24 x 24 consistent mass of a unit-density brick, DOF = 3 * corner + component: the exact integral of
    N_i N_j, which for trilinear shape functions is V / 216 * prod over the axes of (2 if corners i and j share
    that coordinate else 1), times the 3 x 3 identity (the matrix hexahedron_matrices forms by 2x2x2 Gauss
    quadrature, topology_optimization.py:269-299)."""
    volume = float(np.prod(np.asarray(spacing, dtype=float)))
    same = (CORNERS[:, None, :] == CORNERS[None, :, :]).astype(float) + 1.0    # (8, 8, 3)
    m8 = volume / 216.0 * np.prod(same, axis=2)
    return np.kron(m8, np.eye(3))


def node_index(shape):
    nodes = np.asarray(shape) + 1
    return lambda i, j, l: (np.asarray(i) * nodes[1] + np.asarray(j)) * nodes[2] + np.asarray(l)


def node_points(shape, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    axes = [np.arange(n + 1) * h + o for n, h, o in zip(shape, spacing, origin)]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
