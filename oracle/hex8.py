"""Hex8 element, C-order grid numbering and CSR assembly (FP64 oracle).

Restates ``deep_frame/topology_optimization.py``:
  * corner order ``_CORNERS`` (:246-249),
  * isotropic / orthotropic constitutive matrices (:251-267),
  * ``hexahedron_matrices`` 2x2x2 Gauss stiffness and consistent mass (:269-299),
  * ``regular_grid`` C-order node / element / DOF numbering (:301-319),
  * ``select_nodes`` box selector with 1e-7 slack (:321-331),
  * ``HexElasticity.matrix`` COO->CSR assembly and symmetrisation (:638-650).
Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
from scipy.sparse import coo_matrix

# corner k of an element sits at node offset CORNERS[k] (topology_optimization.py:246-249)
CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)


def isotropic_d(nu):
    """Unit-modulus isotropic D in Voigt order xx,yy,zz,xy,yz,xz (:251-258)."""
    if not (np.isfinite(nu) and -1.0 < nu < 0.5):
        raise ValueError("Poisson ratio must be between -1 and 0.5")
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu2 = (1 - 2 * nu) / ((1 + nu) * (1 - 2 * nu))
    d = np.zeros((6, 6))
    d[:3, :3] = lam
    d[[0, 1, 2], [0, 1, 2]] = lam + mu2
    d[[3, 4, 5], [3, 4, 5]] = mu2 / 2
    return d


def orthotropic_d(c):
    """Inverse of the engineering compliance (:260-267)."""
    exy, ez = c["e_xy_mpa"], c["e_z_mpa"]
    nxy, nxz = c["nu_xy"], c["nu_xz"]
    s = np.zeros((6, 6))
    s[0, 0] = s[1, 1] = 1 / exy
    s[2, 2] = 1 / ez
    s[0, 1] = s[1, 0] = -nxy / exy
    s[0, 2] = s[2, 0] = s[1, 2] = s[2, 1] = -nxz / exy
    s[3, 3] = 1 / c["g_xy_mpa"]
    s[4, 4] = s[5, 5] = 1 / c["g_z_mpa"]
    if np.any(np.linalg.eigvalsh(s) <= 0):
        raise ValueError("Orthotropic engineering constants are not positive definite")
    return np.linalg.inv(s)


def hex8(spacing, d):
    """Ke (24x24), Me (24x24) and the 8 strain matrices B_g (6x24).

    Trilinear shape functions N_k = prod_a (1 + s_ka xi_a)/2 on [-1,1]^3, Gauss
    points xi = +-1/sqrt(3), weight 1, Jacobian h/2 per axis (:269-299).
    DOF order 3*corner + component.
    """
    h = np.asarray(spacing, dtype=float)
    if h.shape != (3,) or np.any(h <= 0) or not np.all(np.isfinite(h)):
        raise ValueError("Hexahedron spacing must contain three finite positive lengths")
    s = 2.0 * CORNERS - 1.0                     # (8,3) corner signs
    det = np.prod(h) / 8.0
    ke = np.zeros((24, 24))
    me = np.zeros((24, 24))
    bs = []
    g = 1.0 / np.sqrt(3.0)
    for gp in CORNERS:                           # Gauss point signs in the same order
        xi = (2.0 * gp - 1.0) * g
        f = (1.0 + s * xi) / 2.0                 # (8,3) 1D factors
        n = f.prod(axis=1)
        dn = np.empty((8, 3))
        for a in range(3):
            others = [b for b in range(3) if b != a]
            dn[:, a] = (s[:, a] / 2.0) * f[:, others[0]] * f[:, others[1]] * (2.0 / h[a])
        b = np.zeros((6, 24))
        for k in range(8):
            gx, gy, gz = dn[k]
            c = 3 * k
            b[0, c] = gx
            b[1, c + 1] = gy
            b[2, c + 2] = gz
            b[3, c], b[3, c + 1] = gy, gx
            b[4, c + 1], b[4, c + 2] = gz, gy
            b[5, c], b[5, c + 2] = gz, gx
        ke += b.T @ d @ b * det
        me += np.kron(np.outer(n, n), np.eye(3)) * det
        bs.append(b)
    return ke, me, np.array(bs)


def grid(shape, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """Node points, element connectivity (Ne,8) and DOFs (Ne,24), C order (:301-319)."""
    shape = np.asarray(shape, dtype=np.int64)
    if shape.shape != (3,) or np.any(shape < 1):
        raise ValueError("Grid shape must contain three positive integers")
    nodes = shape + 1
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in nodes], indexing="ij"), -1).reshape(-1, 3)
    points = idx * np.asarray(spacing, float) + np.asarray(origin, float)
    e = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).reshape(-1, 3)
    corner = e[:, None, :] + CORNERS[None]
    conn = (corner[..., 0] * nodes[1] + corner[..., 1]) * nodes[2] + corner[..., 2]
    dofs = (3 * conn[:, :, None] + np.arange(3)).reshape(-1, 24)
    return points, conn, dofs


def select_nodes(points, lo, hi):
    """Nodes inside the closed box [lo, hi] with 1e-7 slack (:321-331)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    sel = np.flatnonzero(np.all(points >= lo - 1e-7, axis=1) & np.all(points <= hi + 1e-7, axis=1))
    if not len(sel):
        raise ValueError("Empty topology node selector")
    return sel


def simp(density, young=1.0, p=3.0, eps=1e-6):
    """E = E0 (eps + (1-eps) rho^p), dE/drho (topology_optimization.py:657-658)."""
    rho = np.asarray(density, float)
    return young * (eps + (1 - eps) * rho ** p), young * (1 - eps) * p * rho ** (p - 1)


def assemble(dofs, ke, moduli, ndof):
    """Symmetrised CSR K = sum_e E_e P_e^T Ke P_e over elements with E_e != 0 (:638-650)."""
    moduli = np.asarray(moduli, float).ravel()
    live = moduli != 0
    el = dofs[live]
    rows = np.repeat(el, 24, axis=1).ravel()
    cols = np.tile(el, (1, 24)).ravel()
    vals = (moduli[live, None] * ke.ravel()[None, :]).ravel()
    k = coo_matrix((vals, (rows, cols)), shape=(ndof, ndof)).tocsr()
    return ((k + k.T) * 0.5).tocsr()
