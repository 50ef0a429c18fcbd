"""Synthetic SIMP workloads (inputs only): grids, densities, fixtures and loads.

* ``cantilever``: 48 x 24 x 24 unit hex8 cantilever, x = 0 face clamped, unit -z
  point load at the bottom far-edge midpoint, rho = 0.3 + U(-0.05, 0.05)
  (the fixture of deep_frame/topology_problem.py:398-409 reduced to a point load).
* ``frame``: frame-like block (96 x 96 x 16 or 256 x 256 x 48): Gaussian-smoothed
  random density clip(0.3 + 0.2 g / std(g), 1e-3, 1) with sigma = 3 cells,
  eight clamped 4 x 4-node pads on the bottom face, and 7 interface patches on
  the top face each loaded by the 6 unit wrenches (Fx, Fy, Fz, Mx, My, Mz) about
  the patch centre -> 42 right-hand sides, as InterfaceCovariance.wrenches
  (deep_frame/topology_problem.py:624-638) builds them.
* ``variants``: seeded uneven grids for batched solves.
"""
import numpy as np

from deep_frame_amd import element

MATERIAL = {"young": 1.0, "emin_ratio": 1e-9, "nu": 0.3, "penal": 3.0}

# patch centres as fractions of the (x, y) extent
PADS = [(0.125, 0.125), (0.125, 0.875), (0.875, 0.125), (0.875, 0.875),
        (0.5, 0.125), (0.5, 0.875), (0.125, 0.5), (0.875, 0.5)]
INTERFACES = {"motor_front_left": (0.2, 0.8), "motor_front_right": (0.8, 0.8), "motor_rear_left": (0.2, 0.2),
              "motor_rear_right": (0.8, 0.2), "stack": (0.5, 0.5), "battery": (0.5, 0.3), "camera": (0.5, 0.72)}


def patch_nodes(shape, centre, layer, size=4):
    """Node indices of a size x size node patch on the z = layer plane around fractional centre."""
    nx, ny, _ = shape
    i0 = int(round(centre[0] * nx)) - size // 2
    j0 = int(round(centre[1] * ny)) - size // 2
    i0 = min(max(i0, 0), nx + 1 - size)
    j0 = min(max(j0, 0), ny + 1 - size)
    ii, jj = np.meshgrid(np.arange(i0, i0 + size), np.arange(j0, j0 + size), indexing="ij")
    return element.node_index(shape)(ii.ravel(), jj.ravel(), np.full(ii.size, layer))


def relief_support(points, candidates):
    """Statically determinate (3-2-1) support DOFs of an inertia-relief fixture on candidate nodes
    (HexElasticity._relief_support, deep_frame/topology_optimization.py:511-525): node a (least y, then z)
    holds all three DOFs, node b (farthest from a) the two DOFs across a-b, node c (largest area with a-b)
    the DOF along the plane normal."""
    candidates = np.asarray(candidates, dtype=np.int64)
    pts = np.asarray(points, float)[candidates]
    a = int(np.argmin(pts[:, 1] + 1e-3 * pts[:, 2]))
    b = int(np.argmax(np.linalg.norm(pts - pts[a], axis=1)))
    normal = np.cross(pts - pts[a], pts[b] - pts[a])
    c = int(np.argmax(np.linalg.norm(normal, axis=1)))
    if np.linalg.norm(normal[c]) < 1e-9:
        raise ValueError("Inertia-relief support requires three non-collinear nodes")
    along, normal = np.abs(pts[b] - pts[a]), np.abs(np.cross(pts[c] - pts[a], pts[b] - pts[a]))
    nodes = candidates[[a, b, c]]
    dofs = [3 * nodes[0] + np.arange(3), 3 * nodes[1] + np.delete(np.arange(3), np.argmax(along)),
            [3 * nodes[2] + np.argmax(normal)]]
    return np.sort(np.concatenate(dofs))


def min_norm_wrenches(points, nodes, reference):
    """(3 * len(nodes), 6): minimum-norm nodal forces of the six unit wrenches about ``reference``.

    Solves B f = e_j with B = [I ... I; skew(r_1) ... skew(r_n)] in the least-norm sense,
    f = B^T (B B^T)^-1 e_j.
    """
    r = points[nodes] - np.asarray(reference, float)
    n = len(nodes)
    b = np.zeros((6, 3 * n))
    b[0, 0::3] = b[1, 1::3] = b[2, 2::3] = 1.0
    x, y, z = r[:, 0], r[:, 1], r[:, 2]
    b[3, 1::3], b[3, 2::3] = -z, y
    b[4, 0::3], b[4, 2::3] = z, -x
    b[5, 0::3], b[5, 1::3] = -y, x
    gram = b @ b.T
    return b.T @ np.linalg.inv(gram)


def clamp_dofs(nodes):
    nodes = np.asarray(nodes, dtype=np.int64)
    return (3 * nodes[:, None] + np.arange(3)).ravel()


def cantilever(seed=0, shape=(48, 24, 24)):
    rng = np.random.default_rng(seed)
    nelem = int(np.prod(shape))
    rho = 0.3 + rng.uniform(-0.05, 0.05, nelem)
    nx, ny, nz = shape
    idx = element.node_index(shape)
    jj, ll = np.meshgrid(np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    fixed = clamp_dofs(idx(np.zeros(jj.size, int), jj.ravel(), ll.ravel()))
    ndof = 3 * int(np.prod(np.asarray(shape) + 1))
    force = np.zeros((ndof, 1))
    force[3 * int(idx(nx, ny // 2, 0)) + 2, 0] = -1.0
    return {"shape": tuple(shape), "spacing": (1.0, 1.0, 1.0), "density": rho, "fixed": fixed, "forces": force,
            "material": dict(MATERIAL), "filter_radius": 1.5}


def smooth_density(shape, seed, sigma=3.0):
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(np.random.default_rng(seed).standard_normal(shape), sigma, mode="wrap")
    return np.clip(0.3 + 0.2 * g / g.std(), 1e-3, 1.0).ravel()


def frame(shape=(96, 96, 16), seed=0, rhs=42):
    shape = tuple(int(n) for n in shape)
    points = element.node_points(shape)
    fixed = clamp_dofs(np.concatenate([patch_nodes(shape, c, 0) for c in PADS]))
    ndof = 3 * len(points)
    columns = []
    for name, centre in INTERFACES.items():
        nodes = patch_nodes(shape, centre, shape[2])
        reference = points[nodes].mean(axis=0)
        forces = min_norm_wrenches(points, nodes, reference)
        block = np.zeros((ndof, 6))
        block[clamp_dofs(nodes)] = forces
        columns.append(block)
    forces = np.hstack(columns)[:, :rhs]
    return {"shape": shape, "spacing": (1.0, 1.0, 1.0), "density": smooth_density(shape, seed), "fixed": fixed,
            "forces": forces, "material": dict(MATERIAL), "filter_radius": 1.5}


def variants(count=64, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for v in range(count):
        shape = (int(rng.integers(48, 161)), int(rng.integers(48, 161)), int(rng.integers(12, 33)))
        out.append(frame(shape, seed=int(rng.integers(1 << 30)), rhs=int(rng.integers(6, 43))))
    return out


def moduli(density, material=MATERIAL):
    """E = E0 (eps + (1 - eps) rho^p) and dE/drho on the host (set-up and checks only)."""
    rho = np.asarray(density, float)
    eps, e0, p = material["emin_ratio"], material["young"], material["penal"]
    return e0 * (eps + (1 - eps) * rho ** p), e0 * (1 - eps) * p * rho ** (p - 1)
