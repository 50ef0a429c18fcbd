"""Geometric multigrid PCG semantics on assembled sparse matrices (FP64 oracle).

Restates ``deep_frame/topology_multigrid.py``:
  * level count and padding (:58-63),
  * fine mask = active * (1 - constrained) (:138-144), coarse dead-DOF mask by
    diagonal ratio (:180-181),
  * trilinear prolongation / full-weighting restriction, stencil (1/2, 1, 1/2)^3
    (:69, :314-316); restriction = prolongation^T,
  * Galerkin coarse operator A_c = P^T (M A M) P (:174-181),
  * Chebyshev smoother with Jacobi scaling, degree 3 * 2^level, bounds
    [1.1 lmax / 20, 1.1 lmax] (:284-305), or damped Jacobi with step damping / lmax and
    round(sweeps * growth^level) sweeps (:284-290),
  * V-cycle, or W-cycle (two coarse-grid corrections on every level whose coarser level is not the
    coarsest) (:306-317), with an exact coarsest solve,
  * rediscretised coarse levels (``coarse = "rediscretize"``, :160-166): moduli averaged over 2x2x2
    children, Ke scaled by the spacing ratio, mask injected from the finer level and cut by the dead-DOF
    ratio,
  * (no counterpart in the reference) the K-cycle variant of the fine level's V-cycle: the level-1
    system of the coarse-grid correction is solved by a few flexible-CG steps preconditioned by the
    level-1 V-cycle (``fcg``, ``kcycle``),
  * flexible PCG with Polak-Ribiere beta and true-residual confirmation (:407-454).
The smoothers and cycles also take an (ndof, k) block of independent columns.
Matrices are SciPy sparse over padded node grids with DOF = 3*node + comp.
Test infrastructure only.
"""
import numpy as np
from scipy.sparse import coo_matrix, diags, kron, identity
from scipy.sparse.linalg import splu

from oracle.hex8 import assemble, grid


def level_count(shape, coarsest_dofs=80000, max_levels=8):
    shape = np.asarray(shape)
    count = 2
    while (count < max_levels and np.all(np.ceil(shape / 2 ** count) >= 1)
           and np.prod(np.ceil(shape / 2 ** (count - 1)) + 1) * 3 > coarsest_dofs):
        count += 1
    factor = 2 ** (count - 1)
    padded = tuple(int(np.ceil(n / factor) * factor) for n in shape)
    return count, padded


def prolongation_1d(nc):
    """(2 nc + 1) x (nc + 1) linear interpolation."""
    rows, cols, vals = [], [], []
    for i in range(2 * nc + 1):
        if i % 2 == 0:
            rows.append(i), cols.append(i // 2), vals.append(1.0)
        else:
            rows += [i, i]
            cols += [i // 2, i // 2 + 1]
            vals += [0.5, 0.5]
    return coo_matrix((vals, (rows, cols)), shape=(2 * nc + 1, nc + 1)).tocsr()


def prolongation(coarse_elems):
    """Fine-DOF x coarse-DOF trilinear prolongation, node-major, comp-minor."""
    px, py, pz = (prolongation_1d(n) for n in coarse_elems)
    return kron(kron(kron(px, py), pz), identity(3)).tocsr()


def fine_operator(shape_padded, ke, moduli_padded):
    ndof = 3 * int(np.prod(np.asarray(shape_padded) + 1))
    _, _, dofs = grid(shape_padded)
    return assemble(dofs, ke, moduli_padded.ravel(), ndof)


def fine_mask(shape_padded, moduli_padded, fixed_padded):
    _, conn, _ = grid(shape_padded)
    nnode = int(np.prod(np.asarray(shape_padded) + 1))
    active = np.zeros(nnode, bool)
    active[np.unique(conn[moduli_padded.ravel() > 0])] = True
    mask = np.repeat(active, 3).astype(float)
    mask[np.asarray(fixed_padded, dtype=np.int64)] = 0.0
    return mask


def pool(moduli):
    """Mean of the 2x2x2 children of every coarse element (avg_pool3d, topology_multigrid.py:161)."""
    m = np.asarray(moduli, float)
    nx, ny, nz = (n // 2 for n in m.shape)
    return m.reshape(nx, 2, ny, 2, nz, 2).mean(axis=(1, 3, 5))


def inject(mask, shape):
    """DOF mask of the coarse node grid (shape elements) taken at the even nodes of the finer one."""
    fine_nodes = tuple(2 * n + 1 for n in shape)
    grid3 = np.asarray(mask).reshape(fine_nodes + (3,))
    return grid3[::2, ::2, ::2].reshape(-1).copy()


def hierarchy(shape, ke, moduli, fixed, levels, dead_ratio=1e-12, coarse="galerkin"):
    """List of dicts with 'A' (masked), 'mask', 'dinv', 'P' (to the next finer level)."""
    _, padded = levels
    count = levels[0]
    mpad = np.zeros(padded)
    mpad[:shape[0], :shape[1], :shape[2]] = np.asarray(moduli, float).reshape(shape)
    a = fine_operator(padded, ke, mpad)
    m = fine_mask(padded, mpad, fixed)
    a = diags(m) @ a @ diags(m)
    out = [{"A": a.tocsr(), "mask": m, "shape": padded, "moduli": mpad}]
    for lvl in range(1, count):
        cshape = tuple(n // 2 ** lvl for n in padded)
        p = prolongation(cshape)
        if coarse == "rediscretize":
            # coarse operator of the averaged moduli with Ke(2^l h) = 2^l Ke(h) (3-D elasticity), assembled
            # with the injected mask on both sides; the dead-DOF ratio then cuts the injected mask
            moduli_c = pool(out[-1]["moduli"])
            injected = inject(out[-1]["mask"], cshape)
            ac = fine_operator(cshape, ke * 2.0 ** lvl, moduli_c)
            ac = (diags(injected) @ ac @ diags(injected)).tocsr()
            dg = ac.diagonal()
            cm = injected * (dg > dead_ratio * dg.max())
        else:
            moduli_c = None
            ac = (p.T @ out[-1]["A"] @ p).tocsr()
            ac = ((ac + ac.T) * 0.5).tocsr()
            dg = ac.diagonal()
            cm = (dg > dead_ratio * dg.max()).astype(float)
        ac = (diags(cm) @ ac @ diags(cm)).tocsr()
        out.append({"A": ac, "mask": cm, "shape": cshape, "P": p, "moduli": moduli_c})
    for lv in out:
        dg = lv["A"].diagonal()
        lv["dinv"] = np.where(lv["mask"] > 0, 1.0 / np.where(dg > 0, dg, 1.0), 0.0)
    return out


def spectrum(level, iterations=60, seed=0):
    """lambda_max(D^-1 A) by power iteration (the reference uses 20 steps, :214-223)."""
    v = (np.random.default_rng(seed).random(level["A"].shape[0]) - 0.5) * level["mask"]
    est = 1.0
    for _ in range(iterations):
        v = level["dinv"] * (level["A"] @ v)
        est = np.linalg.norm(v)
        v = v / est
    return est


def _cols(v, like):
    """A per-DOF vector broadcast against a vector or an (ndof, k) block of columns."""
    return v if np.ndim(like) == 1 else v[:, None]


def chebyshev(level, rhs, x, count, lmax, ratio=20.0, kind=1):
    """``count`` Chebyshev smoothing steps on level A x = rhs from x (None: zero), Jacobi-preconditioned.
    kind 1: the reference's first-kind polynomial on [1.1 lmax / ratio, 1.1 lmax] (topology_multigrid.py:291-304);
    kind 4: the fourth-kind polynomial (Lottes 2023, no reference counterpart), d_0 = 4 / (3 upper) Dinv r_0,
    d_i = (2i - 1) / (2i + 3) d_{i-1} + (8i + 4) / ((2i + 3) upper) Dinv r_i, upper = 1.1 lmax."""
    a, dinv = level["A"], _cols(level["dinv"], rhs)
    upper = 1.1 * lmax
    if kind == 4:
        r = rhs.copy() if x is None else rhs - a @ x
        d = 4.0 / (3.0 * upper) * dinv * r
        x = d.copy() if x is None else x + d
        for i in range(1, count):
            r = r - a @ d
            d = (2 * i - 1) / (2 * i + 3) * d + (8 * i + 4) / ((2 * i + 3) * upper) * dinv * r
            x = x + d
        return x
    lower = upper / ratio
    theta, delta = (upper + lower) / 2, (upper - lower) / 2
    sigma = theta / delta
    rho = 1 / sigma
    r = rhs.copy() if x is None else rhs - a @ x
    d = dinv * r / theta
    x = d.copy() if x is None else x + d
    for _ in range(count - 1):
        r = r - a @ d
        new = 1 / (2 * sigma - rho)
        d = new * rho * d + 2 * new / delta * dinv * r
        rho = new
        x = x + d
    return x


def jacobi(level, rhs, x, count, lmax, damping=1.0):
    """Damped Jacobi sweeps x += (damping / lmax) Dinv (rhs - A x), the first from x = 0 when x is None."""
    step = damping / lmax
    dinv = _cols(level["dinv"], rhs)
    for _ in range(count):
        x = step * dinv * rhs if x is None else x + step * dinv * (rhs - level["A"] @ x)
    return x


def vcycle(levels, lmax, rhs, index=0, degree=3, growth=2.0, smoother="chebyshev", sweeps=2, damping=1.0,
           cycle="V", kind=1):
    lv = levels[index]
    if index == len(levels) - 1:
        live = np.flatnonzero(lv["mask"] > 0)
        out = np.zeros_like(rhs)
        sub = lv["A"][live][:, live].tocsc()
        out[live] = splu(sub).solve(rhs[live])
        return out
    count = max(1, int(round((sweeps if smoother == "jacobi" else degree) * growth ** index)))

    def smooth(x):
        if smoother == "jacobi":
            return jacobi(lv, rhs, x, count, lmax[index], damping)
        return chebyshev(lv, rhs, x, count, lmax[index], kind=kind)

    x = smooth(None)
    p = levels[index + 1]["P"]
    for _ in range(2 if cycle == "W" and index + 2 < len(levels) else 1):
        res = (rhs - lv["A"] @ x) * _cols(lv["mask"], rhs)
        corr = vcycle(levels, lmax, (p.T @ res) * _cols(levels[index + 1]["mask"], rhs), index + 1, degree, growth,
                      smoother, sweeps, damping, cycle, kind)
        x = x + (p @ corr) * _cols(lv["mask"], rhs)
    return smooth(x)


def kvcycle(levels, lmax, rhs, index, steps, krylov_levels, degree=3, growth=2.0, kind=1):
    """V-cycle from level ``index`` whose coarse-grid correction on the first ``krylov_levels`` levels (counted
    from the fine level 0) is ``steps`` flexible-CG steps on the next level (``fcg``), as long as that level is
    not the coarsest; below, the plain V-cycle."""
    lv = levels[index]
    if index == len(levels) - 1 or not (index < krylov_levels and index + 2 < len(levels)):
        return vcycle(levels, lmax, rhs, index, degree, growth, kind=kind)
    count = max(1, int(round(degree * growth ** index)))
    x = chebyshev(lv, rhs, None, count, lmax[index], kind=kind)
    res = (rhs - lv["A"] @ x) * _cols(lv["mask"], rhs)
    p = levels[index + 1]["P"]
    corr = fcg(levels, lmax, index + 1, (p.T @ res) * _cols(levels[index + 1]["mask"], rhs), steps, degree, growth,
               krylov_levels, kind)
    x = x + (p @ corr) * _cols(lv["mask"], rhs)
    return chebyshev(lv, rhs, x, count, lmax[index], kind=kind)


def fcg(levels, lmax, index, rhs, steps, degree=3, growth=2.0, krylov_levels=0, kind=1):
    """``steps`` iterations of flexible CG on A_index x = rhs from x = 0, each preconditioned by the V-cycle
    from level ``index`` (itself a K-cycle while index < krylov_levels): alpha = r.z / p.A p,
    beta = r_new.(z_new - z_old) / r_old.z_old."""
    a = levels[index]["A"]
    x = np.zeros_like(rhs)
    r = rhs.copy()
    z = kvcycle(levels, lmax, r, index, steps, krylov_levels, degree, growth, kind)
    p = z.copy()
    rz = r @ z
    for step in range(steps):
        ap = a @ p
        pap = p @ ap
        if not (rz > 0 and pap > 0):
            break
        alpha = rz / pap
        x += alpha * p
        if step == steps - 1:
            break
        r = r - alpha * ap
        znew = kvcycle(levels, lmax, r, index, steps, krylov_levels, degree, growth, kind)
        beta = (r @ (znew - z)) / rz
        rz = r @ znew
        p = znew + beta * p
        z = znew
    return x


def kcycle(levels, lmax, rhs, steps=2, degree=3, growth=2.0, krylov_levels=1, kind=1):
    """The fine level's cycle with the coarse-grid corrections of the first ``krylov_levels`` levels solved by
    ``fcg`` (``steps`` flexible-CG steps) instead of one V-cycle; with two levels the level-1 solve is the
    exact coarsest one."""
    if len(levels) < 3:
        return vcycle(levels, lmax, rhs, 0, degree, growth, kind=kind)
    return kvcycle(levels, lmax, rhs, 0, steps, krylov_levels, degree, growth, kind)


def pcg(levels, lmax, rhs, tol=1e-8, maxit=2000, precondition=None):
    """Flexible PCG, Polak-Ribiere beta, per-column; returns (x, iterations, rel. residual).
    ``precondition(r)`` defaults to the V-cycle."""
    if precondition is None:
        precondition = lambda r: vcycle(levels, lmax, r)
    a = levels[0]["A"]
    b = rhs * levels[0]["mask"]
    nb = np.linalg.norm(b)
    x = np.zeros_like(b)
    r = b.copy()
    z = precondition(r)
    p = z.copy()
    rz = r @ z
    for it in range(1, maxit + 1):
        ap = a @ p
        alpha = rz / (p @ ap)
        x += alpha * p
        r -= alpha * ap
        rel = np.linalg.norm(r) / nb
        if rel < tol:
            true = b - a @ x
            rel = np.linalg.norm(true) / nb
            if rel < tol:
                return x, it, rel
            r = true
        znew = precondition(r)
        beta = (r @ (znew - z)) / rz
        rz = r @ znew
        p = znew + beta * p
        z = znew
    return x, maxit, np.linalg.norm(b - a @ x) / nb
