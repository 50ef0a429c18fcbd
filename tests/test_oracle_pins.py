"""Closed-form pins of the oracle's smoothers, cycles and rediscretised levels (CPU only).

* Chebyshev (topology_multigrid.py:291-304): on a generalised eigenvector v of (A, D), D = diag(A), with
  D^-1 A v = lam v, `count` steps from zero leave the error T_d((theta - lam) / delta) / T_d(theta / delta) v
  (the Chebyshev semi-iteration's residual polynomial, d = count, theta and delta the centre and half-width
  of [1.1 lmax / ratio, 1.1 lmax]), evaluated here with numpy's Chebyshev series, not the recurrence.
* Jacobi (:286-290): the same eigenvector keeps (1 - damping lam / lmax)^count of its error.
* V and W cycles (:306-317): the cycle operator b -> x assembled column by column equals the one built from
  matrix algebra: smoothers as the polynomials r(D^-1 A) D^-1 with r(t) = (1 - p(t)) / t, the coarse-grid
  correction x += M P B_c M_c P^T M (b - A x) applied once (V) or twice (W), B at the coarsest level the
  inverse on the live DOFs.
* Rediscretised levels (:160-166): for moduli constant on every coarsest-level element the trilinear
  coarse spaces integrate the fine bilinear form exactly, so with no Dirichlet DOFs every rediscretised
  level equals the Galerkin level P^T A P; with a clamped face the coarse mask is the injected fine mask.
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as cheb
from numpy.polynomial import polynomial as poly
from scipy.linalg import eigh

from deep_frame_amd import element, problems
from oracle import gmg


def small_level(seed=0):
    """A 4x2x2 cantilever level: masked FP64 operator, mask and inverse diagonal."""
    shape = (4, 2, 2)
    ke = element.stiffness((1.0, 1.0, 1.0), element.isotropic(0.3))
    rho = np.random.default_rng(seed).uniform(0.2, 1.0, int(np.prod(shape)))
    pts = element.node_points(shape, (1.0, 1.0, 1.0))
    fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    levels = gmg.hierarchy(shape, ke, rho ** 3, fixed, (2, shape))
    return levels[0]


def live_eigenpairs(level):
    live = np.flatnonzero(level["mask"] > 0)
    a = level["A"].toarray()[np.ix_(live, live)]
    lam, vec = eigh(a, np.diag(np.diag(a)))
    return live, lam, vec


@pytest.mark.parametrize("count", [1, 2, 3, 6])
def test_chebyshev_error_is_the_scaled_chebyshev_polynomial(count):
    level = small_level()
    live, lam, vec = live_eigenpairs(level)
    lmax = lam.max() * 0.97  # the power-iteration estimate sits below the true maximum
    upper = 1.1 * lmax
    lower = upper / 20.0
    theta, delta = (upper + lower) / 2, (upper - lower) / 2
    series = np.zeros(count + 1)
    series[count] = 1.0
    for i in (0, len(lam) // 2, len(lam) - 1):
        e = np.zeros(level["A"].shape[0])
        e[live] = vec[:, i]
        x = gmg.chebyshev(level, level["A"] @ e, None, count, lmax)
        factor = cheb.chebval((theta - lam[i]) / delta, series) / cheb.chebval(theta / delta, series)
        assert np.linalg.norm((e - x) - factor * e) < 1e-10 * np.linalg.norm(e), (i, factor)
        # from a nonzero start the same polynomial acts on the initial error
        x0 = 0.3 * e
        x = gmg.chebyshev(level, level["A"] @ e, x0, count, lmax)
        assert np.linalg.norm((e - x) - factor * 0.7 * e) < 1e-10 * np.linalg.norm(e)


@pytest.mark.parametrize("count", [1, 2, 3, 6])
def test_fourth_kind_chebyshev_error_is_the_closed_form(count):
    """The fourth-kind smoother's error polynomial is W_k(1 - 2t) / (2k + 1), t = lam / upper, which for
    t in (0, 1] equals sin((2k + 1) asin(sqrt t)) / ((2k + 1) sqrt t); evaluated here from the sine, not
    from the three-term recurrence the smoother runs."""
    level = small_level(2)
    live, lam, vec = live_eigenpairs(level)
    lmax = lam.max() * 0.97
    upper = 1.1 * lmax
    for i in (0, len(lam) // 2, len(lam) - 1):
        e = np.zeros(level["A"].shape[0])
        e[live] = vec[:, i]
        t = lam[i] / upper
        factor = np.sin((2 * count + 1) * np.arcsin(np.sqrt(t))) / ((2 * count + 1) * np.sqrt(t))
        x = gmg.chebyshev(level, level["A"] @ e, None, count, lmax, kind=4)
        assert np.linalg.norm((e - x) - factor * e) < 1e-10 * np.linalg.norm(e), (i, factor)
        x = gmg.chebyshev(level, level["A"] @ e, 0.3 * e, count, lmax, kind=4)
        assert np.linalg.norm((e - x) - factor * 0.7 * e) < 1e-10 * np.linalg.norm(e)


@pytest.mark.parametrize("count,damping", [(1, 1.0), (2, 1.0), (4, 0.7)])
def test_jacobi_error_factor(count, damping):
    level = small_level(1)
    live, lam, vec = live_eigenpairs(level)
    lmax = lam.max()
    for i in (0, len(lam) // 3, len(lam) - 1):
        e = np.zeros(level["A"].shape[0])
        e[live] = vec[:, i]
        x = gmg.jacobi(level, level["A"] @ e, None, count, lmax, damping)
        assert np.linalg.norm((e - x) - (1 - damping * lam[i] / lmax) ** count * e) < 1e-11


def residual_polynomial(kind, count, lmax, damping=1.0):
    """Power-basis coefficients of the smoother's error polynomial p(t) in t = eigenvalue of D^-1 A."""
    if kind == "jacobi":
        return poly.polypow([1.0, -damping / lmax], count)
    upper = 1.1 * lmax
    lower = upper / 20.0
    theta, delta = (upper + lower) / 2, (upper - lower) / 2
    series = np.zeros(count + 1)
    series[count] = 1.0
    inner = cheb.cheb2poly(series)  # T_d(s) in powers of s, s = (theta - t) / delta
    out = np.zeros(1)
    for power, c in enumerate(inner):
        out = poly.polyadd(out, c * poly.polypow([theta / delta, -1.0 / delta], power))
    return out / cheb.chebval(theta / delta, series)


def smoother_matrix(level, coefficients):
    """r(D^-1 A) D^-1 with r(t) = (1 - p(t)) / t (p(0) = 1, so the division is exact)."""
    one_minus = poly.polysub([1.0], coefficients)
    assert abs(one_minus[0]) < 1e-12
    r = one_minus[1:]
    a = level["A"].toarray()
    dinv = np.diag(level["dinv"])
    t = dinv @ a
    out = np.zeros_like(a)
    power = np.eye(len(a))
    for c in r:
        out += c * power
        power = power @ t
    return out @ dinv


def cycle_matrix(levels, lmax, index, kind, cycle, degree=3, growth=2.0, sweeps=2):
    lv = levels[index]
    n = lv["A"].shape[0]
    if index == len(levels) - 1:
        live = np.flatnonzero(lv["mask"] > 0)
        out = np.zeros((n, n))
        out[np.ix_(live, live)] = np.linalg.inv(lv["A"].toarray()[np.ix_(live, live)])
        return out
    count = max(1, int(round((sweeps if kind == "jacobi" else degree) * growth ** index)))
    s = smoother_matrix(lv, residual_polynomial(kind, count, lmax[index]))
    a = lv["A"].toarray()
    m = np.diag(lv["mask"])
    p = levels[index + 1]["P"].toarray()
    mc = np.diag(levels[index + 1]["mask"])
    coarse = cycle_matrix(levels, lmax, index + 1, kind, cycle, degree, growth, sweeps)
    correction = m @ p @ coarse @ mc @ p.T @ m
    b = s.copy()  # x = S b
    for _ in range(2 if cycle == "W" and index + 2 < len(levels) else 1):
        b = b + correction @ (np.eye(n) - a @ b)
    return b + s @ (np.eye(n) - a @ b)


@pytest.mark.parametrize("kind", ["chebyshev", "jacobi"])
@pytest.mark.parametrize("cycle", ["V", "W"])
def test_cycle_operator_matches_matrix_recursion(kind, cycle):
    shape = (8, 4, 4)
    ke = element.stiffness((1.0, 1.0, 1.0), element.isotropic(0.3))
    rho = np.random.default_rng(4).uniform(0.1, 1.0, int(np.prod(shape)))
    pts = element.node_points(shape, (1.0, 1.0, 1.0))
    fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    levels = gmg.hierarchy(shape, ke, 1e-3 + rho ** 3, fixed, (3, shape))
    lmax = [gmg.spectrum(lv) for lv in levels[:-1]]
    n = levels[0]["A"].shape[0]
    live = np.flatnonzero(levels[0]["mask"] > 0)
    expected = cycle_matrix(levels, lmax, 0, kind, cycle)
    got = np.zeros((n, n))
    for j in live:
        e = np.zeros(n)
        e[j] = 1.0
        got[:, j] = gmg.vcycle(levels, lmax, e, smoother=kind, cycle=cycle)
    scale = np.abs(expected[np.ix_(live, live)]).max()
    assert np.abs(got[np.ix_(live, live)] - expected[np.ix_(live, live)]).max() < 1e-9 * scale
    if cycle == "W":  # the second correction changes the operator
        v = cycle_matrix(levels, lmax, 0, kind, "V")
        assert np.abs(v[np.ix_(live, live)] - expected[np.ix_(live, live)]).max() > 1e-4 * scale


def test_rediscretized_levels_equal_galerkin_for_blockwise_constant_moduli():
    shape = (8, 8, 4)
    ke = element.stiffness((1.0, 1.0, 1.0), element.isotropic(0.3))
    blocks = np.random.default_rng(5).uniform(0.1, 1.0, (2, 2, 1))
    moduli = np.kron(blocks, np.ones((4, 4, 4)))  # constant on every coarsest-level element
    counts = (3, shape)
    galerkin = gmg.hierarchy(shape, ke, moduli.ravel(), [], counts)
    redisc = gmg.hierarchy(shape, ke, moduli.ravel(), [], counts, coarse="rediscretize")
    for g, r in zip(galerkin[1:], redisc[1:]):
        diff = abs(g["A"] - r["A"]).max()
        assert diff < 1e-12 * abs(g["A"]).max()
        assert np.array_equal(g["mask"], r["mask"])
    # a non-uniform field breaks the identity (the rediscretised operator is not P^T A P)
    rough = gmg.hierarchy(shape, ke, np.random.default_rng(6).uniform(0.1, 1.0, 256), [], counts,
                          coarse="rediscretize")
    rough_g = gmg.hierarchy(shape, ke, np.random.default_rng(6).uniform(0.1, 1.0, 256), [], counts)
    assert abs(rough[1]["A"] - rough_g[1]["A"]).max() > 1e-3 * abs(rough_g[1]["A"]).max()


def test_rediscretized_mask_is_the_injected_fine_mask():
    shape = (8, 4, 4)
    ke = element.stiffness((1.0, 1.0, 1.0), element.isotropic(0.3))
    pts = element.node_points(shape, (1.0, 1.0, 1.0))
    fixed = problems.clamp_dofs(np.flatnonzero(np.abs(pts[:, 0]) < 1e-9))
    levels = gmg.hierarchy(shape, ke, np.ones(128), fixed, (3, shape), coarse="rediscretize")
    coarse_pts = element.node_points((4, 2, 2), (2.0, 2.0, 2.0))
    face = np.flatnonzero(np.abs(coarse_pts[:, 0]) < 1e-9)
    mask = levels[1]["mask"].reshape(-1, 3)
    assert np.all(mask[face] == 0) and mask.sum() == 3 * (len(coarse_pts) - len(face))
    # Galerkin keeps the coarse face DOFs (their prolongation reaches the free x = 1 fine nodes)
    galerkin = gmg.hierarchy(shape, ke, np.ones(128), fixed, (3, shape))
    assert np.all(galerkin[1]["mask"].reshape(-1, 3)[face] == 1)


# ---- MMA (Svanberg 1987 / KTH note 2007; mmasub as called by topology_optimizers.py:55-67) ----------------
# Hand-derived numbers: bounds [0, 1] (span 1), asyinit 0.5, asydecr 0.7, asyincr 1.2, asymptote clamps
# x -+ 10 span and x -+ 0.01 span, albefa 0.1, raa0 1e-5, the 0.001 (p + q) term.

def test_mma_asymptotes_over_four_iterations():
    from oracle import mma
    x = np.array([0.3, 0.6])
    for it in (1, 2):  # low = x - 0.5, upp = x + 0.5
        low, upp = mma.asymptotes(it, x, x, x, None, None, 0.0, 1.0)
        assert np.allclose(low, [-0.2, 0.1], atol=1e-15) and np.allclose(upp, [0.8, 1.1], atol=1e-15)
    # iteration 3: variable 0 keeps its direction (factor 1.2), variable 1 stalled (factor 1)
    low, upp = mma.asymptotes(3, np.array([0.5, 0.5]), np.array([0.4, 0.6]), np.array([0.3, 0.6]),
                              np.array([0.0, 0.2]), np.array([0.9, 1.0]), 0.0, 1.0)
    assert np.allclose(low, [0.02, 0.1], atol=1e-15) and np.allclose(upp, [1.1, 0.9], atol=1e-15)
    # iteration 4: variable 0 oscillates (factor 0.7, upp clamped up to x + 0.01), variable 1 keeps its
    # direction (factor 1.2, low clamped up to x - 10)
    low, upp = mma.asymptotes(4, np.array([0.5, 0.45]), np.array([0.6, 0.5]), np.array([0.4, 0.52]),
                              np.array([0.3, -30.0]), np.array([0.605, 0.9]), 0.0, 1.0)
    assert np.allclose(low, [0.29, -9.55], atol=1e-14) and np.allclose(upp, [0.51, 0.93], atol=1e-14)


def test_mma_approximation_terms():
    from oracle import mma
    # x = 0.5 between asymptotes 0 and 1, df0 = 2, dg = -1 (already scaled), g = -0.1, move 0.1
    alfa, beta, p0, q0, p, q, b = mma.approximation(np.array([0.5]), np.array([0.0]), np.array([1.0]), 0.0, 1.0,
                                                    np.array([2.0]), np.array([-0.1]), np.array([[-1.0]]), 0.1)
    assert alfa[0] == pytest.approx(0.4) and beta[0] == pytest.approx(0.6)  # move limits inside the albefa bounds
    assert p0[0] == pytest.approx((2 + 0.002 + 1e-5) * 0.25, rel=1e-14)  # 0.5005025
    assert q0[0] == pytest.approx((0.002 + 1e-5) * 0.25, rel=1e-14)      # 0.0005025
    assert p[0, 0] == pytest.approx((0.001 + 1e-5) * 0.25, rel=1e-14)    # 0.0002525
    assert q[0, 0] == pytest.approx((1.001 + 1e-5) * 0.25, rel=1e-14)    # 0.2502525
    assert b[0] == pytest.approx(0.0002525 / 0.5 + 0.2502525 / 0.5 + 0.1, rel=1e-14)


def _root(a2, a1, a0):
    """Larger root of a2 x^2 + a1 x + a0."""
    return (-a1 + np.sqrt(a1 * a1 - 4 * a2 * a0)) / (2 * a2)


@pytest.mark.parametrize("case", ["one_active", "second_inactive", "both_active"])
def test_mma_subproblem_closed_form_solutions(case):
    """Two variables at x = 0.5 (iteration 1: asymptotes 0 and 1, move 0.1 -> [0.4, 0.6]), objective
    slope -1 and constraint slopes +1 (scaled), so each term is p / (1 - x) + q / x with p = 0.2502525,
    q = 0.0002525 (constraints) and p0 = 0.0002525, q0 = 0.2502525 (objective).  An active constraint j
    fixes its variables by sum_i [p / (1 - x_i) + q / x_i] = b_j, a quadratic in x."""
    from oracle import mma
    x = np.array([0.5, 0.5])
    df0 = np.array([-1.0, -1.0])
    if case == "one_active":  # x1 + x2 <= 1.1: symmetric, 2 [p / (1 - x) + q / x] = 1.10202
        g, dg = np.array([-0.1]), np.array([[1.0, 1.0]])
        expected = np.full(2, _root(0.55101, 0.2502525 - 0.0002525 - 0.55101, 0.0002525))
    elif case == "second_inactive":  # and x1 <= 0.9: its bound 0.90101 is not reached at the same point
        g, dg = np.array([-0.1, -0.4]), np.array([[1.0, 1.0], [1.0, 0.0]])
        expected = np.full(2, _root(0.55101, 0.2502525 - 0.0002525 - 0.55101, 0.0002525))
    else:  # and x1 <= 0.52: x1 from p / (1 - x1) + q / x1 = 0.52101, x2 from the rest 1.10202 - 0.52101
        g, dg = np.array([-0.1, -0.02]), np.array([[1.0, 1.0], [1.0, 0.0]])
        expected = np.array([_root(0.52101, 0.2502525 - 0.0002525 - 0.52101, 0.0002525),
                             _root(0.58101, 0.2502525 - 0.0002525 - 0.58101, 0.0002525)])
    assert np.all((expected > 0.4) & (expected < 0.6))
    xn, low, upp, lam, _ = mma.mma_step(1, x, x, x, None, None, 0.0, 1.0, 0.0, df0, g, dg, move=0.1, scale=1.0)
    assert np.allclose(low, 0.0) and np.allclose(upp, 1.0)
    assert np.max(np.abs(xn - expected)) < 1e-6, (xn, expected)
    if case == "second_inactive":
        assert lam[1] < 1e-5 and lam[0] > 0.1
    if case == "both_active":
        assert lam[0] > 0.1 and lam[1] > 0.1


def _dense_cone(allowed, spacing, radius):
    """Row-normalised cone matrix from its definition: H_ij = a_i a_j max(r - |x_i - x_j|, 0) / sum_j (...),
    rows of non-allowed cells zero."""
    cells = np.argwhere(np.ones(allowed.shape, bool)) * np.asarray(spacing, float)
    a = allowed.ravel().astype(float)
    dist = np.linalg.norm(cells[:, None, :] - cells[None, :, :], axis=2)
    w = np.maximum(radius - dist, 0.0) * a[:, None] * a[None, :]
    sums = w.sum(axis=1)
    sums[sums == 0] = 1.0
    return w / sums[:, None]


def test_corridor_filter_is_the_blended_dense_matrix_and_its_adjoint():
    """oracle.filters.Cone with a corridor equals (1 - b) H_r + b H_R built from the definition, and its
    pullback is the transposed matrix on the free cells (DensityMap corridor,
    topology_optimization.py:1125-1129, :1142, :1176)."""
    from oracle import filters
    rng = np.random.default_rng(7)
    shape, spacing = (5, 4, 3), (1.0, 0.8, 1.2)
    allowed = rng.random(shape) < 0.8
    preserve = allowed & (rng.random(shape) < 0.1)
    weights = np.where(rng.random(shape) < 0.5, rng.uniform(-0.3, 1.4, shape), 0.0)
    cone = filters.Cone(allowed, spacing, 1.5, preserve, corridor=(2.6, weights))
    b = (np.clip(weights, 0, 1) * allowed).ravel()
    h = (1 - b)[:, None] * _dense_cone(allowed, spacing, 1.5) + b[:, None] * _dense_cone(allowed, spacing, 2.6)
    x = rng.uniform(0, 1, allowed.size)
    assert np.max(np.abs(cone.filtered(x) - h @ x)) < 1e-14
    s, d = rng.normal(size=allowed.size), rng.uniform(0.5, 1.5, allowed.size)
    free = (allowed & ~preserve).ravel()
    assert np.max(np.abs(cone.pullback(s, d) - np.where(free, h.T @ (s * d), 0.0))) < 1e-13


def test_volume_weights_of_soft_discs():
    from deep_frame_amd.density import volume_weights
    discs = {"mode": "soft", "motors_mm": [[0.0, 0.0], [10.0, 0.0]], "radius_mm": 4.0, "weight": 2.0,
             "plane_mm": 1.0, "length_mm": 3.0}
    points = np.array([[0.0, 0.0, 1.0],    # on a motor axis in the disc plane: 1 + weight
                       [10.0, 3.0, 4.0],   # inside the second disc, 3 mm off the plane: 1 + 2 e^-1
                       [5.0, 0.0, 1.0],    # between the discs (r = 5 > 4): 1
                       [0.0, 4.0, 1.0]])   # on the disc rim (r = radius counts as inside): 1 + weight
    assert np.allclose(volume_weights(points, discs), [3.0, 1 + 2 * np.exp(-1.0), 1.0, 3.0])
    ramp = {**discs, "hub_radius_mm": 1.0, "radial_power": 2.0}  # ((r - 1) / 3)^2 inside the disc
    assert np.allclose(volume_weights(points, ramp), [1.0, 1 + 2 * (2 / 3) ** 2 * np.exp(-1.0), 1.0, 3.0])
    assert np.all(volume_weights(points, {**discs, "mode": "hard"}) == 1.0)
    assert np.all(volume_weights(points, None) == 1.0)
