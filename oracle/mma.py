"""Method of Moving Asymptotes step (FP64 oracle).

The reference calls ``mmapy.mmasub`` (mmapy 0.3.1, a Python port of
Svanberg's MMA; not vendored under the reference and not installed here) from
``MMAOptimizer.step`` (``deep_frame/topology_optimizers.py:55-67``) with
a0 = 1, a = 0, c = 1e4, d = 1, f0/df0/g/dg scaled by 100, asyinit 0.5,
asydecr 0.7, asyincr 1.2, raa0 1e-5, albefa 0.1, move 0.1 (0.05 late) (:7).

This module restates the published method (K. Svanberg, "The method of moving
asymptotes", IJNME 24 (1987); "MMA and GCMMA - two methods for nonlinear
optimization", KTH note 2007): asymptote update, move-limited bounds
alpha/beta, the convex separable approximation p0, q0, P, Q, b, and the
primal-dual interior-point subproblem solver (epsilon continuation 1 -> 1e-7).
Parity against mmapy itself is unpinned (absent); the tests pin the
subproblem solution by its KKT conditions and by an independent dual solver.
Test infrastructure only.
"""
import numpy as np


def asymptotes(it, x, xold1, xold2, low, upp, xmin, xmax, asyinit=0.5, asydecr=0.7, asyincr=1.2):
    span = np.broadcast_to(np.asarray(xmax - xmin, float), x.shape)
    if it <= 2:
        return x - asyinit * span, x + asyinit * span
    sign = (x - xold1) * (xold1 - xold2)
    f = np.where(sign > 0, asyincr, np.where(sign < 0, asydecr, 1.0))
    low = x - f * (xold1 - low)
    upp = x + f * (upp - xold1)
    low = np.minimum(np.maximum(low, x - 10 * span), x - 0.01 * span)
    upp = np.maximum(np.minimum(upp, x + 10 * span), x + 0.01 * span)
    return low, upp


def approximation(x, low, upp, xmin, xmax, df0, g, dg, move, raa0=1e-5, albefa=0.1):
    """alpha, beta, p0, q0, P (m,n), Q (m,n), b (m,)."""
    span = xmax - xmin
    alfa = np.maximum(np.maximum(low + albefa * (x - low), x - move * span), xmin)
    beta = np.minimum(np.minimum(upp - albefa * (upp - x), x + move * span), xmax)
    inv = np.broadcast_to(1.0 / np.maximum(span, 1e-5), x.shape)
    ux2, xl2 = (upp - x) ** 2, (x - low) ** 2
    p0, q0 = np.maximum(df0, 0), np.maximum(-df0, 0)
    pq0 = 0.001 * (p0 + q0) + raa0 * inv
    p0, q0 = (p0 + pq0) * ux2, (q0 + pq0) * xl2
    p, q = np.maximum(dg, 0), np.maximum(-dg, 0)
    pq = 0.001 * (p + q) + raa0 * inv[None, :]
    p, q = (p + pq) * ux2[None, :], (q + pq) * xl2[None, :]
    b = p @ (1.0 / (upp - x)) + q @ (1.0 / (x - low)) - g
    return alfa, beta, p0, q0, p, q, b


def subsolve(low, upp, alfa, beta, p0, q0, p, q, a0, a, b, c, d, epsimin=1e-7):
    """Primal-dual Newton on the KKT system of the MMA subproblem (Svanberg 2007, sec. 5)."""
    m, n = p.shape
    x = 0.5 * (alfa + beta)
    y, z, lam = np.ones(m), 1.0, np.ones(m)
    xsi = np.maximum(1.0 / (x - alfa), 1.0)
    eta = np.maximum(1.0 / (beta - x), 1.0)
    mu, zet, s = np.maximum(1.0, 0.5 * c), 1.0, np.ones(m)
    epsi = 1.0

    def residual(x, y, z, lam, xsi, eta, mu, zet, s, epsi):
        ux, xl = upp - x, x - low
        plam, qlam = p0 + p.T @ lam, q0 + q.T @ lam
        gvec = p @ (1 / ux) + q @ (1 / xl)
        r = [plam / ux ** 2 - qlam / xl ** 2 - xsi + eta, c + d * y - mu - lam, [a0 - zet - a @ lam],
             gvec - a * z - y + s - b, xsi * (x - alfa) - epsi, eta * (beta - x) - epsi,
             mu * y - epsi, [zet * z - epsi], lam * s - epsi]
        r = np.concatenate([np.atleast_1d(v) for v in r])
        return np.sqrt(r @ r), np.max(np.abs(r))

    while epsi > epsimin:
        norm, rmax = residual(x, y, z, lam, xsi, eta, mu, zet, s, epsi)
        it = 0
        while rmax > 0.9 * epsi and it < 200:
            it += 1
            ux, xl = upp - x, x - low
            plam, qlam = p0 + p.T @ lam, q0 + q.T @ lam
            gvec = p @ (1 / ux) + q @ (1 / xl)
            gg = p / ux[None] ** 2 - q / xl[None] ** 2
            delx = plam / ux ** 2 - qlam / xl ** 2 - epsi / (x - alfa) + epsi / (beta - x)
            dely = c + d * y - lam - epsi / y
            delz = a0 - a @ lam - epsi / z
            dellam = gvec - a * z - y - b + epsi / lam
            diagx = 2 * (plam / ux ** 3 + qlam / xl ** 3) + xsi / (x - alfa) + eta / (beta - x)
            diagy = d + mu / y
            diaglamyi = s / lam + 1.0 / diagy
            blam = dellam + dely / diagy - gg @ (delx / diagx)
            alam = np.diag(diaglamyi) + (gg / diagx[None]) @ gg.T
            aa = np.block([[alam, a[:, None]], [a[None, :], np.array([[-zet / z]])]])
            sol = np.linalg.solve(aa, np.concatenate([blam, [delz]]))
            dlam, dz = sol[:m], sol[m]
            dx = -delx / diagx - (gg.T @ dlam) / diagx
            dy = -dely / diagy + dlam / diagy
            dxsi = -xsi + epsi / (x - alfa) - xsi * dx / (x - alfa)
            deta = -eta + epsi / (beta - x) + eta * dx / (beta - x)
            dmu = -mu + epsi / y - mu * dy / y
            dzet = -zet + epsi / z - zet * dz / z
            ds = -s + epsi / lam - s * dlam / lam
            xx = np.concatenate([y, [z], lam, xsi, eta, mu, [zet], s])
            dxx = np.concatenate([dy, [dz], dlam, dxsi, deta, dmu, [dzet], ds])
            stm = max(np.max(-1.01 * dxx / xx), np.max(-1.01 * dx / (x - alfa)), np.max(1.01 * dx / (beta - x)), 1.0)
            step = 1.0 / stm
            old = (x, y, z, lam, xsi, eta, mu, zet, s)
            inner, new = 0, 2 * norm
            while new > norm and inner < 50:
                inner += 1
                x, y, z, lam, xsi, eta, mu, zet, s = (o + step * dd for o, dd in zip(old, (dx, dy, dz, dlam, dxsi, deta, dmu, dzet, ds)))
                new, rmax = residual(x, y, z, lam, xsi, eta, mu, zet, s, epsi)
                step /= 2
            norm = new
        epsi *= 0.1
    return x, y, z, lam


def mma_step(it, x, xold1, xold2, low, upp, xmin, xmax, f0, df0, g, dg, move=0.1, scale=100.0,
             a0=1.0, c=1e4, d=1.0, asyinit=0.5, asydecr=0.7, asyincr=1.2, raa0=1e-5):
    """One MMA update; returns (x_new, low, upp) like mmasub's (xmma, ..., low, upp)."""
    g = np.atleast_1d(np.asarray(g, float)) * scale
    dg = np.atleast_2d(np.asarray(dg, float)) * scale
    df0 = np.asarray(df0, float) * scale
    m = len(g)
    low, upp = asymptotes(it, x, xold1, xold2, low, upp, xmin, xmax, asyinit, asydecr, asyincr)
    alfa, beta, p0, q0, p, q, b = approximation(x, low, upp, xmin, xmax, df0, g, dg, move, raa0)
    xn, _, _, lam = subsolve(low, upp, alfa, beta, p0, q0, p, q, a0, np.zeros(m), b, np.full(m, c), np.full(m, d))
    return xn, low, upp, lam, (alfa, beta, p0, q0, p, q, b)


def dual_subsolve(low, upp, alfa, beta, p0, q0, p, q, b, c, d, tol=1e-14):
    """Independent check for m = 1, a = 0: maximise the dual over lambda >= 0.

    For fixed lambda the separable primal minimiser is closed form:
    x_j = clip((sqrt(P) L + sqrt(Q) U) / (sqrt(P) + sqrt(Q)), alpha, beta) with
    P = p0 + lambda p, Q = q0 + lambda q; y = max(0, (lambda - c) / d).
    The dual derivative g(x(lambda)) - y(lambda) - b is decreasing; bisect on it.
    """
    p, q = p[0], q[0]

    def primal(lmb):
        sp, sq = np.sqrt(p0 + lmb * p), np.sqrt(q0 + lmb * q)
        x = np.clip((sp * low + sq * upp) / (sp + sq), alfa, beta)
        y = max(0.0, (lmb - c[0]) / d[0])
        return x, y, np.sum(p / (upp - x) + q / (x - low)) - y - b[0]

    if primal(0.0)[2] <= 0:
        return primal(0.0)[0], 0.0
    lo, hi = 0.0, 1.0
    while primal(hi)[2] > 0:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if primal(mid)[2] > 0:
            lo = mid
        else:
            hi = mid
        if hi - lo <= tol * hi:
            break
    return primal(0.5 * (lo + hi))[0], 0.5 * (lo + hi)
