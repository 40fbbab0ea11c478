"""Shared test configurations and independent quadrature helpers."""
import numpy as np
import scipy.special as sp

import lippmann_schwinger_solver_amd as L


def inputs(depth=3, kappa=10.0, q=0, seed=None, nbumps=5, angle=0.0, find_phi=1, tol_pow=10, amp_scale=1.0):
    """Uniform-tree inputs; q=6 draws `nbumps` seeded Gaussian bumps."""
    inp = L.InputsToDFMM(kappa=kappa, nLevelsUniform=depth, Qchoice=q, find_phi=find_phi, TOL_POW=tol_pow,
                         inc_angle=angle)
    if q == 6:
        c = L.bumps_from_seed(0 if seed is None else seed, nbumps)
        for i in range(c.n_bumps):
            c.amp[i] *= amp_scale
        inp.contrast = c
    return inp


def rand_vec(n, seed=0):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ------------------------------------------------------------------------
# Independent quadrature (numpy/scipy) used to pin the oracle's tables.
def cheb_nodes(n):
    return -np.cos((np.arange(n) + 0.5) / n * np.pi)


def basis_ab(deg):
    return [(a, d - a) for d in range(deg) for a in range(d + 1)]


def cheb(x, k):
    """T_k(x) as a polynomial (also defined outside [-1, 1])."""
    c = np.zeros(k + 1)
    c[k] = 1.0
    return np.polynomial.chebyshev.chebval(x, c)


def vandermonde(n, deg):
    x1 = cheb_nodes(n)
    X = np.tile(x1, n)          # index j*n+k -> x = node[k]
    Y = np.repeat(x1, n)        # y = node[j]
    Q = np.stack([cheb(X, a) * cheb(Y, b) for a, b in basis_ab(deg)], axis=1)
    return Q, X, Y


def far_table_indep(n, deg, kappa, r, ng=48):
    """M = C pinv(Q), C(n,l) = r^2 int J_n(kappa r|P|) e^{-in theta} T_a T_b over [-1,1]^2."""
    Q, _, _ = vandermonde(n, deg)
    g, w = np.polynomial.legendre.leggauss(ng)
    PX, PY = np.meshgrid(g, g, indexing="ij")
    W = np.outer(w, w)
    rho, th = np.hypot(PX, PY), np.arctan2(PY, PX)
    ab = basis_ab(deg)
    C = np.zeros((31, len(ab)), dtype=complex)
    for i, m in enumerate(range(-15, 16)):
        base = sp.jv(m, kappa * r * rho) * np.exp(-1j * m * th)
        for l, (a, b) in enumerate(ab):
            C[i, l] = r * r * np.sum(W * base * cheb(PX, a) * cheb(PY, b))
    return C @ np.linalg.pinv(Q)


def fan_quad(f, X, nphi=20, nr=20, grade=0.25, levels=14):
    """int_{[-1,1]^2} f(P) dP by a polar fan around X: every edge of the square is
    cut into segments graded geometrically toward the foot of the perpendicular
    from X, each segment spans a signed triangle with apex X, integrated in
    polar coordinates with r = R t^2 (exact for the log singularity of H0 at X)."""
    g, w = np.polynomial.legendre.leggauss(nphi)
    gr, wr = np.polynomial.legendre.leggauss(nr)
    t, wt = 0.5 * (gr + 1), 0.5 * wr
    corners = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)
    total = 0j
    for e in range(4):
        A, B = corners[e], corners[(e + 1) % 4]
        d = B - A
        nrm = np.array([d[1], -d[0]]) / np.hypot(*d)
        h = np.dot(A - X, nrm)
        if abs(h) < 1e-300:
            continue
        s0 = np.clip(np.dot(X - A, d) / np.dot(d, d), 0.0, 1.0)
        cuts = {0.0, 1.0, s0}
        for k in range(levels):
            off = grade ** k
            cuts.update({min(1.0, s0 + off), max(0.0, s0 - off)})
        cuts = np.array(sorted(cuts))
        for sa, sb in zip(cuts[:-1], cuts[1:]):
            if sb - sa < 1e-16:
                continue
            Pa, Pb = A + sa * d, A + sb * d
            a0 = np.arctan2(*(Pa - X)[::-1])
            a1 = np.arctan2(*(Pb - X)[::-1])
            da = (a1 - a0 + np.pi) % (2 * np.pi) - np.pi
            phis = a0 + 0.5 * (g + 1) * da
            for phi, wphi in zip(phis, w * 0.5 * da):
                u = np.array([np.cos(phi), np.sin(phi)])
                R = h / np.dot(u, nrm)
                r = R * t * t
                P = X[None, :] + r[:, None] * u[None, :]
                total += wphi * np.sum(wt * f(P[:, 0], P[:, 1]) * r * 2 * R * t)
    return total


def near_entry_indep(n, deg, kappa, r, slot_center, i, j, rB=None):
    """Near-field operator entry T(i, j): target node i of the box of radius r at
    the origin, source node j of the neighbour of radius rB (default r) centred
    at slot_center * r."""
    rB = r if rB is None else rB
    Q, X, Y = vandermonde(n, deg)
    P = np.linalg.pinv(Q)
    Xi = np.array([(r * X[i] - slot_center[0] * r) / rB, (r * Y[i] - slot_center[1] * r) / rB])
    val = 0j
    for l, (a, b) in enumerate(basis_ab(deg)):
        f = lambda px, py: sp.hankel1(0, kappa * rB * np.hypot(Xi[0] - px, Xi[1] - py)) * cheb(px, a) * cheb(py, b)
        val += fan_quad(f, Xi) * P[l, j]
    return 1j * rB * rB / 4.0 * val


def near_entry_gl(n, deg, kappa, r, slot_center, i, j, rB=None, ng=64):
    """near_entry_indep by a tensor Gauss-Legendre rule over the source box, for
    sources well separated from the target node (smooth integrand)."""
    rB = r if rB is None else rB
    Q, X, Y = vandermonde(n, deg)
    P = np.linalg.pinv(Q)
    Xi = np.array([(r * X[i] - slot_center[0] * r) / rB, (r * Y[i] - slot_center[1] * r) / rB])
    g, w = np.polynomial.legendre.leggauss(ng)
    PX, PY = np.meshgrid(g, g, indexing="ij")
    W = np.outer(w, w)
    Hk = sp.hankel1(0, kappa * rB * np.hypot(Xi[0] - PX, Xi[1] - PY))
    val = sum(np.sum(W * Hk * cheb(PX, a) * cheb(PY, b)) * P[l, j] for l, (a, b) in enumerate(basis_ab(deg)))
    return 1j * rB * rB / 4.0 * val
