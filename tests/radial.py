"""Exact scattering of a plane wave by a radial contrast, for known-answer tests.

The Lippmann-Schwinger system of GMRES/GMRES.cpp (A = I + diag(k^2 q) V with
the kernel (i/4) H0(k|x-y|), b = -k^2 q u_inc, u = u_inc + V Phi) is the
integral form of

    Laplace(u) + k^2 (1 - q) u = 0,   u - u_inc outgoing,

so for q = q0 exp(-a r^2) (Qchoice 0, GMRES/GMRES.cpp:5-8) the total field
splits into Fourier modes u = sum_n u_n(r) e^{i n theta}. Each mode solves

    u_n'' + u_n'/r + (k^2 (1 - q(r)) - n^2/r^2) u_n = 0,

regular at r = 0. Outside the scatterer it equals i^n J_n(kr) + a_n H_n(kr)
(Jacobi-Anger for e^{ikx}). Each mode is integrated outward from its Frobenius
series and matched to that form at the outermost radius. The solution has no
part in common with the solver: no tree, no tables, no Graf series. With
q0 = 0 it reproduces the plane wave to about 1e-11.
"""
import math

import numpy as np
import scipy.special as sp
from scipy.integrate import solve_ivp


def _series_coeffs(kappa, q0, a, n, nterms):
    """b_k of u_n = r^n sum_k b_k r^{2k}. The coefficient k^2 (1 - q(r)) is
    sum_j c_j r^{2j}, with c_j = k^2 (delta_j0 - q0 (-a)^j / j!)."""
    c = [kappa * kappa * ((1.0 if j == 0 else 0.0) - q0 * (-a) ** j / math.factorial(j)) for j in range(nterms)]
    b = [1.0]
    for k in range(1, nterms):
        b.append(-sum(c[j] * b[k - 1 - j] for j in range(k)) / (4.0 * k * (n + k)))
    return b


def _modes(kappa, q0, a, R, rtol):
    """Per mode n = 0..nmax: (n, r0, series coefficients, ODE solution of
    w = u_n / (s r0^n), s, a_n), with u_n = i^n J_n(kr) + a_n H_n(kr) for r >= R."""
    kR = kappa * R
    nmax = int(kR) + 1
    while abs(sp.jv(nmax, kR)) > 1e-16:
        nmax += 1
    out = []
    for n in range(nmax + 1):
        b = _series_coeffs(kappa, q0, a, n, 40)
        # start where the series is still accurate: sqrt(a) r <= 1.3 and
        # k r well inside the evanescent zone of mode n
        r0 = min(1.3 / math.sqrt(a), max(1e-3, 0.3 * n / kappa))
        # w = (r / r0)^n sum_k b_k r^{2k}: w(r0) and w'(r0)
        w0 = sum(bk * r0 ** (2 * k) for k, bk in enumerate(b))
        dw0 = n * w0 / r0 + sum(bk * 2 * k * r0 ** (2 * k - 1) for k, bk in enumerate(b) if k)

        def f(t, z, n=n):
            return [z[1], -z[1] / t - (kappa * kappa * (1.0 - q0 * math.exp(-a * t * t)) - n * n / (t * t)) * z[0]]

        sol = solve_ivp(f, (r0, R), [w0, dw0], method="DOP853", rtol=rtol, atol=1e-300, dense_output=True)
        if sol.status != 0:
            raise RuntimeError(f"radial mode {n}: {sol.message}")
        wR, dwR = sol.y[0, -1], sol.y[1, -1]
        M = np.array([[wR, -sp.hankel1(n, kR)], [dwR, -kappa * sp.h1vp(n, kR)]], dtype=complex)
        s, an = np.linalg.solve(M, (1j ** n) * np.array([sp.jv(n, kR), kappa * sp.jvp(n, kR)]))
        out.append((n, r0, b, sol, s, an))
    return out


def scattering_coefficients(kappa, q0=1.5, a=160.0, R=0.5, rtol=1e-11):
    """a_n, n = 0..nmax, of the scattered field sum_n a_n H_n(kr) e^{in theta}
    (a_{-n} = (-1)^n a_n)."""
    return np.array([m[5] for m in _modes(kappa, q0, a, R, rtol)])


def total_field(kappa, x, y, q0=1.5, a=160.0, rtol=1e-11):
    """u(x, y) for the incident wave e^{i kappa x} and q = q0 exp(-a r^2).

    The match radius is max(0.5, max r); q0 exp(-a r^2) is below 1e-17 there
    for the reference's Gaussian."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    r = np.hypot(x, y)
    th = np.arctan2(y, x)
    R = max(float(r.max()), 0.5) * 1.0001
    order = np.argsort(r)
    rs = r[order]
    u = np.zeros(r.size, dtype=complex)
    for n, r0, b, sol, s, _ in _modes(kappa, q0, a, R, rtol):
        w = np.empty(rs.size)
        inner = rs < r0
        w[~inner] = sol.sol(rs[~inner])[0]
        ri = rs[inner]
        w[inner] = (ri / r0) ** n * sum(bk * ri ** (2 * k) for k, bk in enumerate(b))
        un = np.empty(r.size, dtype=complex)
        un[order] = s * w
        u += (1.0 if n == 0 else 2.0) * un * np.cos(n * th)     # u_{-n} = u_n
    return u


def contrast(x, y, q0=1.5, a=160.0):
    return q0 * np.exp(-a * (np.asarray(x) ** 2 + np.asarray(y) ** 2))
