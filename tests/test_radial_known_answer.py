"""Known answer for the whole solver chain: scattering by the radial Gaussian.

The reference's Gaussian contrast (Qchoice 0, q = 1.5 exp(-160 r^2),
GMRES/GMRES.cpp:5-8) is radial, so its Lippmann-Schwinger solution follows
from one ODE per Fourier mode (tests/radial.py). Three checks pin that exact
solution:
- q0 = 0 gives the plane wave back;
- a real contrast gives unitary mode coefficients;
- the small-contrast limit is the closed-form Born amplitude of a Gaussian.

The oracle is then checked against the exact solution:
- Phi = -k^2 q u on the uniform trees converges with depth to the
  discretisation floor;
- the README's adaptive run (GMRES/Readme.md:43-69) lands at its discretisation
  level; its tables and neighbour slots come from quadrature, not from the
  exact solution.
"""
import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H
from tests import radial as R

README = dict(nCones_LFR=16, nChebNodes=6, treeAdaptivity=5, L=0.5, kappa=40.0, yes2DFMM=1, degreeOfBases=6,
              TOL_POW=8, m=10, restart=1, nLevelsUniform=5, Qchoice=0, tree_mode=1)


def born_pattern(kappa, q0, theta, a=160.0):
    """Far-field pattern F of u_s ~ F(theta) sqrt(2/(pi k r)) e^{i(kr - pi/4)} in
    the Born approximation: u_s = V(-k^2 q e^{ikx}), and the Fourier transform
    of q0 exp(-a |y|^2) at k(e_x - x_hat) is q0 (pi/a) exp(-k^2 (1 - cos theta)/(2a))."""
    return (1j / 4) * (-kappa ** 2) * q0 * (np.pi / a) * np.exp(-kappa ** 2 * (1 - np.cos(theta)) / (2 * a))


def pattern(an, theta):
    # H_n(kr) ~ sqrt(2/(pi k r)) e^{i(kr - n pi/2 - pi/4)}, a_{-n} (-i)^{-n} = a_n (-i)^n
    return an[0] + 2 * sum(an[m] * (-1j) ** m * np.cos(m * theta) for m in range(1, an.size))


def test_exact_field_reproduces_plane_wave():
    g = np.linspace(-0.5, 0.5, 31)
    X, Y = np.meshgrid(g, g)
    for kappa in (10.0, 40.0):
        u = R.total_field(kappa, X.ravel(), Y.ravel(), q0=0.0)
        err = np.abs(u - np.exp(1j * kappa * X.ravel())).max()
        print(f"kappa {kappa}: q0 = 0 field vs e^(ikx) {err:.1e}")
        assert err < 1e-9


def test_exact_coefficients_are_unitary():
    """Real q conserves energy: u_n is proportional to H2_n + S_n H1_n outside,
    with S_n = 1 + 2 a_n / i^n, and |S_n| = 1."""
    for kappa in (10.0, 40.0):
        an = R.scattering_coefficients(kappa)
        S = 1 + 2 * an / (1j ** np.arange(an.size))
        print(f"kappa {kappa}: max ||S_n| - 1| = {np.abs(np.abs(S) - 1).max():.1e} over {an.size} modes")
        assert np.abs(np.abs(S) - 1).max() < 1e-12
        assert np.abs(S - 1).max() > 0.1          # the contrast scatters


def test_exact_field_has_the_born_limit():
    """The first-order term in q0 equals the closed-form Born amplitude, which
    fixes the sign of q and the (i/4) H0 scale; the remainder is second order
    (shrinks tenfold with q0)."""
    th = np.linspace(0, 2 * np.pi, 13)
    for kappa in (10.0, 40.0):
        d = []
        for q0 in (1e-3, 1e-4):
            F = pattern(R.scattering_coefficients(kappa, q0=q0, rtol=1e-13), th)
            FB = born_pattern(kappa, q0, th)
            d.append(np.abs(F - FB).max() / np.abs(FB).max())
        print(f"kappa {kappa}: Born relative difference {d[0]:.2e} (q0 1e-3), {d[1]:.2e} (q0 1e-4)")
        assert d[1] < 2e-4
        assert 8 < d[0] / d[1] < 12


def exact_phi(kappa, x, y):
    u = R.total_field(kappa, x, y)
    return -kappa ** 2 * R.contrast(x, y) * u, u


def test_oracle_converges_to_the_exact_solution():
    """Uniform trees at kappa 10: Phi from the oracle's GMRES, and the total field
    u = u_inc + V Phi through the oracle's findPhi = 0 operator, against the
    exact field. Depth 3 to 4 gains two orders (6.1e-5 to 2.6e-7)."""
    errs = {}
    for depth in (3, 4):
        o = O.Oracle(H.inputs(depth=depth, kappa=10.0).to_c())
        x, y = o.grid()
        b = o.rhs()
        phi, _ = o.gmres(b, 4, 50)
        assert H.relerr(o.matvec(phi), b) < 1e-12
        pe, ue = exact_phi(10.0, x, y)
        ov = O.Oracle(H.inputs(depth=depth, kappa=10.0, find_phi=0).to_c())
        u = np.exp(10j * x) + ov.matvec(phi)
        errs[depth] = (H.relerr(phi, pe), H.relerr(u, ue))
        print(f"depth {depth}, N {o.N}: Phi vs exact {errs[depth][0]:.2e}, u vs exact {errs[depth][1]:.2e}")
        o.close()
        ov.close()
    assert errs[3][0] < 1.2e-4 and errs[4][0] < 6e-7
    assert errs[3][0] / errs[4][0] > 50
    assert errs[3][1] < 3e-5 and errs[4][1] < 6e-6


def test_readme_run_against_the_exact_solution():
    """The README's adaptive run (kappa 40, 148 leaves on levels 2-5), as
    measured:
    - GMRES(2,10) Phi is 1.6e-5 from the exact one;
    - the exact Phi has residual 7.5e-6 in the dense operator (getMatrixEntry2:
      tables for touching leaves, Graf series otherwise) and 1.2e-5 in the
      DAFMM;
    - u through the dense V is 8e-6 from the exact field.
    Through the compressed V at TOL_POW 8, u is 8e-4 off: the compression
    error of V alone."""
    inp = L.InputsToDFMM(**README)
    o = O.Oracle(inp.to_c())
    x, y = o.grid()
    b = o.rhs()
    pe, ue = exact_phi(40.0, x, y)
    phi, _ = o.gmres(b, 2, 10)
    ov = O.Oracle(L.InputsToDFMM(**dict(README, find_phi=0)).to_c())
    e_phi = H.relerr(phi, pe)
    r_dense, r_fmm = H.relerr(o.dense_matvec(pe), b), H.relerr(o.matvec(pe), b)
    e_u = H.relerr(np.exp(40j * x) + ov.dense_matvec(phi), ue)
    print(f"N {o.N}: Phi vs exact {e_phi:.2e}; exact-Phi residual dense {r_dense:.2e}, DAFMM {r_fmm:.2e}; "
          f"u (dense V) vs exact {e_u:.2e}")
    assert e_phi < 3e-5
    assert r_dense < 1.5e-5 and r_fmm < 2.5e-5
    assert e_u < 1.6e-5
    o.close()
    ov.close()
