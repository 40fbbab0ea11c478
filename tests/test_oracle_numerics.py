"""Pins of the oracle's numerical primitives and level tables against
independent sources (scipy/mpmath fixtures, numpy, high-order quadrature)."""
import os

import numpy as np
import pytest

import oracle as O
from tests import helpers as H

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bessel_jy.npz")


def test_bessel_against_scipy_mpmath_fixture():
    """besselJ / besselY (G_FMM2DTree.hpp:73-100) for n = 0..15 over the whole
    argument range; negative orders by (-1)^n symmetry."""
    d = np.load(GOLD)
    worst = 0.0
    for i, x in enumerate(d["x"]):
        for n in range(16):
            j, y = O.bessel(n, float(x))
            jr, yr = d["J"][i, n], d["Y"][i, n]
            scale_j = max(abs(jr), 1e-3 / np.sqrt(max(x, 1.0)))   # absolute near zeros of J
            worst = max(worst, abs(j - jr) / scale_j)
            if np.isfinite(yr) and abs(yr) < 1e250:
                scale_y = max(abs(yr), 1e-3 / np.sqrt(max(x, 1.0)))
                assert abs(y - yr) <= 1e-12 * scale_y, (x, n, y, yr)
            jm, ym = O.bessel(-n, float(x))
            assert jm == (-j if n % 2 else j) and ym == (-y if n % 2 else y)
    assert worst < 1e-12, worst


def test_gauss_legendre_10_matches_numpy():
    x, w = O.gauss_legendre(10)
    xr, wr = np.polynomial.legendre.leggauss(10)
    assert np.max(np.abs(x - xr)) < 1e-15 and np.max(np.abs(w - wr)) < 1e-15


@pytest.fixture(scope="module")
def small_oracle():
    return O.Oracle(H.inputs(depth=3, kappa=10.0).to_c())


def test_qpinv_equals_numpy_pinv(small_oracle):
    """Q_pinv (FarFieldInteraction.hpp:138-150) against numpy's SVD pseudo-inverse."""
    Q, _, _ = H.vandermonde(6, 6)
    P = small_oracle.qpinv().reshape(21, 36)
    assert np.max(np.abs(P - np.linalg.pinv(Q))) < 1e-15


def test_far_table_against_independent_quadrature(small_oracle):
    """M = C Q_pinv of the leaf level (depth 3, kappa 10, r = 1/16), 6-digit
    rounded, against a 48x48 Gauss-Legendre evaluation with scipy's J_n."""
    M = small_oracle.far_table().reshape(31, 36)
    Mi = H.far_table_indep(6, 6, 10.0, 0.5 / 8, ng=48)
    err = np.abs(M - Mi)
    assert np.all(err <= 6e-6 * np.abs(Mi) + 2e-11), err.max()


@pytest.mark.parametrize("slot,i,j", [(8, 0, 0), (8, 14, 21), (8, 35, 7), (3, 5, 30), (0, 0, 35), (4, 17, 2)])
def test_near_table_against_polar_quadrature(small_oracle, slot, i, j):
    """Colleague operators (getOperator, G_FMM2DTree.hpp:1773-1805): singular
    self slot and adjacent slots against a polar-fan quadrature of H0^(1)."""
    T = small_oracle.near_tables().reshape(9, 36, 36)
    centers = [(-2, -2), (0, -2), (2, -2), (2, 0), (2, 2), (0, 2), (-2, 2), (-2, 0), (0, 0)]
    ref = H.near_entry_indep(6, 6, 10.0, 0.5 / 8, centers[slot], i, j)
    # 6-digit table rounding (5e-6 relative) plus the reference's adaptive rule,
    # which accepts panels at |I - sum children| < 1e-7 absolute; a wrong basis
    # index, sign, centre or scale is an O(1) error.
    assert abs(T[slot, i, j] - ref) <= 2e-5 * abs(ref), (T[slot, i, j], ref)
