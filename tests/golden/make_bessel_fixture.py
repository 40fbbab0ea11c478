"""Write tests/golden/bessel_jy.npz: J_n(x), Y_n(x), n = 0..15, from scipy.special
(cross-checked against mpmath at 30 digits), at arguments covering the ranges the
operator build evaluates: near-field H0 arguments (small x), leaf-level Graf
arguments and the series / recurrence / asymptotic switch points (1, 25)."""
import os

import mpmath
import numpy as np
import scipy.special as sp

xs = np.array([1e-9, 1e-4, 0.01, 0.3, 0.7, 0.999, 1.0, 1.001, 1.9, 2.5, 4.0, 7.3, 10.0, 12.5, 15.0, 19.9,
               24.99, 25.0, 25.01, 31.4, 40.0, 63.2, 80.0, 113.0, 160.0, 226.0, 300.0])
n = np.arange(16)
J = sp.jv(n[None, :], xs[:, None])
Y = sp.yv(n[None, :], xs[:, None])
mpmath.mp.dps = 30
for i, x in enumerate(xs):
    for k in n:
        jm = float(mpmath.besselj(int(k), mpmath.mpf(x)))
        ym = float(mpmath.bessely(int(k), mpmath.mpf(x)))
        assert abs(J[i, k] - jm) <= 1e-13 * max(1.0, abs(jm)), (x, k)
        if np.isfinite(ym) and abs(ym) < 1e250:
            assert abs(Y[i, k] - ym) <= 1e-13 * max(1.0, abs(ym)), (x, k)
            Y[i, k] = ym
        J[i, k] = jm
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bessel_jy.npz")
np.savez(out, x=xs, J=J, Y=Y)
print(out)
