"""CPU checks of the learned preconditioner's numpy reference (oracle/learned_ref.py):
its backward against central finite differences, its BF16 rounding against
known values, and the Chebyshev Lagrange basis against polynomial exactness."""
import numpy as np

from oracle import learned_ref as R


def test_backward_matches_finite_differences():
    lay = R.Layout(3, 4, (1, 2, 1))
    rng = np.random.default_rng(0)
    p = rng.standard_normal(lay.nparams) * 0.3
    x = rng.standard_normal((2, 8, 8, 3))
    t = rng.standard_normal((2, 8, 8, 2))
    loss, g = R.loss_grad(lay, p, x, t, quantize=False)
    h = 1e-6
    idx = rng.choice(lay.nparams, 40, replace=False)
    fd = np.array([(R.loss_grad(lay, p + h * np.eye(lay.nparams)[i], x, t, quantize=False)[0] -
                    R.loss_grad(lay, p - h * np.eye(lay.nparams)[i], x, t, quantize=False)[0]) / (2 * h) for i in idx])
    assert np.max(np.abs(fd - g[idx])) <= 1e-6 * max(1.0, np.max(np.abs(fd)))
    # every layer's weights and biases are covered by the sample
    assert loss > 0


def test_conv_matches_direct_sum():
    """conv() against a literal triple loop on a tiny case (dilation 2, zero padding)."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 5, 5, 2))
    w = rng.standard_normal((3, 3, 3, 2))
    b = rng.standard_normal(3)
    y = R.conv(x, w, b, 2)
    ref = np.zeros((1, 5, 5, 3))
    for yy in range(5):
        for xx in range(5):
            for co in range(3):
                s = b[co]
                for ky in range(3):
                    for kx in range(3):
                        sy, sx = yy + (ky - 1) * 2, xx + (kx - 1) * 2
                        if 0 <= sy < 5 and 0 <= sx < 5:
                            s += np.dot(w[co, ky, kx], x[0, sy, sx])
                ref[0, yy, xx, co] = s
    assert np.allclose(y, ref, atol=1e-12)


def test_bf16_rounding():
    """BF16 keeps 8 significant bits: spacing 2^-7 just above 1."""
    v = np.array([1.0, 1 + 2 ** -7, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 5 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = R.bf16(v)
    expect = [1.0, 1 + 2 ** -7, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5]   # ties go to the even mantissa
    assert np.array_equal(r[:6], np.array(expect, dtype=np.float32))
    assert abs(r[6] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_cheb_lagrange_is_exact_for_polynomials():
    nc = 6
    tn = -np.cos((np.arange(nc) + 0.5) * np.pi / nc)
    t = np.linspace(-1, 1, 13)
    W = R.cheb_lagrange(nc, t)
    for deg in range(nc):
        assert np.allclose(W @ tn ** deg, t ** deg, atol=1e-12)
    assert np.allclose(W.sum(axis=1), 1.0)
