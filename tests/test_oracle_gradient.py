"""Pins of the gradient oracle (CPU): dense K_L products and finite differences."""
import numpy as np
import pytest

from oracle import gradient as OG, semisep
from dwba_amd.grid import exp_grid


def _small(n=48, P=3, L_max=6, seed=3):
    g = exp_grid(n, 1e-3, 50.0)
    rng = np.random.default_rng(seed)
    ra = rng.standard_normal((P, n)) * np.exp(-g.r / 5)
    rb = rng.standard_normal((P, n))
    c = rng.standard_normal((P, L_max + 1))
    return g, ra, rb, c


@pytest.mark.parametrize("L", [0, 1, 4, 9])
def test_kmatvec_equals_dense_kernel_product(L):
    g, ra, rb, _ = _small(n=80)
    x = rb.T
    Y = OG.kmatvec_all_L(g.r, x, L)[L].astype(np.float64)
    K = semisep.k_dense(g.r, L)
    dense = K @ x
    scale = np.abs(K) @ np.abs(x)
    assert np.all(np.abs(Y - dense) <= 1e-13 * scale + 1e-300)


def test_gradient_equals_central_finite_differences():
    # the loss is bilinear in (rho_a, rho_b): a central difference is exact up to rounding
    L_max = 6
    g, ra, rb, c = _small(L_max=L_max)
    ga, gb = OG.density_gradients(g.r, g.w_simpson, ra, rb, c, L_max)

    def loss(a, b):
        return float(np.sum(c * OG.plain_forward(g.r, g.w_simpson, a, b, L_max)))

    rng = np.random.default_rng(11)
    h = 1e-3
    for _ in range(24):
        p = int(rng.integers(0, ra.shape[0]))
        i = int(rng.integers(0, ra.shape[1]))
        for which, grad in ((0, ga), (1, gb)):
            ap, am = ra.copy(), ra.copy()
            bp, bm = rb.copy(), rb.copy()
            if which == 0:
                ap[p, i] += h
                am[p, i] -= h
            else:
                bp[p, i] += h
                bm[p, i] -= h
            fd = (loss(ap, bp) - loss(am, bm)) / (2 * h)
            assert abs(fd - grad[p, i]) <= 1e-9 * max(abs(grad[p, i]), 1e-12) + 1e-11, (which, p, i, fd, grad[p, i])


def test_gradient_detects_swapped_densities():
    # a plausible mistake (K_L applied to w rho_a for the rho_a gradient) fails the FD pin
    L_max = 4
    g, ra, rb, c = _small(L_max=L_max)
    ga, gb = OG.density_gradients(g.r, g.w_simpson, ra, rb, c, L_max)
    wrong_a, _ = OG.density_gradients(g.r, g.w_simpson, rb, ra, c, L_max)
    assert np.allclose(wrong_a, gb)
    assert not np.allclose(wrong_a, ga, rtol=1e-3)


def test_absolute_magnitude_bounds_gradient():
    L_max = 5
    g, ra, rb, c = _small(L_max=L_max)
    ga, gb = OG.density_gradients(g.r, g.w_simpson, ra, rb, c, L_max)
    ma, mb = OG.density_gradients(g.r, g.w_simpson, ra, rb, c, L_max, absolute=True)
    assert np.all(np.abs(ga) <= ma * (1 + 1e-12)) and np.all(np.abs(gb) <= mb * (1 + 1e-12))
