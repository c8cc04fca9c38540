"""Pins of the oracle's ``gradient`` and diffusion operator (CPU).

The reference has no test of ``cardiax.solve.gradient`` (cardiax/solve.py:
225-254); these checks pin it by what the mathematics fixes: exactness on
polynomials of the formulas' orders, the rational boundary tables of the
composite operators (derived by hand from the reference's rules and typed here
as known answers), symmetry, and the undamped checkerboard mode.
"""
from fractions import Fraction as F

import numpy as np
import pytest

from oracle import fk_oracle as o


def _poly(x, coeffs):
    return sum(c * x ** k for k, c in enumerate(coeffs))


def _dpoly(x, coeffs):
    return sum(k * c * x ** (k - 1) for k, c in enumerate(coeffs) if k > 0)


@pytest.mark.parametrize("n", [5, 6, 9, 17])
def test_exact_on_cubics_everywhere(n):
    x = np.arange(n, dtype=np.float64)
    c = [0.3, -1.7, 0.25, 0.125]
    a = np.tile(_poly(x, c)[:, None], (1, 3))
    g = o.gradient(a, 0)[:, 1]
    np.testing.assert_allclose(g, _dpoly(x, c), rtol=0, atol=1e-9)


def test_exact_on_quartics_in_the_interior_only():
    n = 12
    x = np.arange(n, dtype=np.float64)
    c = [0.0, 0.0, 0.0, 0.0, 0.01]
    g = o.gradient(np.tile(_poly(x, c)[None, :], (3, 1)), 1)[1]
    exact = _dpoly(x, c)
    np.testing.assert_allclose(g[2:-2], exact[2:-2], atol=1e-9)
    # the one-sided edge rules are 3rd order: a quartic exposes them
    assert abs(g[0] - exact[0]) > 1e-3 and abs(g[-1] - exact[-1]) > 1e-3


def test_constant_has_zero_derivative_and_rows_sum_to_zero():
    a = np.full((7, 9), 3.25)
    # zero up to the rounding of the float coefficients (-11/6 + 3 - 3/2 + 1/3)
    np.testing.assert_allclose(o.gradient(a, 0), 0, atol=1e-14)
    np.testing.assert_allclose(o.gradient(a, 1), 0, atol=1e-14)
    A1, A2 = o.effective_operators(11)
    np.testing.assert_allclose(A1.sum(1), 0, atol=1e-13)
    np.testing.assert_allclose(A2.sum(1), 0, atol=1e-13)


def test_reflection_antisymmetry():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(10, 13))
    np.testing.assert_allclose(o.gradient(a[::-1], 0), -o.gradient(a, 0)[::-1], atol=1e-13)
    np.testing.assert_allclose(o.gradient(a[:, ::-1], 1), -o.gradient(a, 1)[:, ::-1],
                               atol=1e-13)


def test_axis_selection():
    # a function of the column index only has no derivative along axis 0
    a = np.tile(np.arange(8.0) ** 2, (6, 1))
    np.testing.assert_allclose(o.gradient(a, 0), 0, atol=1e-13)
    assert np.abs(o.gradient(a, 1)).max() > 1


# Effective coefficients of crop(G(G(pad_edge(u)))) on original cells,
# derived by hand with exact rationals from the reference's rules
# (one-sided -11/6, 3, -3/2, 1/3 on the two outer padded entries; central
# 1/12, -2/3, 2/3, -1/12 elsewhere; edge padding by one cell).
D2_TABLE = [
    [F(107, 72), F(-161, 36), F(163, 36), F(-67, 36), F(25, 72), F(-1, 36)],
    [F(11, 8), F(-371, 144), F(13, 12), F(2, 9), F(-1, 9), F(1, 144)],
    [F(17, 72), F(11, 36), F(-49, 48), F(5, 36), F(4, 9), F(-1, 9), F(1, 144)],
    [F(-5, 48), F(4, 9), F(1, 9), F(-65, 72), F(1, 9), F(4, 9), F(-1, 9), F(1, 144)],
]
D2_INTERIOR = [F(1, 144), F(-1, 9), F(4, 9), F(1, 9), F(-65, 72), F(1, 9), F(4, 9), F(-1, 9),
               F(1, 144)]
D1_TABLE = [
    [F(-11, 6), F(3), F(-3, 2), F(1, 3)],
    [F(-7, 12), F(0), F(2, 3), F(-1, 12)],
]
D1_INTERIOR = [F(1, 12), F(-2, 3), F(0), F(2, 3), F(-1, 12)]


def test_composite_boundary_tables():
    n = 14
    A1, A2 = o.effective_operators(n)
    for r, row in enumerate(D2_TABLE):
        expect = np.zeros(n)
        expect[:len(row)] = [float(x) for x in row]
        np.testing.assert_allclose(A2[r], expect, atol=1e-13, err_msg="row %d" % r)
        # far edge mirrors: D2[N-1-r][N-1-j] = D2[r][j]
        np.testing.assert_allclose(A2[n - 1 - r][::-1], expect, atol=1e-13)
    for r in range(4, n - 4):
        expect = np.zeros(n)
        expect[r - 4:r + 5] = [float(x) for x in D2_INTERIOR]
        np.testing.assert_allclose(A2[r], expect, atol=1e-13)
    for r, row in enumerate(D1_TABLE):
        expect = np.zeros(n)
        expect[:len(row)] = [float(x) for x in row]
        np.testing.assert_allclose(A1[r], expect, atol=1e-13)
        np.testing.assert_allclose(A1[n - 1 - r][::-1], -expect, atol=1e-13)
    for r in range(2, n - 2):
        expect = np.zeros(n)
        expect[r - 2:r + 3] = [float(x) for x in D1_INTERIOR]
        np.testing.assert_allclose(A1[r], expect, atol=1e-13)


def test_ghost_row_first_derivative():
    # The padded ghost entry's first derivative feeds row 1's second derivative:
    # (7/6 u0 - 3/2 u1 + 1/3 u2).
    u = np.array([0.7, -0.2, 1.3, 0.4, 2.0, -1.0])
    g = o.gradient(np.pad(np.tile(u[:, None], (1, 5)), 1, mode="edge"), 0)[0, 2]
    assert g == pytest.approx(7 / 6 * u[0] - 3 / 2 * u[1] + 1 / 3 * u[2], abs=1e-14)


def test_checkerboard_is_undamped_in_the_interior():
    # interior symbol -f(theta)^2 with f(pi) = 0
    n = 24
    u = np.fromfunction(lambda i, j: (-1.0) ** (i + j), (n, n))
    D = np.ones((n, n))
    lap = o.diffusion(np.pad(u, 1, mode="edge"), np.pad(D, 1, mode="edge"), 1.0)[1:-1, 1:-1]
    np.testing.assert_allclose(lap[4:-4, 4:-4], 0, atol=1e-12)
    assert np.abs(lap[0]).max() > 1  # the boundary rows do act on it


def test_diffusion_linear_zero_on_uniform_and_uniform_D_has_no_drift():
    rng = np.random.default_rng(2)
    n, m = 12, 15
    u1, u2 = rng.normal(size=(n, m)), rng.normal(size=(n, m))
    D = rng.uniform(0.5, 1.5, size=(n, m))
    L = lambda u: o.diffusion(np.pad(u, 1, mode="edge"), np.pad(D, 1, mode="edge"), 0.1)
    np.testing.assert_allclose(L(2 * u1 - 3 * u2), 2 * L(u1) - 3 * L(u2), atol=1e-9)
    np.testing.assert_allclose(L(np.full((n, m), 0.37)), 0, atol=1e-12)
    Dp = np.pad(np.full((n, m), 0.8), 1, mode="edge")
    np.testing.assert_allclose(o.gradient(Dp, 0), 0, atol=1e-14)
    np.testing.assert_allclose(o.gradient(Dp, 1), 0, atol=1e-14)


def test_stability_bound_of_the_interior_symbol():
    # max over theta of f^2 with f = 4/3 sin - 1/6 sin 2theta is 1.883..., so
    # explicit Euler on D * (-f^2)/dx^2 per axis is stable iff
    # D dt / dx^2 < 2 / (2 * 1.883) = 0.531.
    th = np.linspace(0, np.pi, 200001)
    f2 = (4 / 3 * np.sin(th) - 1 / 6 * np.sin(2 * th)) ** 2
    assert f2.max() == pytest.approx(1.883, abs=1e-3)
    assert 1 / f2.max() == pytest.approx(0.531, abs=1e-3)
