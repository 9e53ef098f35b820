"""Pins the quadrature rules and the Lagrange basis (oracle and library tables) to closed forms.

Reference: element/GaussQuadrature.h:46-99, element/LagrangeShapeFuncs.h:7-252,
element/LagrangePolyBenchmark.cpp:49-121 (hand-factored order-2/3 forms),
element/lagrangePolyCoeffs.py:6-18 (the coefficient derivation).
"""

import math

import numpy as np
import pytest

import textbook

ORDERS = [1, 2, 3, 4]


def closed_form_rule(n):
    """Gauss-Legendre points/weights in closed form for n = 2..5 points."""
    if n == 2:
        x = [-1 / math.sqrt(3), 1 / math.sqrt(3)]
        w = [1.0, 1.0]
    elif n == 3:
        x = [-math.sqrt(3 / 5), 0.0, math.sqrt(3 / 5)]
        w = [5 / 9, 8 / 9, 5 / 9]
    elif n == 4:
        a = math.sqrt(3 / 7 - 2 / 7 * math.sqrt(6 / 5))
        b = math.sqrt(3 / 7 + 2 / 7 * math.sqrt(6 / 5))
        wa = (18 + math.sqrt(30)) / 36
        wb = (18 - math.sqrt(30)) / 36
        x = [-b, -a, a, b]
        w = [wb, wa, wa, wb]
    else:
        a = math.sqrt(5 - 2 * math.sqrt(10 / 7)) / 3
        b = math.sqrt(5 + 2 * math.sqrt(10 / 7)) / 3
        wa = (322 + 13 * math.sqrt(70)) / 900
        wb = (322 - 13 * math.sqrt(70)) / 900
        x = [-b, -a, 0.0, a, b]
        w = [wb, wa, 128 / 225, wa, wb]
    return np.array(x), np.array(w)


@pytest.mark.parametrize("order", ORDERS)
def test_oracle_gauss_closed_form(orc, order):
    xs, ws = closed_form_rule(order + 1)
    got = np.array([orc.gauss(order, i) for i in range(order + 1)])
    # The reference's constants carry 15 significant digits.
    np.testing.assert_allclose(got[:, 0], xs, rtol=0, atol=1e-15)
    np.testing.assert_allclose(got[:, 1], ws, rtol=0, atol=1e-15)


@pytest.mark.parametrize("order", ORDERS)
def test_gauss_exact_for_degree_2n_minus_1(orc, order):
    n = order + 1
    pts = [orc.gauss(order, i) for i in range(n)]
    for k in range(2 * n):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        got = sum(w * x**k for x, w in pts)
        assert abs(got - exact) < 2e-15, (order, k)
    # ... and not for degree 2n (a rule with a wrong point count fails this).
    exact = 2.0 / (2 * n + 1)
    assert abs(sum(w * x ** (2 * n) for x, w in pts) - exact) > 1e-6


@pytest.mark.parametrize("order", ORDERS)
def test_library_tables_match_oracle_bitwise(fem, orc, order):
    phi, dphi, xg, wg = fem.basis_table(order)
    for i in range(order + 1):
        x, w = orc.gauss(order, i)
        assert xg[i] == x and wg[i] == w
        for k in range(order + 1):
            N, dN = orc.lagrange1d(order, x, k)
            assert phi[k, i] == N and dphi[k, i] == dN, (order, k, i)


@pytest.mark.parametrize("order", ORDERS)
def test_basis_kronecker_at_cosine_knots(orc, order):
    knots = textbook.cosine_knots(order)
    for j, xj in enumerate(knots):
        for k in range(order + 1):
            N, _ = orc.lagrange1d(order, xj, k)
            # order-4 coefficients use sqrt(2) ~ 1.41421356237310 (15 digits)
            assert abs(N - (1.0 if j == k else 0.0)) < 1e-13, (order, j, k, N)


@pytest.mark.parametrize("order", ORDERS)
def test_basis_partition_of_unity_and_linear_reproduction(orc, order):
    knots = textbook.cosine_knots(order)
    for x in np.linspace(-1, 1, 17):
        vals = [orc.lagrange1d(order, x, k) for k in range(order + 1)]
        N = np.array([v[0] for v in vals])
        dN = np.array([v[1] for v in vals])
        assert abs(N.sum() - 1) < 1e-13
        assert abs(dN.sum()) < 1e-12
        assert abs((knots * N).sum() - x) < 1e-13
        assert abs((knots * dN).sum() - 1) < 1e-12


@pytest.mark.parametrize("order", ORDERS)
def test_basis_matches_independent_product_formula(orc, order):
    x = np.linspace(-1, 1, 23)
    N_ref, dN_ref = textbook.basis_1d(order, x)
    for k in range(order + 1):
        got = np.array([orc.lagrange1d(order, xi, k) for xi in x])
        np.testing.assert_allclose(got[:, 0], N_ref[k], atol=5e-14)
        np.testing.assert_allclose(got[:, 1], dN_ref[k], atol=5e-13)


def _hand_factored(order, x, k):
    """LagrangePolyBenchmark.cpp:49-121 restated: factored order-2/3 bases and derivatives."""
    if order == 2:
        N = [-0.5 * x * (1 - x), (1 - x) * (1 + x), 0.5 * (1 + x) * x][k]
        dN = [-0.5 + x, -2 * x, 0.5 + x][k]
    else:
        N = [
            -(2 / 3) * (0.5 + x) * (0.5 - x) * (1 - x),
            (4 / 3) * (1 + x) * (0.5 - x) * (1 - x),
            (4 / 3) * (1 + x) * (0.5 + x) * (1 - x),
            -(2 / 3) * (1 + x) * (0.5 + x) * (0.5 - x),
        ][k]
        dN = [
            -2 * x * x + (4 / 3) * x + 1 / 6,
            4 * x * x - (4 / 3) * x - 4 / 3,
            -4 * x * x - (4 / 3) * x + 4 / 3,
            2 * x * x + (4 / 3) * x - 1 / 6,
        ][k]
    return N, dN


@pytest.mark.parametrize("order", [2, 3])
def test_basis_matches_hand_factored_forms(orc, order):
    for x in np.linspace(-1, 1, 13):
        for k in range(order + 1):
            N, dN = orc.lagrange1d(order, x, k)
            Nh, dNh = _hand_factored(order, x, k)
            assert abs(N - Nh) < 1e-14 and abs(dN - dNh) < 1e-13


@pytest.mark.parametrize("order", ORDERS)
def test_monomial_coefficients_independently_derived(orc, order):
    # Recover monomial coefficients from the oracle by interpolation at order+1 points and compare with
    # the expansion of the product formula (lagrangePolyCoeffs.py:6-18 derivation, restated in numpy).
    xs = np.linspace(-0.9, 0.9, order + 1)
    V = np.vander(xs, order + 1, increasing=True)
    polys = textbook.lagrange_polys(order)
    for k in range(order + 1):
        vals = np.array([orc.lagrange1d(order, x, k)[0] for x in xs])
        coeffs = np.linalg.solve(V, vals)
        np.testing.assert_allclose(coeffs, polys[k], atol=1e-12)
