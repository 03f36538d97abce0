"""Pins the oracle's RKF78 stepper (restating boost::odeint
runge_kutta_fehlberg78::do_step as called at src/tdse/tdse.cpp:190,273).

The pins are mathematical, not a retyped formula:
  * the oracle's own tableau satisfies every Runge-Kutta order condition
    (all 200 rooted trees up to order 8 for b, up to order 7 for b-hat) and
    b-hat fails at least one order-8 condition (else the error estimate
    would vanish);
  * a step on a diagonal system equals the stability polynomial
    R(z) = 1 + z b^T (I - zA)^{-1} 1, and R(z) - e^z = O(z^9);
  * a step on a constant-field two-level system equals R(hM) and tends to
    expm(hM) with an error ratio of about 2^9 per halving of h (local) and
    2^8 (global).
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg


def rooted_trees(max_order):
    """All rooted trees up to max_order as canonical nested tuples."""
    def canon(t):
        return tuple(sorted(canon(c) for c in t))

    def add_leaf_everywhere(t):
        out = {canon(t + ((),))}
        for i, c in enumerate(t):
            for c2 in add_leaf_everywhere(c):
                out.add(canon(t[:i] + (c2,) + t[i + 1:]))
        return out

    levels = {1: {()}}
    for n in range(2, max_order + 1):
        s = set()
        for t in levels[n - 1]:
            s |= add_leaf_everywhere(t)
        levels[n] = s
    return levels


def order(t):
    return 1 + sum(order(c) for c in t)


def gamma(t):
    g = order(t)
    for c in t:
        g *= gamma(c)
    return g


def elementary_weights(t, A):
    """Phi_i(t) for every stage i (vector)."""
    phi = np.ones(A.shape[0], dtype=object)
    for c in t:
        phi = phi * A.dot(elementary_weights(c, A))
    return phi


def _rational(x, maxden=10_000):
    f = Fraction(x).limit_denominator(maxden)
    assert abs(float(f) - x) < 1e-15 * max(1.0, abs(x)), x
    return f


@pytest.fixture(scope="module")
def tab(oracle):
    a, b, db, c = oracle.tableau()
    A = np.zeros((13, 13))
    A[:, :12] = a
    return A, b, b - db, c


def test_tree_counts():
    lv = rooted_trees(8)
    assert [len(lv[n]) for n in range(1, 9)] == [1, 1, 2, 4, 9, 20, 48, 115]


def test_explicit_and_row_sums(tab):
    A, b, bh, c = tab
    assert np.all(np.triu(A) == 0.0), "tableau must be explicit"
    np.testing.assert_allclose(A.sum(axis=1), c, atol=1e-15)


def test_order_conditions_exact(tab):
    """Exact-rational order conditions on the oracle's tableau."""
    A, b, bh, c = tab
    Aq = np.array([[_rational(v) for v in row] for row in A], dtype=object)
    bq = np.array([_rational(v) for v in b], dtype=object)
    bhq = np.array([_rational(v) for v in bh], dtype=object)
    levels = rooted_trees(8)
    failures_bh_8 = 0
    for n in range(1, 9):
        for t in levels[n]:
            phi = elementary_weights(t, Aq)
            target = Fraction(1, gamma(t))
            assert sum(bq * phi) == target, ("b", n, t)
            if n <= 7:
                assert sum(bhq * phi) == target, ("bhat", n, t)
            elif sum(bhq * phi) != target:
                failures_bh_8 += 1
    assert failures_bh_8 > 0


def stability_poly(A, b, z):
    n = A.shape[0]
    return 1.0 + z * b.dot(np.linalg.solve(np.eye(n) - z * A, np.ones(n)))


def test_diagonal_system_matches_R(oracle, tab):
    A, b, bh, c = tab
    E = [np.array([-2.5, -0.3, 0.7, 4.9]), np.array([1.1, -1.9])]
    D = [np.zeros((4, 2))]
    sys_ = oracle.System(E, D, gauge="v")
    rng = np.random.default_rng(1)
    x = rng.standard_normal(2 * sys_.N)
    h = 0.17
    y = sys_.step(0.3, x, h)  # zero dipoles: the field has no effect
    xc = x[0::2] + 1j * x[1::2]
    yc = y[0::2] + 1j * y[1::2]
    Ecat = np.concatenate(E)
    R = np.array([stability_poly(A.astype(complex), b, -1j * e * h)
                  for e in Ecat])
    np.testing.assert_allclose(yc, R * xc, rtol=0, atol=1e-14)
    # R(z) - exp(z) = O(z^9)
    for zabs in (0.05, 0.1):
        z = -1j * zabs
        assert abs(stability_poly(A.astype(complex), b, z) - np.exp(z)) < 1e-3 * zabs ** 9


def two_level(oracle, gauge, E0=-0.9, E1=0.4, d=0.8):
    return oracle.System([np.array([E0]), np.array([E1])], [np.array([[d]])],
                         gauge=gauge)


def two_level_matrix(gauge, F, E0=-0.9, E1=0.4, d=0.8):
    if gauge == "v":   # dc/dt = -iE c + F (D^T c_{L-1} - D c_{L+1})
        return np.array([[-1j * E0, -F * d], [F * d, -1j * E1]])
    return -1j * np.array([[E0, F * d], [F * d, E1]])


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_two_level_step_is_R_of_hM(oracle, tab, gauge):
    A, b, bh, c = tab
    F, h = 0.35, 0.3
    sys_ = two_level(oracle, gauge)
    M = two_level_matrix(gauge, F)
    x = np.array([0.6, -0.2, 0.3, 0.7])
    y = sys_.step(F, x, h)
    # R(hM) = sum_k (hM)^k b^T A^{k-1} 1
    Rm = np.eye(2, dtype=complex)
    Ak1 = np.ones(13)
    P = np.eye(2, dtype=complex)
    for k in range(1, 14):
        P = P @ (h * M)
        Rm = Rm + P * b.dot(Ak1)
        Ak1 = A.dot(Ak1)
    xc = x[0::2] + 1j * x[1::2]
    yc = y[0::2] + 1j * y[1::2]
    np.testing.assert_allclose(yc, Rm @ xc, atol=1e-15)
    np.testing.assert_allclose(yc, scipy.linalg.expm(h * M) @ xc, atol=1e-9)


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_global_convergence_order_8(oracle, gauge):
    F = 0.5
    sys_ = two_level(oracle, gauge, E0=-2.0, E1=3.0, d=1.3)
    M = two_level_matrix(gauge, F, E0=-2.0, E1=3.0, d=1.3)
    x0 = np.array([1.0, 0.0, 0.0, 0.0])
    T = 4.0
    exact = scipy.linalg.expm(T * M) @ np.array([1.0, 0.0])
    errs = []
    for n in (20, 40):
        h = T / n
        x = x0.copy()
        for _ in range(n):
            x = sys_.step(F, x, h)
        errs.append(np.abs((x[0::2] + 1j * x[1::2]) - exact).max())
    ratio = errs[0] / errs[1]
    assert 2 ** 7.5 < ratio < 2 ** 8.5, (errs, ratio)


def test_embedded_error_is_b_minus_bhat(oracle, tab):
    """err returned by the stage step equals h (R8 - R7) on a diagonal system."""
    A, b, bh, c = tab
    E = [np.array([0.9, -1.7])]
    sys_ = oracle.System(E, [], gauge="v")
    x = np.array([1.0, 0.5, -0.25, 0.75])
    h = 0.4
    _, err, k1 = sys_.step_stages(np.zeros(13), x, h)
    xc = x[0::2] + 1j * x[1::2]
    ec = err[0::2] + 1j * err[1::2]
    Ac = A.astype(complex)
    for i, e in enumerate(E[0]):
        z = -1j * e * h
        R8 = stability_poly(Ac, b, z)
        R7 = stability_poly(Ac, bh, z)
        assert abs(ec[i] - (R8 - R7) * xc[i]) < 1e-15
    # k1 is the derivative at the start of the step
    np.testing.assert_allclose(k1[0::2] + 1j * k1[1::2], -1j * E[0] * xc, atol=1e-15)
