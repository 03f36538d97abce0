"""Pins the oracle's RHS (MatVecV / MatVecL, src/tdse/tdse.cpp:14-143)
against a dense numpy operator built from the physics, not from the dgemv
call pattern:

  velocity gauge:  dc/dt = -i E c + F K c,  K[L,L+1] = -D_L, K[L+1,L] = D_L^T
  length gauge:    dc/dt = -i (E + F S) c,  S[L,L+1] = D_L,  S[L+1,L] = D_L^T

plus the invariant that makes the flow unitary: Re<x, f(x)> = 0 for any x
(K is real antisymmetric, S is real symmetric).
"""
import numpy as np
import pytest

SHAPES = [[3, 2, 4], [1, 1], [5, 1], [2, 3, 1, 4], [7], [4, 6, 5, 3, 2]]


def make_sys(oracle, s, gauge, seed):
    rng = np.random.default_rng(seed)
    E = [np.sort(rng.uniform(-2.9, 5.0, n)) for n in s]
    D = [rng.standard_normal((s[L], s[L + 1])) for L in range(len(s) - 1)]
    return oracle.System(E, D, gauge=gauge), E, D


def dense_operator(E, D, gauge, F):
    s = [len(e) for e in E]
    off = np.concatenate([[0], np.cumsum(s)])
    N = off[-1]
    C = np.zeros((N, N))
    for L, d in enumerate(D):
        a, b, c = off[L], off[L + 1], off[L + 2]
        if gauge == "v":
            C[a:b, b:c] = -d
            C[b:c, a:b] = d.T
        else:
            C[a:b, b:c] = d
            C[b:c, a:b] = d.T
    Ed = np.diag(np.concatenate(E)).astype(complex)
    if gauge == "v":
        return -1j * Ed + F * C
    return -1j * (Ed + F * C)


def to_c(x):
    return x[0::2] + 1j * x[1::2]


@pytest.mark.parametrize("gauge", ["v", "l"])
@pytest.mark.parametrize("s", SHAPES)
def test_rhs_matches_dense_operator(oracle, gauge, s):
    sys_, E, D = make_sys(oracle, s, gauge, seed=sum(s))
    rng = np.random.default_rng(7)
    x = rng.standard_normal(2 * sys_.N)
    F = 0.37
    M = dense_operator(E, D, gauge, F)
    np.testing.assert_allclose(to_c(sys_.rhs(F, x)), M @ to_c(x), rtol=0,
                               atol=1e-13)


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_coupling_is_unit_field_minus_diagonal(oracle, gauge):
    sys_, E, D = make_sys(oracle, [4, 3, 5], gauge, seed=3)
    x = np.random.default_rng(2).standard_normal(2 * sys_.N)
    M1 = dense_operator(E, D, gauge, 1.0) - dense_operator(E, D, gauge, 0.0)
    np.testing.assert_allclose(to_c(sys_.coupling(x)), M1 @ to_c(x), atol=1e-13)


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_flow_is_norm_preserving(oracle, gauge):
    sys_, _, _ = make_sys(oracle, [6, 5, 7, 4], gauge, seed=11)
    x = np.random.default_rng(5).standard_normal(2 * sys_.N)
    assert abs(np.dot(x, sys_.rhs(0.8, x))) < 1e-12


def test_zero_field_is_diagonal(oracle):
    sys_, E, _ = make_sys(oracle, [3, 4], "v", seed=1)
    x = np.random.default_rng(3).standard_normal(2 * sys_.N)
    np.testing.assert_allclose(to_c(sys_.rhs(0.0, x)),
                               -1j * np.concatenate(E) * to_c(x), atol=1e-15)
