"""Pins of the CPU restatement of the w2e stage (oracle/ci_oracle.py) that
do not restate its formulas: an independent packer for the V_12 triangle,
eigen-residuals, orthonormality, the 2 x 2 closed form, and a diagonal
Hamiltonian whose CI dipoles equal the product-basis ones.  Also: the CI
library loads without a GPU, exports its header, and refuses to compute."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle import ci_oracle as co

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pack_upper(H):
    """Row-major enumeration of the upper triangle, i.e. w2e's V_12."""
    return H[np.triu_indices(H.shape[0])]


@pytest.mark.parametrize("n", [1, 2, 5, 33])
def test_unpack_matches_independent_packer(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    H = A + A.T
    assert np.array_equal(co.unpack_upper(pack_upper(H), n), H)


@pytest.mark.parametrize("n", [1, 3, 40, 120])
def test_basis_is_an_eigendecomposition(n):
    rng = np.random.default_rng(10 + n)
    ens = np.sort(rng.uniform(-2.9, 5, n))
    A = rng.standard_normal((n, n)) * 0.1
    H = A + A.T
    w, V = co.ci_basis(ens, pack_upper(H))
    Hf = H + np.diag(ens)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(V @ V.T - np.eye(n)).max() < 1e-12
    assert np.abs(Hf @ V.T - V.T * w).max() < 1e-11 * max(1, np.abs(w).max())
    for k in range(n):
        i = int(np.argmax(np.abs(V[k])))
        assert V[k, i] > 0


def test_two_by_two_closed_form():
    a, b, c = 0.7, -0.3, -1.1
    w, V = co.ci_basis([a, c], [0.0, b, 0.0])
    mid, rad = (a + c) / 2, math.hypot((a - c) / 2, b)
    assert abs(w[0] - (mid - rad)) < 1e-15 and abs(w[1] - (mid + rad)) < 1e-15
    # eigenvector of the lower eigenvalue is proportional to (b, lam - a)
    v = np.array([b, w[0] - a])
    v /= np.linalg.norm(v)
    assert abs(abs(V[0] @ v) - 1.0) < 1e-14


def test_diagonal_hamiltonian_keeps_dipoles():
    rng = np.random.default_rng(4)
    n0, n1 = 6, 4
    e0, e1 = np.sort(rng.uniform(-1, 1, n0)), np.sort(rng.uniform(-1, 1, n1))
    _, V0 = co.ci_basis(e0, np.zeros(n0 * (n0 + 1) // 2))
    _, V1 = co.ci_basis(e1, np.zeros(n1 * (n1 + 1) // 2))
    assert np.array_equal(V0, np.eye(n0)) and np.array_equal(V1, np.eye(n1))
    D = rng.standard_normal((n0, n1))
    assert np.array_equal(co.ci_dipoles(D, V0, V1), D)


def test_dipoles_transform_as_operators():
    """CI_dip = V_L^T D V_{L+1}: a dipole operator between the two blocks
    keeps its singular values under the orthogonal change of basis."""
    rng = np.random.default_rng(5)
    n0, n1 = 9, 7
    A0, A1 = rng.standard_normal((n0, n0)), rng.standard_normal((n1, n1))
    _, V0 = co.ci_basis(np.zeros(n0), pack_upper(A0 + A0.T))
    _, V1 = co.ci_basis(np.zeros(n1), pack_upper(A1 + A1.T))
    D = rng.standard_normal((n0, n1))
    Dci = co.ci_dipoles(D, V0, V1)
    assert np.allclose(np.linalg.svd(Dci, compute_uv=False),
                       np.linalg.svd(D, compute_uv=False), rtol=1e-12)
    assert np.allclose(V0.T @ Dci @ V1, D, atol=1e-12)


def test_ci_library_exports_and_refuses_without_gpu():
    from mpci import ci
    txt = open(os.path.join(ROOT, "include", "mpci_ci.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(mpci_(?:ci|d1e)_[a-z0-9_]+)\s*\(", txt)))
    assert names == sorted(ci.EXPORTED)
    lib = ctypes.CDLL(ci.CI_LIB)
    for n in names:
        assert hasattr(lib, n), n
    import mpci
    if mpci.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(Exception) as e:
        ci.ci_basis([0.0, 1.0], [0.0, 0.1, 0.0])
    assert "GPU" in str(e.value)


def hydrogen_on_quadrature(q=200, R=60.0):
    """Gauss-Legendre nodes and weights on (0, R); P_10 = 2 r e^-r and
    P_21 = r^2 e^(-r/2) / (2 sqrt 6) with P_21'."""
    x, w = np.polynomial.legendre.leggauss(q)
    r, wr = 0.5 * R * (x + 1), 0.5 * R * w
    p10 = 2 * r * np.exp(-r)
    p21 = r * r * np.exp(-r / 2) / (2 * math.sqrt(6))
    dp21 = (2 * r - r * r / 2) * np.exp(-r / 2) / (2 * math.sqrt(6))
    return r, wr, p10[None, :], p21[None, :], dp21[None, :]


def test_d1e_hydrogen_1s_2p_both_gauges():
    r, w, p10, p21, dp21 = hydrogen_on_quadrature()
    k0 = 1 / math.sqrt(3)           # k_l of l = 0
    radial = 128 * math.sqrt(6) / 243  # <1s| r |2p> radial integral
    Dl = co.d1e_dipoles(0, "l", r, w, p10, p21)
    assert abs(Dl[0, 0] - k0 * radial) < 1e-12
    # velocity form: <1s| d/dr + (l+1)/r |2p> = (E_2p - E_1s) <1s| r |2p>
    Dv = co.d1e_dipoles(0, "v", r, w, p10, p21, dp21)
    assert abs(Dv[0, 0] - k0 * 0.375 * radial) < 1e-12
