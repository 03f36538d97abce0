"""CPU restatement of MP-CITDSE's w2e stage (CI two-electron basis) and of
its d1e stage (one-electron dipole blocks).

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may use it, as the oracle the GPU path
(lib/libmpci_ci.so) is compared with, never as the product.

Follows src/w2e/w2e.cpp:17-125:
  * formCIh0 (w2e.cpp:17-82): H = diag(e_2e) + V_12, V_12 read as the packed
    upper triangle (row i holds columns i..n-1, element (i, j) at
    (2n - i - 1) i / 2 + j, w2e.cpp:45-52), then LAPACKE_dsyevd (numpy's
    eigh here: the same LAPACK syevd family, eigenvalues ascending).
  * formCIDipoles (w2e.cpp:84-125): CI_dip = V_L^T d_if V_{L+1}, formed as
    (V_L^T d_if) then times V_{L+1} like the two cblas_dgemm calls.
Eigenvector signs: LAPACK leaves them to the implementation, so the
reference has no canonical sign; this restatement applies the GPU library's
convention (largest-magnitude component positive, first index on a tie) so
the two can be compared element by element.

d1e follows src/d1e/dmx1e.cpp:24-139 and src/d1e/integrator.cpp:3-31: for
every pair (n1, n2) the quadrature sum of tlenGL (qw x P1 P2) or tvelGL
(qw P1 (P2' - v2 P2 / x)), times k_l = sqrt((l+1)^2 / (4 (l+1)^2 - 1)).

Pins (tests/test_ci_oracle.py): an independent packer (numpy.triu_indices)
for the triangle layout; eigen-residual, orthonormality and ascending order;
the closed form of the 2 x 2 symmetric eigenproblem; a diagonal H whose
eigenvectors are the unit vectors, for which CI_dip equals d_if exactly;
the hydrogen 1s-2p radial dipole in both gauges (128 sqrt(6) / 243 and
3/8 of it, the velocity form being (E_2p - E_1s) times the length form).
"""
from __future__ import annotations

import numpy as np


def unpack_upper(v12, n):
    """Full symmetric matrix from the packed upper triangle (w2e.cpp:45-52)."""
    v12 = np.asarray(v12, dtype=np.float64)
    H = np.empty((n, n))
    for i in range(n):
        row = v12[(2 * n - i - 1) * i // 2 + i:(2 * n - i - 1) * i // 2 + n]
        H[i, i:] = row
        H[i:, i] = row
    return H


def fix_signs(vecs):
    """Largest-magnitude component of each eigenvector (row) positive; the
    first index wins a tie."""
    vecs = np.array(vecs, dtype=np.float64, copy=True)
    for k in range(vecs.shape[0]):
        i = int(np.argmax(np.abs(vecs[k])))  # argmax: first index on ties
        if vecs[k, i] < 0:
            vecs[k] = -vecs[k]
    return vecs


def ci_basis(ens, v12):
    """formCIh0 for one L: (eigenvalues ascending, eigenvectors as rows)."""
    ens = np.asarray(ens, dtype=np.float64)
    n = len(ens)
    H = unpack_upper(v12, n) + np.diag(ens)
    w, V = np.linalg.eigh(H, UPLO="L")
    return w, fix_signs(V.T)


def ci_dipoles(D, vecs0, vecs1):
    """formCIDipoles for one pair: (V_L^T D) V_{L+1} in the reference's
    order; the eigenvector matrices hold eigenvectors as rows."""
    temp = np.asarray(vecs0) @ np.asarray(D)
    return temp @ np.asarray(vecs1).T


def d1e_dipoles(l, gauge, qx, qw, wfn_l, wfn_l1, wfnp_l1=None):
    """dmx1e::genDipole for one l: D[n1][n2] (nen x nen)."""
    qx, qw = np.asarray(qx, dtype=np.float64), np.asarray(qw, dtype=np.float64)
    w0, w1 = np.asarray(wfn_l), np.asarray(wfn_l1)
    nen = w0.shape[0]
    lp1 = l + 1
    kl = np.sqrt((lp1 * lp1) / (4.0 * lp1 * lp1 - 1))
    v2 = (l * (l + 1) - lp1 * (lp1 + 1)) * 0.5
    D = np.empty((nen, nen))
    for n1 in range(nen):
        for n2 in range(nen):
            if gauge == "l":  # tlenGL
                t = np.sum(qw * qx * w0[n1] * w1[n2])
            else:             # tvelGL
                t = np.sum(qw * (w0[n1] * wfnp_l1[n2] - v2 * w0[n1] * w1[n2]
                                 / qx))
            D[n1, n2] = kl * t
    return D
