"""Python side of the producers of the tdse dipole inputs (the w2e CI basis
and the d1e one-electron dipoles) over the C ABI in include/mpci_ci.h; the
arithmetic runs in lib/libmpci_ci.so on the GPU.

lib/libmpci_ci.so links ROCm's rocSOLVER and rocBLAS.  The torch wheel
carries its own copies of those libraries under the same sonames (another
ROCm release); in a process that imported torch first, they would be bound
instead.  Call this module from a process without torch (the `w2e` program
is one)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

CI_LIB = os.path.join(_lib.LIB_DIR, "libmpci_ci.so")
EXPORTED = ["mpci_ci_last_error", "mpci_ci_basis", "mpci_ci_dipoles",
            "mpci_d1e_dipoles"]
_ci = None


def ci_lib():
    global _ci
    if _ci is None:
        if not os.path.exists(CI_LIB):
            raise RuntimeError(f"{CI_LIB} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; "
                               "g.build()'`")
        lib = C.CDLL(CI_LIB)
        lib.mpci_ci_last_error.restype = C.c_char_p
        lib.mpci_ci_last_error.argtypes = []
        lib.mpci_ci_basis.restype = C.c_int
        lib.mpci_ci_basis.argtypes = [C.c_int] + [_lib._dp] * 4
        lib.mpci_ci_dipoles.restype = C.c_int
        lib.mpci_ci_dipoles.argtypes = [C.c_int, C.c_int] + [_lib._dp] * 4
        lib.mpci_d1e_dipoles.restype = C.c_int
        lib.mpci_d1e_dipoles.argtypes = ([C.c_int, C.c_int, C.c_int, C.c_char]
                                         + [_lib._dp] * 6)
        _ci = lib
    return _ci


def _check(rc):
    if rc != 0:
        msg = ci_lib().mpci_ci_last_error()
        raise _lib.MpciError(rc, msg.decode() if msg else "")


def ci_basis(ens, v12_packed):
    """Eigenvalues (ascending) and eigenvectors (rows) of diag(ens) + V12,
    V12 given as the packed upper triangle of w2e's V_12 dataset."""
    ens = np.ascontiguousarray(ens, dtype=np.float64)
    v12 = np.ascontiguousarray(v12_packed, dtype=np.float64)
    n = len(ens)
    if v12.shape != (n * (n + 1) // 2,):
        raise ValueError("V12 must hold n(n+1)/2 values")
    eig = np.empty(n)
    vecs = np.empty((n, n))
    _check(ci_lib().mpci_ci_basis(n, _lib.dptr(ens), _lib.dptr(v12),
                                  _lib.dptr(eig), _lib.dptr(vecs)))
    return eig, vecs


def ci_dipoles(D, vecs0, vecs1):
    """V_L^T D V_{L+1} (row-major), eigenvector matrices with eigenvectors
    as rows."""
    D = np.ascontiguousarray(D, dtype=np.float64)
    n0, n1 = D.shape
    v0 = np.ascontiguousarray(vecs0, dtype=np.float64)
    v1 = np.ascontiguousarray(vecs1, dtype=np.float64)
    if v0.shape != (n0, n0) or v1.shape != (n1, n1):
        raise ValueError("eigenvector matrices must match the dipole block")
    out = np.empty((n0, n1))
    _check(ci_lib().mpci_ci_dipoles(n0, n1, _lib.dptr(D), _lib.dptr(v0),
                                    _lib.dptr(v1), _lib.dptr(out)))
    return out


def d1e_dipoles(l, gauge, qx, qw, wfn_l, wfn_l1, wfnp_l1=None):
    """One-electron dipole block between l and l + 1 (the d1e stage):
    nen x nen from nen x qsz samples of the radial functions (and, for the
    velocity gauge, of the derivatives of the l + 1 functions)."""
    qx = np.ascontiguousarray(qx, dtype=np.float64)
    qw = np.ascontiguousarray(qw, dtype=np.float64)
    w0 = np.ascontiguousarray(wfn_l, dtype=np.float64)
    w1 = np.ascontiguousarray(wfn_l1, dtype=np.float64)
    nen, qsz = w0.shape
    if w1.shape != (nen, qsz) or qx.shape != (qsz,) or qw.shape != (qsz,):
        raise ValueError("function samples must be nen x qsz")
    wp = None
    if gauge == "v":
        wp = np.ascontiguousarray(wfnp_l1, dtype=np.float64)
        if wp.shape != (nen, qsz):
            raise ValueError("derivative samples must be nen x qsz")
    out = np.empty((nen, nen))
    _check(ci_lib().mpci_d1e_dipoles(nen, qsz, l, gauge.encode(),
                                     _lib.dptr(qx), _lib.dptr(qw),
                                     _lib.dptr(w0), _lib.dptr(w1),
                                     _lib.dptr(wp) if wp is not None else None,
                                     _lib.dptr(out)))
    return out
