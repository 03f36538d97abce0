"""CPU-side checks of the C ABI: the built library loads without a GPU and
exports every function include/mpci_tdse.h declares; the host-side pulse
helpers agree with the oracle's restatement of src/tdse/pulse.cpp; the
product refuses to run without a GPU instead of falling back to the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "mpci_tdse.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpci_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    from mpci import _lib
    lib = ctypes.CDLL(_lib.TDSE_LIB)
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


def test_pulses_match_oracle(oracle):
    import mpci
    for gauge in "vl":
        a = mpci.Pulse.sine(0.0009, 0.7, 10, gauge=gauge, cepd=0.1)
        b = oracle.Pulse.sine(0.0009, 0.7, 10, gauge=gauge, cepd=0.1)
        g1 = mpci.Pulse.gauss(0.0009, 0.7, 89.76, 10, gauge=gauge)
        g2 = oracle.Pulse.gauss(0.0009, 0.7, 89.76, 10, gauge=gauge)
        env = np.linspace(0, 0.04, 9)
        p1 = mpci.Pulse.pwl(env, 0.7, 50.0, gauge=gauge)
        p2 = oracle.Pulse.pwl(env, 0.7, 50.0, gauge=gauge)
        for t in np.linspace(-1, 95, 301):
            assert a(t) == b(t)
            assert g1(t) == g2(t)
            assert p1(t) == p2(t)
    assert mpci.to_au(1e14, 15.0) == oracle.to_au(1e14, 15.0)


def test_no_gpu_fails_loudly():
    import mpci
    if mpci.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(mpci.MpciError) as ei:
        mpci.System([np.zeros(2), np.zeros(3)], [np.zeros((2, 3))])
    assert ei.value.code == -2 and "GPU" in str(ei.value)


def test_synthetic_generator_is_seeded():
    from mpci import synth
    E1, D1 = synth.system_arrays([50, 40, 30], seed=99)
    E2, D2 = synth.system_arrays([50, 40, 30], seed=99)
    E3, _ = synth.system_arrays([50, 40, 30], seed=98)
    assert all(np.array_equal(a, b) for a, b in zip(E1, E2))
    assert all(np.array_equal(a, b) for a, b in zip(D1, D2))
    assert not np.array_equal(E1[0], E3[0])
    for e in E1:
        assert np.all(np.diff(e) >= 0) and e.min() >= -2.9 and e.max() < 5.0
    d = synth.system_arrays([400, 400], seed=1)[1][0]
    assert abs(d.std() * np.sqrt(400) - 1.0) < 0.01 and abs(d.mean()) < 2e-3


def test_system_create_refuses_bad_arguments():
    """mpci_system_create checks its arguments before any device call: a
    block count outside 1..32, an unknown gauge, an empty block, a NULL
    energy or dipole block, and a state beyond 2^30 coefficients are each
    refused with MPCI_ERR_ARG and a message, and no system is returned.
    The pointers are never read on these paths."""
    from mpci import _lib
    lib = ctypes.CDLL(_lib.TDSE_LIB)
    f = lib.mpci_system_create
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_char,
                  ctypes.POINTER(ctypes.c_void_p),
                  ctypes.POINTER(ctypes.c_void_p),
                  ctypes.POINTER(ctypes.c_void_p)]
    lib.mpci_last_error.restype = ctypes.c_char_p
    buf = (ctypes.c_double * 4)()
    ok = ctypes.addressof(buf)

    def call(s, gauge=b"v", e_null=-1, d_null=-1, nb=None):
        n = len(s) if nb is None else nb
        m = max(len(s), 1)
        S = (ctypes.c_int * m)(*(s or [1]))
        E = (ctypes.c_void_p * m)(*[None if i == e_null else ok
                                    for i in range(m)])
        D = (ctypes.c_void_p * m)(*[None if i == d_null else ok
                                    for i in range(m)])
        out = ctypes.c_void_p()
        rc = f(n, S, gauge, E, D, ctypes.byref(out))
        return rc, lib.mpci_last_error().decode(), out.value

    cases = [call([], nb=0), call([4] * 33), call([4, 4], gauge=b"x"),
             call([4, 0, 4]), call([4, 4], e_null=1), call([4, 4], d_null=0),
             call([(1 << 29) + 1, (1 << 29) + 1])]
    for rc, msg, out in cases:
        assert rc == -1 and msg and out is None, (rc, msg)
    assert "too large" in cases[-1][1]


def test_new_entries_refuse_bad_arguments_before_device_calls():
    """mpci_system_rhs_noalloc refuses a NULL system or state, and
    mpci_multi_adaptive_grouped refuses an empty group, more than eight
    members and NULL arrays, each with MPCI_ERR_ARG and a message, before
    any pointer is read or the device touched."""
    from mpci import _lib
    lib = ctypes.CDLL(_lib.TDSE_LIB)
    lib.mpci_last_error.restype = ctypes.c_char_p
    na = lib.mpci_system_rhs_noalloc
    na.restype = ctypes.c_int
    na.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p]
    buf = (ctypes.c_double * 4)()
    ok = ctypes.addressof(buf)
    for args in [(None, 0.01, ok, ok, None), (ok, 0.01, None, ok, None),
                 (ok, 0.01, ok, None, None)]:
        assert na(*args) == -1
        assert b"NULL" in lib.mpci_last_error()
    g = lib.mpci_multi_adaptive_grouped
    g.restype = ctypes.c_int
    g.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 4 + [
        ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    arr = (ctypes.c_void_p * 9)(*([None] * 9))
    a = ctypes.addressof(arr)
    for n, p in [(0, a), (9, a), (2, None)]:
        assert g(n, p, a, a, a, 1e-10, 1e-10, 10, None) == -1
        assert b"1 to 8 members" in lib.mpci_last_error()
    assert g(2, a, a, a, a, 1e-10, 1e-10, 10, None) == -1
    assert b"NULL member" in lib.mpci_last_error()
