"""The allocation-free RHS entry point, mpci_system_rhs_noalloc, against
the allocating one it is a variant of, mpci_prop_rhs (one pulse, the
large-system path: the same kernels, so the results are bitwise equal), and
against the oracle's RHS (MatVecV / MatVecL, src/tdse/tdse.cpp:14-143) to
1e-12 relative.  Two calls in a row on one device, the second on a smaller
system, so the second runs over the scratch the first left behind.  Input
and output live in device buffers the test allocates itself through the
HIP runtime the library is linked to."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    import mpci
    from mpci import _lib
    assert mpci.device_count() >= 1
    _lib.tdse_lib()  # binds the runtime the engine uses
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DevArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip = hip
        self.n = host.nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), self.n) == 0
        assert hip.hipMemcpy(self.p, host.ctypes.data, self.n, H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def noalloc_rhs(hip, sysg, F, x):
    from mpci import _lib
    xd = DevArray(hip, x)
    out = DevArray(hip, np.full_like(x, np.nan))
    try:
        rc = _lib.tdse_lib().mpci_system_rhs_noalloc(sysg.h, F, xd.p, out.p,
                                                     None)
        _lib.check(rc)
        return out.get(x)
    finally:
        xd.free()
        out.free()


def test_noalloc_rhs_matches_prop_rhs(hip, oracle):
    import mpci
    from mpci import synth
    cases = (([900, 1100, 700, 1300], "v", 0.031, 11),
             ([300, 257, 129], "l", -0.02, 12))
    for s, gauge, F, seed in cases:
        print(f"case s={s} gauge={gauge} F={F}", flush=True)
        E, D = synth.system_arrays(s, seed=seed)
        sysg = mpci.System(E, D, gauge=gauge)
        rng = np.random.default_rng(seed)
        x = rng.standard_normal(2 * sysg.N)
        x /= np.linalg.norm(x)
        got = noalloc_rhs(hip, sysg, F, x)
        prop = mpci.Propagator(sysg)
        prop.set_small_path(0)
        ref = prop.rhs([F], x)
        prop.close()
        want = oracle.System(E, D, gauge=gauge).rhs(F, x)
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"  max|noalloc - prop_rhs| = {np.abs(got - ref).max():.3e}, "
              f"relative to the oracle {err:.3e}", flush=True)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
        assert err < 1e-12


def test_noalloc_rejects_oversized_system(hip):
    import mpci
    from mpci import _lib, synth
    # 2 x 7500: 59 row tiles and 30 column tiles of partials over 7500 rows
    # need 10.68 MB of scratch, beyond the 10 MiB the library holds
    E, D = synth.system_arrays([7500, 7500], seed=5)
    sysg = mpci.System(E, D, gauge="v")
    x = np.zeros(2 * sysg.N)
    x[0] = 1.0
    xd = DevArray(hip, x)
    out = DevArray(hip, x)
    try:
        rc = _lib.tdse_lib().mpci_system_rhs_noalloc(sysg.h, 0.01, xd.p,
                                                     out.p, None)
        assert rc == _lib.MPCI_ERR_ARG
        msg = _lib.tdse_lib().mpci_last_error().decode()
        assert "scratch" in msg and "10485760" in msg
    finally:
        xd.free()
        out.free()
