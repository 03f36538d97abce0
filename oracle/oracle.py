"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the oracle.  The product path
(mp-citdse_amd/) never imports this module.

The oracle restates MP-CITDSE's `tdse` propagation path
(src/tdse/tdse.cpp:14-310, src/tdse/pulse.cpp:3-97, src/pes/pes.cpp:151-262);
the header oracle/tdse_oracle.h maps every function to its reference lines.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrSys(C.Structure):
    _fields_ = [("nb", C.c_int), ("s", _ip), ("E", C.POINTER(_dp)),
                ("D", C.POINTER(_dp)), ("gauge", C.c_char)]


class OrPulse(C.Structure):
    _fields_ = [("kind", C.c_int), ("Eo", C.c_double), ("w", C.c_double),
                ("cepd", C.c_double), ("a", C.c_double), ("b", C.c_double),
                ("n_env", C.c_int), ("env", _dp), ("t_end", C.c_double),
                ("scale", C.c_double)]


class OrStats(C.Structure):
    _fields_ = [("accepted", C.c_int), ("rejected", C.c_int),
                ("t", C.c_double), ("h", C.c_double)]


SINE2, GAUSS, PWL = 0, 1, 2

_lib = None


def build(force: bool = False) -> None:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "all", "-C", HERE],
                       check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_field.restype = C.c_double
        L.or_field.argtypes = [C.POINTER(OrPulse), C.c_double]
        L.or_period.restype = C.c_double
        L.or_dznrm2.restype = C.c_double
        L.or_prop_fixed.restype = C.c_double
        L.or_prop_adaptive.restype = C.c_int
        L.or_state_size.restype = C.c_int
        L.or_use_blas.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class System:
    """A block-tridiagonal system as the oracle sees it.

    E: list of nb float64 arrays; D: list of nb-1 row-major (s_L, s_L+1)
    float64 arrays; gauge 'v' or 'l'.
    """

    def __init__(self, E, D, gauge="v"):
        self.E = [np.ascontiguousarray(e, dtype=np.float64) for e in E]
        self.D = [np.ascontiguousarray(d, dtype=np.float64) for d in D]
        self.s = np.array([len(e) for e in self.E], dtype=np.int32)
        for L, d in enumerate(self.D):
            assert d.shape == (self.s[L], self.s[L + 1]), (L, d.shape)
        self.gauge = gauge
        self.N = int(self.s.sum())
        self.offs = np.concatenate([[0], np.cumsum(self.s)]).astype(np.int64)
        nb = len(self.E)
        self._Ep = (_dp * nb)(*[_ptr(e) for e in self.E])
        self._Dp = (_dp * max(nb - 1, 1))(*[_ptr(d) for d in self.D])
        self.c = OrSys(nb, self.s.ctypes.data_as(_ip), self._Ep, self._Dp,
                       gauge.encode())

    # --- RHS -------------------------------------------------------------
    def rhs(self, F: float, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(x)
        lib().or_rhs(C.byref(self.c), C.c_double(F), _ptr(x), _ptr(out))
        return out

    def coupling(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(x)
        lib().or_coupling(C.byref(self.c), _ptr(x), _ptr(out))
        return out

    def step(self, F: float, x: np.ndarray, dt: float) -> np.ndarray:
        y = np.array(x, dtype=np.float64, copy=True)
        work = np.empty(16 * 2 * self.N)
        lib().or_rkf78_step(C.byref(self.c), C.c_double(F), _ptr(y),
                            C.c_double(dt), _ptr(work))
        return y

    def step_stages(self, F13, x, dt):
        """One step with stage fields; returns (x_new, err, k1)."""
        y = np.array(x, dtype=np.float64, copy=True)
        F13 = np.ascontiguousarray(F13, dtype=np.float64)
        work = np.empty(16 * 2 * self.N)
        err = np.empty(2 * self.N)
        k1 = np.empty(2 * self.N)
        lib().or_rkf78_step_stages(C.byref(self.c), _ptr(F13), _ptr(y),
                                   C.c_double(dt), _ptr(work), _ptr(err),
                                   _ptr(k1))
        return y, err, k1

    # --- propagation -----------------------------------------------------
    def prop_fixed(self, pulse: "Pulse", t0, dt, steps, x, track=0):
        x = np.array(x, dtype=np.float64, copy=True)
        pop = np.empty(max(steps, 1))
        fld = np.empty(max(steps, 1))
        t = lib().or_prop_fixed(C.byref(self.c), C.byref(pulse.c),
                                C.c_double(t0), C.c_double(dt), steps, _ptr(x),
                                track, _ptr(pop), _ptr(fld))
        return x, t, pop[:steps], fld[:steps]

    def prop_adaptive(self, pulse, t0, t1, h0, atol, rtol, x,
                      max_attempts=10_000_000, max_log=1_000_000,
                      allow_limit=False):
        x = np.array(x, dtype=np.float64, copy=True)
        st = OrStats()
        hlog = np.empty(max_log)
        rc = lib().or_prop_adaptive(C.byref(self.c), C.byref(pulse.c),
                                    C.c_double(t0), C.c_double(t1),
                                    C.c_double(h0), C.c_double(atol),
                                    C.c_double(rtol), max_attempts, _ptr(x),
                                    C.byref(st), _ptr(hlog), max_log)
        if rc != 0 and not allow_limit:
            raise RuntimeError("oracle adaptive: attempt limit reached")
        return x, st, hlog[:min(st.accepted, max_log)].copy()

    def prop_schedule(self, pulse, t0, hs, x):
        x = np.array(x, dtype=np.float64, copy=True)
        hs = np.ascontiguousarray(hs, dtype=np.float64)
        lib().or_prop_schedule(C.byref(self.c), C.byref(pulse.c),
                               C.c_double(t0), len(hs), _ptr(hs), _ptr(x))
        return x

    def observables(self, x, thr):
        """(ground pop, 2e yield, excitation, norm), pes.cpp:151-262."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(4)
        lib().or_observables(C.byref(self.c), _ptr(x), C.c_double(thr),
                             _ptr(out))
        return out

    def observables_1e(self, x):
        """(ground pop, 1e binned yield, 0, norm), pes.cpp:48-149."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(4)
        lib().or_observables_1e(C.byref(self.c), _ptr(x), _ptr(out))
        return out


class Pulse:
    """pulse::params (src/tdse/include/pulse.hpp:20-26) + PWL extension."""

    def __init__(self):
        self.c = OrPulse()
        self._env = None

    @staticmethod
    def sine(Io_au, w_au, cycles, gauge="v", cepd=0.0):
        p = Pulse()
        fn = lib().or_sine_a_setup if gauge == "v" else lib().or_sine_e_setup
        fn(C.c_double(Io_au), C.c_double(w_au), C.c_double(cepd), cycles,
           C.byref(p.c))
        return p

    @staticmethod
    def gauss(Io_au, w_au, tau, cycles, gauge="v", cepd=0.0):
        p = Pulse()
        fn = lib().or_gauss_a_setup if gauge == "v" else lib().or_gauss_e_setup
        fn(C.c_double(Io_au), C.c_double(w_au), C.c_double(tau), cycles,
           C.c_double(cepd), C.byref(p.c))
        return p

    @staticmethod
    def pwl(env, w_au, t_end, gauge="v", cepd=0.0):
        p = Pulse()
        p._env = np.ascontiguousarray(env, dtype=np.float64)
        p.c.kind = PWL
        p.c.w = w_au
        p.c.cepd = cepd
        p.c.n_env = len(p._env)
        p.c.env = _ptr(p._env)
        p.c.t_end = t_end
        p.c.scale = 1.0 / w_au if gauge == "v" else 1.0
        return p

    def __call__(self, t: float) -> float:
        return lib().or_field(C.byref(self.c), C.c_double(t))


def to_au(Io_W, w_eV):
    a, b = C.c_double(), C.c_double()
    lib().or_to_au(C.c_double(Io_W), C.c_double(w_eV), C.byref(a), C.byref(b))
    return a.value, b.value


def period(w, cycles):
    return lib().or_period(C.c_double(w), cycles)


def tableau():
    a = np.zeros((13, 12))
    b = np.zeros(13)
    db = np.zeros(13)
    c = np.zeros(13)
    lib().or_tableau(_ptr(a), _ptr(b), _ptr(db), _ptr(c))
    return a, b, db, c


def use_blas(path: str) -> bool:
    return lib().or_use_blas(path.encode()) == 0


def use_builtin_dgemv() -> None:
    lib().or_set_dgemv(None)
