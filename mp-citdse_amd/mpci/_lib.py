"""ctypes binding of the C ABI in include/mpci_tdse.h.

The HIP library (lib/libmpci_tdse.so) is the only compute path: there is no
CPU fallback.  Loading fails loudly when it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_ROOT, "lib")
TDSE_LIB = os.path.join(LIB_DIR, "libmpci_tdse.so")
HOST_LIB = os.path.join(LIB_DIR, "libmpci_host.so")

MPCI_OK = 0
MPCI_ERR_ARG = -1
MPCI_ERR_HIP = -2
MPCI_ERR_STATE = -3
MPCI_ERR_LIMIT = -4
MPCI_ERR_NOTSUP = -5

PULSE_SINE2, PULSE_GAUSS, PULSE_PWL = 0, 1, 2

# rows of the observable weights (MPCI_OBS_*)
OBS_GROUND, OBS_YIELD, OBS_EXCITED, OBS_NORM = 0, 1, 2, 3
OBS_ROWS = 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class MpciPulse(C.Structure):
    _fields_ = [("kind", C.c_int), ("Eo", C.c_double), ("w", C.c_double),
                ("cepd", C.c_double), ("a", C.c_double), ("b", C.c_double),
                ("t_end", C.c_double), ("scale", C.c_double),
                ("n_env", C.c_int), ("env", _dp)]


class MpciStats(C.Structure):
    _fields_ = [("accepted", C.c_int), ("rejected", C.c_int),
                ("attempts", C.c_int), ("status", C.c_int),
                ("t", C.c_double), ("h", C.c_double)]


class MpciAdjointStats(C.Structure):
    _fields_ = [("steps_max", C.c_int), ("segments", C.c_int),
                ("fwd_ms", C.c_double), ("bwd_ms", C.c_double),
                ("peak_bytes", C.c_longlong), ("ckpt_slots", C.c_int)]


class MpciError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mpci error {code}: {msg}")
        self.code = code


_tdse = None
_host = None

# (name, restype, argtypes)
_SIGS = [
    ("mpci_last_error", C.c_char_p, []),
    ("mpci_device_count", C.c_int, [_ip]),
    ("mpci_to_au", None, [C.c_double, C.c_double, _dp, _dp]),
    ("mpci_period", C.c_double, [C.c_double, C.c_int]),
    ("mpci_pulse_sine", C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_char, C.POINTER(MpciPulse)]),
    ("mpci_pulse_gauss", C.c_int, [C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_double, C.c_char,
                                   C.POINTER(MpciPulse)]),
    ("mpci_pulse_eval", C.c_double, [C.POINTER(MpciPulse), C.c_double]),
    ("mpci_system_create", C.c_int, [C.c_int, _ip, C.c_char, C.POINTER(_dp),
                                     C.POINTER(_dp), C.POINTER(_vp)]),
    ("mpci_system_destroy", C.c_int, [_vp]),
    ("mpci_system_set_gemv_grid", C.c_int, [_vp, C.c_int]),
    ("mpci_system_info", C.c_int, [_vp, _ip, C.POINTER(C.c_longlong)]),
    ("mpci_prop_create", C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    ("mpci_prop_destroy", C.c_int, [_vp]),
    ("mpci_prop_set_state", C.c_int, [_vp, C.c_int, _dp]),
    ("mpci_prop_get_state", C.c_int, [_vp, C.c_int, _dp]),
    ("mpci_prop_set_pulse", C.c_int, [_vp, C.c_int, C.POINTER(MpciPulse)]),
    ("mpci_prop_rhs", C.c_int, [_vp, _dp, _dp, _dp]),
    ("mpci_system_rhs_noalloc", C.c_int, [_vp, C.c_double, _vp, _vp, _vp]),
    ("mpci_prop_fixed", C.c_int, [_vp, C.c_double, C.c_double, C.c_int, _dp,
                                  C.c_int, _dp]),
    ("mpci_prop_adaptive", C.c_int, [_vp, _dp, _dp, _dp, C.c_double,
                                     C.c_double, C.c_int, C.c_int,
                                     C.POINTER(MpciStats), _dp, C.c_int]),
    ("mpci_multi_adaptive", C.c_int, [C.c_int, C.POINTER(_vp),
                                      C.POINTER(_dp), C.POINTER(_dp),
                                      C.POINTER(_dp), C.c_double, C.c_double,
                                      C.c_int,
                                      C.POINTER(C.POINTER(MpciStats))]),
    ("mpci_multi_adaptive_grouped", C.c_int, [C.c_int, C.POINTER(_vp),
                                              C.POINTER(_dp), C.POINTER(_dp),
                                              C.POINTER(_dp), C.c_double,
                                              C.c_double, C.c_int,
                                              C.POINTER(C.POINTER(MpciStats))]),
    ("mpci_prop_schedule", C.c_int, [_vp, _dp, C.c_int, _dp]),
    ("mpci_prop_populations", C.c_int, [_vp, C.c_int, _dp, _dp]),
    ("mpci_obs_weights", C.c_int, [C.c_int, _ip, C.POINTER(_dp), C.c_int,
                                   C.c_double, _dp]),
    ("mpci_prop_observables", C.c_int, [_vp, C.c_int, C.c_double, _dp]),
    ("mpci_prop_adjoint", C.c_int, [_vp, _dp, _dp, _dp, C.c_double,
                                    C.c_double, C.c_int, _dp, C.c_int, _dp,
                                    _dp, C.c_int, _ip, _dp, C.c_int,
                                    C.POINTER(MpciAdjointStats)]),
    ("mpci_prop_set_graph", C.c_int, [_vp, C.c_int]),
    ("mpci_prop_set_instrument", C.c_int, [_vp, C.c_int]),
    ("mpci_prop_kernel_time", C.c_int, [_vp, _dp, C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_longlong)]),
    ("mpci_prop_gemm_config", C.c_int,
     [_vp, _ip, _ip, _ip, _dp, _ip, _ip, _ip, _ip]),
    ("mpci_prop_gemm_candidates", C.c_int, [_ip]),
    ("mpci_prop_set_small_path", C.c_int, [_vp, C.c_int, _ip]),
    ("mpci_prop_set_gemm_candidate", C.c_int, [_vp, C.c_int]),
    ("mpci_probe_hbm_read", C.c_int, [C.c_longlong, C.c_int, _dp]),
    ("mpci_probe_mfma_f64", C.c_int, [C.c_int, _dp]),
]

# Every symbol the header declares; tests check the library exports them.
EXPORTED = [s[0] for s in _SIGS]


def tdse_lib():
    """Load the HIP engine; raises if it is not built."""
    global _tdse
    if _tdse is None:
        if not os.path.exists(TDSE_LIB):
            raise RuntimeError(
                f"{TDSE_LIB} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`; "
                "the tdse engine has no CPU fallback")
        lib = C.CDLL(TDSE_LIB)
        for name, res, args in _SIGS:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _tdse = lib
    return _tdse


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is missing (build first)")
        lib = C.CDLL(HOST_LIB)
        lib.mpci_synth_system.restype = C.c_int
        lib.mpci_synth_system.argtypes = [C.c_uint64, C.c_int, _ip, C.c_double,
                                          C.c_double, C.POINTER(_dp),
                                          C.POINTER(_dp)]
        lib.mpci_synth_uniform.restype = None
        lib.mpci_synth_uniform.argtypes = [C.c_uint64, C.c_uint64,
                                           C.c_longlong, C.c_double,
                                           C.c_double, _dp]
        lib.mpci_synth_normal.restype = None
        lib.mpci_synth_normal.argtypes = [C.c_uint64, C.c_uint64,
                                          C.c_longlong, C.c_double, _dp]
        _host = lib
    return _host


def check(rc: int) -> None:
    if rc != MPCI_OK:
        msg = tdse_lib().mpci_last_error()
        raise MpciError(rc, msg.decode() if msg else "")


def dptr(a):
    return a.ctypes.data_as(_dp)
