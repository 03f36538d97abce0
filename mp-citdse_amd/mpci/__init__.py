"""MI355X (gfx950) implementation of MP-CITDSE's `tdse` propagation path.

All compute runs in lib/libmpci_tdse.so (hand-written HIP kernels behind the
C ABI of include/mpci_tdse.h); this package is the Python-side mirror of
that interface.
"""
from ._lib import (EXPORTED, OBS_EXCITED, OBS_GROUND, OBS_NORM,  # noqa: F401
                   OBS_ROWS, OBS_YIELD, MpciError)
from .engine import (Propagator, Pulse, System, device_count,  # noqa: F401
                     multi_adaptive, obs_weights, period, probe_hbm_read,
                     probe_mfma_f64, to_au)
