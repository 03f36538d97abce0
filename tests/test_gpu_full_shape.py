"""GPU parity at the full He shape (4 blocks of 4000 CI states, dense FP64
dipole blocks of 4000 x 4000; synthetic, seeded).

test_adaptive_full_shape_vs_oracle: the single-pulse adaptive mode at
tol 1e-10 (the headline workload) against the oracle run on the host with
the reference's cblas_dgemv call pattern (MKL when present): the first
attempts from t = 0 and from the peak of the pulse.  Bars: equal accepted
and rejected counts; accepted step sizes equal to 1e-6 relative (the error
ratio that sets them is a difference of O(1) terms at the 1e-10 level, so
summation-order rounding of ~1e-16 in the state moves it by ~1e-6 relative
and h = 0.9 E^(-1/8) h by ~1e-7: two correct runs end at times that differ
by ~1e-8, and their states by that time difference times |dc/dt|); and the
GPU state within 1e-10 of the oracle's map replayed over the GPU's own
accepted steps.

test_adjoint_full_config: the adjoint workload at its own configuration
(20-cycle piecewise-linear pulses, w = 0.7, 256 envelope samples
~ U[0, 0.04], a batch of 16 envelopes, checkpoints every 32 accepted
steps, the batched MFMA coupling in the forward, recompute and backward
sweeps): eight gradient entries spread over the pulse, for the
ground-population and the yield objectives, against central finite
differences (h = 1e-6) of the same discrete map on the frozen accepted
schedule (the map of src/tdse/tdse.cpp:186-204 with its renormalisation),
all sixteen perturbed envelopes in one batched schedule replay.  Bar:
relative error <= 1e-6.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

S_HE = [4000, 4000, 4000, 4000]
FD_H = 1e-6


@pytest.fixture(scope="module")
def mpci():
    import mpci as m
    assert m.device_count() >= 1
    return m


@pytest.fixture(scope="module")
def he(mpci):
    from mpci import synth
    E, D = synth.system_arrays(S_HE, seed=synth.HE_SEED)
    return E, D, mpci.System(E, D, gauge="v")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def use_host_blas(oracle):
    for path in ("/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"):
        if os.path.exists(path) and oracle.use_blas(path):
            return path
    return None


@pytest.mark.timeout(600)
def test_adaptive_full_shape_vs_oracle(mpci, oracle, he):
    from mpci import synth
    E, D, g = he
    o = oracle.System(E, D, gauge="v")
    w, e0, cyc, tol = 0.7, 0.03, 10, 1e-10
    tau = mpci.period(w, cyc)
    x0 = synth.ground_state(g.N)
    blas = use_host_blas(oracle)
    print(f"oracle dgemv: {blas or 'built-in loops'}", flush=True)
    try:
        for t0, n in ((0.0, 12), (0.5 * tau, 12)):
            prop = mpci.Propagator(g)
            prop.set_pulse(mpci.Pulse.sine(e0 ** 2, w, cyc))
            prop.set_state(x0)
            st, log = prop.prop_adaptive(t0, tau, 0.05, tol, tol,
                                         max_attempts=n, allow_limit=True,
                                         log_cap=n)
            xg = prop.get_state()
            prop.close()
            pulse = oracle.Pulse.sine(e0 ** 2, w, cyc)
            xo, so, ho = o.prop_adaptive(pulse, t0, tau, 0.05, tol, tol, x0,
                                         max_attempts=n, allow_limit=True)
            hg = log[0, :st[0].accepted, 1]
            xr = o.prop_schedule(pulse, t0, hg, x0)
            print(f"t0={t0:.3f}: GPU {st[0].accepted}/{st[0].rejected}, "
                  f"oracle {so.accepted}/{so.rejected}, "
                  f"max|dh|/h={np.abs(hg - ho).max() / ho.max():.2e}, "
                  f"|dt|={abs(st[0].t - so.t):.2e}, "
                  f"max|dx| on the GPU schedule={np.abs(xg - xr).max():.2e}",
                  flush=True)
            assert st[0].attempts == n
            assert (st[0].accepted, st[0].rejected) == (so.accepted,
                                                        so.rejected)
            assert hg.shape == ho.shape
            assert np.abs(hg - ho).max() <= 1e-6 * ho.max()
            assert abs(st[0].t - so.t) <= 1e-6 * (so.t - t0)
            assert abs(st[0].t - (t0 + hg.sum())) < 1e-12
            assert np.abs(xg - xr).max() < 1e-10
            pg = xg[0::2] ** 2 + xg[1::2] ** 2
            pr = xr[0::2] ** 2 + xr[1::2] ** 2
            assert np.abs(pg - pr).max() < 1e-10
    finally:
        oracle.use_builtin_dgemv()


def yield_weights(mpci, E, thr=-2.0):
    return mpci.obs_weights(E, 2, thr)[mpci.OBS_YIELD]


@pytest.mark.timeout(900)
def test_adjoint_full_config(mpci, he):
    from mpci import synth
    E, _, g = he
    B, K, cyc, w = 16, 256, 20, 0.7
    T = mpci.period(w, cyc)
    x0 = synth.ground_state(g.N)
    envs = [synth.uniform(77 + p, 0, K, 0.0, 0.04) for p in range(B)]
    objectives = {"ground": np.eye(1, g.N)[0], "yield": yield_weights(mpci, E)}
    checked = 16 + 32 * np.arange(8)  # samples spread over the pulse
    prop = mpci.Propagator(g, P=B)
    grads, hlog0 = {}, None
    for name, W in objectives.items():
        for p in range(B):
            prop.set_state(x0, p)
            prop.set_pulse(mpci.Pulse.pwl(envs[p], w, T), p)
        J, grad, nst, hlog, st = prop.adjoint(0.0, T, 0.05, W, ckpt_every=32,
                                              hlog_cap=100000)
        print(f"{name}: J[0]={J[0]:.12e} steps={nst.tolist()} "
              f"segments={st.segments} fwd={st.fwd_ms:.0f} ms "
              f"bwd={st.bwd_ms:.0f} ms", flush=True)
        assert st.segments == math.ceil(nst.max() / 32)
        assert st.segments > 40
        grads[name] = grad[0, checked]
        h0 = hlog[0, :nst[0]].copy()
        if hlog0 is None:
            hlog0 = h0
        else:
            # the same forward map for both objectives
            assert np.array_equal(h0, hlog0)
    prop.close()
    # central differences of pulse 0's map on its frozen schedule: the eight
    # samples +- h in one batched replay of sixteen pulses
    fd = mpci.Propagator(g, P=2 * len(checked))
    keep = []
    for i, k in enumerate(checked):
        for sgn, p in ((1, 2 * i), (-1, 2 * i + 1)):
            e = envs[0].copy()
            e[k] += sgn * FD_H
            pl = mpci.Pulse.pwl(e, w, T)
            keep.append(pl)
            fd.set_pulse(pl, p)
            fd.set_state(x0, p)
    fd.prop_schedule(0.0, np.tile(hlog0, (fd.P, 1)))
    Jfd = fd.populations(np.stack([objectives["ground"],
                                   objectives["yield"]]))
    for r, name in enumerate(("ground", "yield")):
        g_fd = (Jfd[r, 0::2] - Jfd[r, 1::2]) / (2 * FD_H)
        err = rel(grads[name], g_fd)
        print(f"{name}: adjoint {grads[name]}\n{name}: fd      {g_fd}\n"
              f"{name}: relative error {err:.3e}", flush=True)
        assert err <= 1e-6
        assert np.all(np.abs(grads[name] - g_fd) <=
                      1e-6 * np.abs(g_fd).max())
