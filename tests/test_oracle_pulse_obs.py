"""Pins the oracle's pulses (src/tdse/pulse.cpp:3-97, au.hpp:30-32) and
observables (src/pes/pes.cpp:151-262) with closed forms and hand-computed
cases."""
import math

import numpy as np
import pytest


def test_unit_conversions(oracle):
    # 1 Hartree = 27.211396181 eV and 1 a.u. of intensity = 3.509338e16 W/cm^2
    Io, w = oracle.to_au(3.509338e16, 27.211396181)
    assert abs(Io - 1.0) < 1e-15 and abs(w - 1.0) < 1e-15
    assert abs(oracle.period(0.5, 10) - 40 * math.pi) < 1e-12


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_sine_squared_closed_form(oracle, gauge):
    Io, w, cyc = 0.0025, 0.5, 10          # E0 = 0.05 a.u.
    p = oracle.Pulse.sine(Io, w, cyc, gauge=gauge)
    amp = 0.05 / w if gauge == "v" else 0.05
    tau = cyc * 2 * math.pi / w
    for t in np.linspace(0.0, tau, 17):
        env = math.sin(math.pi * t / tau) ** 2
        assert abs(p(t) - amp * env * math.sin(w * t)) < 1e-15
    assert p(-1e-9) == 0.0 and p(tau * (1 + 1e-9)) == 0.0
    # peak of the envelope at tau/2
    assert abs(p(tau / 2) - amp * math.sin(w * tau / 2)) < 1e-15


def test_sine_cepd_is_in_cycles(oracle):
    # pulse.cpp:45 converts the phase from cycles to a time shift * w
    p = oracle.Pulse.sine(1.0, 0.7, 4, gauge="l", cepd=0.25)
    t = 3.0
    tau = 4 * 2 * math.pi / 0.7
    ref = math.sin(math.pi * t / tau) ** 2 * math.sin(0.7 * t + 0.25 * 2 * math.pi / 0.7)
    assert abs(p(t) - ref) < 1e-15


@pytest.mark.parametrize("gauge", ["v", "l"])
@pytest.mark.parametrize("cycles", [4, 12, 25])
def test_gaussian_fwhm(oracle, gauge, cycles):
    """The intensity envelope's FWHM is (2 + cycles/10) optical periods."""
    w = 0.7
    tau = cycles * 2 * math.pi / w
    p = oracle.Pulse.gauss(0.0009, w, tau, cycles, gauge=gauge)
    Eo = 0.03 / w if gauge == "v" else 0.03
    a, b = p.c.a, p.c.b
    assert b == tau and abs(p.c.Eo - Eo) < 1e-17
    fwhm = (2 + cycles // 10) * 2 * math.pi / w
    half = fwhm / 2
    env = lambda t: math.exp(-((t - tau / 2) / a) ** 2)
    assert abs(env(tau / 2 + half) ** 2 - 0.5) < 1e-12
    t = 0.3 * tau
    assert abs(p(t) - Eo * env(t) * math.sin(w * t)) < 1e-15
    assert p(tau * 1.0001) == 0.0


def test_pwl_envelope(oracle):
    env = np.array([0.0, 0.04, 0.01, 0.02])
    w, T = 0.7, 30.0
    p = oracle.Pulse.pwl(env, w, T, gauge="v")
    for k, tk in enumerate(np.linspace(0, T, 4)):
        assert abs(p(tk) - env[k] / w * math.sin(w * tk)) < 1e-15
    t = 15.0  # midway between knots 1 and 2
    assert abs(p(t) - 0.025 / w * math.sin(w * t)) < 1e-15
    assert p(T + 1e-6) == 0.0


def test_observables_hand_case(oracle):
    # two L blocks of sizes 4 and 3, threshold -2.0
    E = [np.array([-2.9, -2.1, -1.5, 0.3]), np.array([-2.5, -1.9, 2.0])]
    D = [np.zeros((4, 3))]
    sys_ = oracle.System(E, D)
    c = np.array([0.5, 0.1j, 0.2, 0.3 - 0.1j, 0.4, 0.25j, 0.6])
    x = np.empty(14)
    x[0::2], x[1::2] = c.real, c.imag
    out = sys_.observables(x, -2.0)
    p = np.abs(c) ** 2
    # yield: L=0 i in [1,2]: E=-2.1 (no), -1.5 (yes); L=1 i in [1,1]: -1.9 (yes)
    assert abs(out[0] - p[0]) < 1e-16
    assert abs(out[1] - (p[2] + p[5])) < 1e-16
    assert abs(out[2] - (p[1] + p[2] + p[5])) < 1e-16
    assert abs(out[3] - p.sum()) < 1e-15
