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


def _pops(c):
    x = np.empty(2 * len(c))
    x[0::2], x[1::2] = c.real, c.imag
    return x, np.abs(c) ** 2


def test_observables_1e_hand_binning(oracle):
    """genPES1e's yield (pes.cpp:90-110), worked out by hand.  Bins are the
    l = 1 energies k = 1 .. s_0 - 2 = 3 of [-0.5, -0.125, 0.125, 0.25, 0.75,
    1.0], each reaching halfway to its neighbours, open at both ends:
      k = 1: centre -0.125, (-0.3125, 0.0)    negative: no yield
      k = 2: centre  0.125, (0.0, 0.1875)     yield
      k = 3: centre  0.25,  (0.1875, 0.5)     yield
    (with the neighbour spacings swapped, k = 3 would be (0.0, 0.3125); with
    bins up to s_1 - 2, the state at 0.75 would count; with i from 0, the
    first state of l = 2 would count).  Every value is a multiple of 1/64,
    so the edges are exact."""
    E = [np.array([-2.0, -0.25, 0.0, 0.4375, 1.0]),
         np.array([-0.5, -0.125, 0.125, 0.25, 0.75, 1.0]),
         np.array([0.0625, 0.15625, 0.1875])]
    D = [np.zeros((5, 6)), np.zeros((6, 3))]
    sys_ = oracle.System(E, D)
    c = np.exp(1j * np.arange(14)) * np.sqrt(np.arange(1, 15) / 105.0)
    x, p = _pops(c)
    out = sys_.observables_1e(x)
    # l = 0: -0.25 in bin 1 (negative), 0.0 on the edge of bins 1 and 2
    #        (in neither), 0.4375 in bin 3, 1.0 in none;
    # l = 1: -0.125 bin 1, 0.125 bin 2, 0.25 bin 3, 0.75 and 1.0 none;
    # l = 2: 0.15625 bin 2, 0.1875 on the edge of bins 2 and 3
    assert abs(out[1] - (p[3] + p[7] + p[8] + p[12])) < 1e-16
    assert abs(out[0] - p[0]) < 1e-16
    assert out[2] == 0.0
    assert abs(out[3] - p.sum()) < 1e-15
