"""Pins the oracle's adaptive extension (no reference counterpart: the
reference only has fixed steps, src/tdse/tdse.cpp:186-215) and its
reference-mode loop.

Pins:
  * zero field: the exact solution is exp(-i E t) per component;
  * a time-dependent two-level problem against scipy's DOP853 at rtol 1e-13
    (field at stage times: a wrong node c_s or a frozen field would miss);
  * replaying the accepted step list reproduces the adaptive result exactly
    (the controller only chooses h, it does not alter the map);
  * reference mode: unit norm after every step and t accumulated by t += dt.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp


def diag_sys(oracle):
    E = [np.array([-2.9, -0.4, 1.3, 5.0]), np.array([-1.0, 2.2, 3.9])]
    return oracle.System(E, [np.zeros((4, 3))]), np.concatenate(E)


def test_zero_field_exact(oracle):
    sys_, E = diag_sys(oracle)
    p = oracle.Pulse.sine(0.0, 0.7, 4)  # Io = 0 -> F == 0
    x0 = np.zeros(2 * sys_.N)
    x0[0::2] = 1.0 / math.sqrt(sys_.N)
    T = 20.0
    x, st, hs = sys_.prop_adaptive(p, 0.0, T, 0.1, 1e-10, 1e-10, x0)
    exact = np.exp(-1j * E * T) / math.sqrt(sys_.N)
    assert np.abs((x[0::2] + 1j * x[1::2]) - exact).max() < 1e-8
    assert st.accepted == len(hs) and abs(hs.sum() - T) < 1e-12
    assert st.t == T


def test_two_level_time_dependent_vs_dop853(oracle):
    E0, E1, d = -0.9, 0.4, 1.2
    sys_ = oracle.System([np.array([E0]), np.array([E1])], [np.array([[d]])],
                         gauge="l")
    w, cyc = 0.7, 4
    Io = 0.2 ** 2
    p = oracle.Pulse.sine(Io, w, cyc, gauge="l")
    T = cyc * 2 * math.pi / w

    def f(t, c):
        F = p(t)
        H = np.array([[E0, F * d], [F * d, E1]])
        return -1j * H @ c

    ref = solve_ivp(f, (0, T), np.array([1.0 + 0j, 0j]), method="DOP853",
                    rtol=1e-13, atol=1e-14).y[:, -1]
    ref = ref / np.linalg.norm(ref)
    x, st, _ = sys_.prop_adaptive(p, 0.0, T, 0.1, 1e-10, 1e-10,
                                  np.array([1.0, 0.0, 0.0, 0.0]))
    got = x[0::2] + 1j * x[1::2]
    assert np.abs(got - ref).max() < 1e-8, (got, ref)
    # frozen field (reference mode) is only first order: it must NOT match
    xf, _, _, _ = sys_.prop_fixed(p, 0.0, T / 400, 400, np.array([1.0, 0, 0, 0]))
    assert np.abs((xf[0::2] + 1j * xf[1::2]) - ref).max() > 1e-6


def test_schedule_replay_reproduces_adaptive(oracle):
    rng = np.random.default_rng(4)
    s = [5, 4, 6]
    E = [np.sort(rng.uniform(-2.9, 5, n)) for n in s]
    D = [rng.standard_normal((s[i], s[i + 1])) / 2 for i in range(2)]
    sys_ = oracle.System(E, D, gauge="v")
    p = oracle.Pulse.sine(0.03 ** 2, 0.7, 3)
    x0 = np.zeros(2 * sys_.N)
    x0[0] = 1.0
    T = 3 * 2 * math.pi / 0.7
    x, st, hs = sys_.prop_adaptive(p, 0.0, T, 0.05, 1e-10, 1e-10, x0)
    assert st.rejected >= 0 and st.accepted > 10
    xr = sys_.prop_schedule(p, 0.0, hs, x0)
    assert np.array_equal(x, xr)


def test_reference_mode_loop(oracle):
    rng = np.random.default_rng(9)
    s = [6, 5, 7]
    E = [np.sort(rng.uniform(-2.9, 5, n)) for n in s]
    D = [rng.standard_normal((s[i], s[i + 1])) / 3 for i in range(2)]
    sys_ = oracle.System(E, D, gauge="v")
    p = oracle.Pulse.sine(0.05 ** 2, 0.5, 2)
    x0 = np.zeros(2 * sys_.N)
    x0[0] = 1.0
    x, t, pop, fld = sys_.prop_fixed(p, 0.0, 0.1, 40, x0, track=0)
    tt = 0.0
    for st in range(40):
        assert fld[st] == p(tt + 0.1)   # field frozen at t + dt (tdse.cpp:187)
        tt += 0.1
    assert t == tt
    assert abs(np.dot(x, x) - 1.0) < 1e-14
    assert abs(pop[-1] - (x[0] ** 2 + x[1] ** 2)) < 1e-16
    # one step by hand: do_step at frozen F, then divide by the 2-norm
    y = sys_.step(p(0.1), x0, 0.1)
    y /= np.linalg.norm(y)
    _, _, pop1, _ = sys_.prop_fixed(p, 0.0, 0.1, 1, x0, track=0)
    assert abs(pop1[0] - (y[0] ** 2 + y[1] ** 2)) < 1e-16
