"""GPU parity of the single-pulse path (fused GEMV RHS, stage kernels,
on-device control) against the CPU oracle, through the C ABI.

Bars: RHS within 1e-13 of the oracle (relative to max|x|); reference-mode
propagation within 1e-12 absolute per coefficient; adaptive mode with the
same accepted/rejected counts and final populations within 1e-8 (observed
~1e-13); full He shape (4 x 4000) checked for two steps and through the
norm invariant.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [[3, 2, 4], [1, 1], [5], [300, 257, 513], [256, 256], [1, 700, 2]]


@pytest.fixture(scope="module")
def mpci():
    import mpci as m
    assert m.device_count() >= 1
    return m


def fixed_fields(pulse, t0, dt, steps):
    """F(t + dt) with t accumulated by t += dt (src/tdse/tdse.cpp:187,192)."""
    F, t = np.empty(steps), t0
    for k in range(steps):
        F[k] = pulse(t + dt)
        t += dt
    return F[None, :]


def build(mpci, oracle, s, gauge, seed=5):
    from mpci import synth
    E, D = synth.system_arrays(s, seed=seed)
    return mpci.System(E, D, gauge=gauge), oracle.System(E, D, gauge=gauge)


@pytest.mark.parametrize("gauge", ["v", "l"])
@pytest.mark.parametrize("s", SHAPES)
def test_rhs_parity(mpci, oracle, gauge, s):
    g, o = build(mpci, oracle, s, gauge, seed=sum(s))
    prop = mpci.Propagator(g)
    x = np.random.default_rng(3).standard_normal(2 * g.N)
    for F in (0.0, 0.37, -1.3):
        got = prop.rhs(F, x)
        ref = o.rhs(F, x)
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() <= 1e-13 * scale, (s, gauge, F)


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_fixed_mode_parity(mpci, oracle, gauge):
    g, o = build(mpci, oracle, [40, 35, 50, 20], gauge)
    prop = mpci.Propagator(g)
    x0 = np.zeros(2 * g.N)
    x0[0] = 1.0
    pulse = oracle.Pulse.sine(0.05 ** 2, 0.5, 2, gauge=gauge)
    steps, dt = 120, 0.1
    F = fixed_fields(pulse, 0.0, dt, steps)
    prop.set_state(x0)
    pop = prop.prop_fixed(0.0, dt, F, track=3)
    x = prop.get_state()
    xr, _, popr, fr = o.prop_fixed(pulse, 0.0, dt, steps, x0, track=3)
    assert np.array_equal(F[0], fr)
    assert np.abs(x - xr).max() < 1e-12
    assert np.abs(pop[0] - popr).max() < 1e-12
    assert abs(np.dot(x, x) - 1.0) < 1e-13


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_adaptive_parity(mpci, oracle, gauge):
    g, o = build(mpci, oracle, [60, 50, 70, 40], gauge, seed=11)
    prop = mpci.Propagator(g)
    x0 = np.zeros(2 * g.N)
    x0[0] = 1.0
    w, cyc = 0.7, 2
    Io = 0.03 ** 2
    T = cyc * 2 * math.pi / w
    po = oracle.Pulse.sine(Io, w, cyc, gauge=gauge)
    pg = mpci.Pulse.sine(Io, w, cyc, gauge=gauge)
    prop.set_state(x0)
    prop.set_pulse(pg)
    st, log = prop.prop_adaptive(0.0, T, 0.1, 1e-10, 1e-10, log_cap=100000)
    x = prop.get_state()
    xr, sr, hr = o.prop_adaptive(po, 0.0, T, 0.1, 1e-10, 1e-10, x0)
    assert st[0].accepted == sr.accepted and st[0].rejected == sr.rejected
    assert st[0].t == T
    # The embedded error estimate cancels ~10 digits, so last-bit differences
    # in the stage derivatives move the controller's h at the 1e-7 level.
    np.testing.assert_allclose(log[0, :st[0].accepted, 1], hr, rtol=1e-5)
    assert np.abs(x - xr).max() < 1e-10
    pops = lambda v: v[0::2] ** 2 + v[1::2] ** 2
    assert np.abs(pops(x) - pops(xr)).max() < 1e-8


def test_schedule_parity(mpci, oracle):
    g, o = build(mpci, oracle, [33, 47, 21], "l", seed=2)
    prop = mpci.Propagator(g)
    x0 = np.zeros(2 * g.N)
    x0[0] = 1.0
    po = oracle.Pulse.gauss(0.02 ** 2, 0.6, 30.0, 3, gauge="l")
    pg = mpci.Pulse.gauss(0.02 ** 2, 0.6, 30.0, 3, gauge="l")
    hs = np.random.default_rng(1).uniform(0.05, 0.15, 80)
    prop.set_state(x0)
    prop.set_pulse(pg)
    prop.prop_schedule(0.0, hs)
    x = prop.get_state()
    xr = o.prop_schedule(po, 0.0, hs, x0)
    assert np.abs(x - xr).max() < 1e-12


def test_populations(mpci, oracle):
    g, o = build(mpci, oracle, [20, 30], "v", seed=4)
    prop = mpci.Propagator(g)
    x = np.random.default_rng(0).standard_normal(2 * g.N)
    prop.set_state(x)
    W = np.random.default_rng(1).uniform(0, 1, (3, g.N))
    got = prop.populations(W)[:, 0]
    ref = W @ (x[0::2] ** 2 + x[1::2] ** 2)
    np.testing.assert_allclose(got, ref, rtol=1e-13)


def test_graph_and_eager_agree(mpci, oracle):
    g, _ = build(mpci, oracle, [130, 90, 110], "v", seed=8)
    x0 = np.zeros(2 * g.N)
    x0[0] = 1.0
    pg = mpci.Pulse.sine(0.04 ** 2, 0.7, 2)
    out = []
    for graph in (True, False):
        prop = mpci.Propagator(g)
        prop.set_graph(graph)
        prop.set_state(x0)
        prop.set_pulse(pg)
        prop.prop_adaptive(0.0, 10.0, 0.1, 1e-10, 1e-10)
        out.append(prop.get_state())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("grid", [1, 3, 7, 0])
def test_persistent_gemv_grid(mpci, oracle, grid):
    """The persistent coupling kernel gives the same RHS whatever the number
    of workgroups (each then streams many tiles, ragged last tiles
    included)."""
    g, o = build(mpci, oracle, [300, 420, 270, 150], "v", seed=12)
    g.set_gemv_grid(grid)
    prop = mpci.Propagator(g)
    assert not prop.set_small_path(0)  # the tile kernel, not the fused path
    rng = np.random.default_rng(grid)
    x = rng.standard_normal(2 * g.N)
    got = prop.rhs(0.37, x)
    ref = o.rhs(0.37, x)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    g.set_gemv_grid(0)


def test_instrumented_runs_agree(mpci, oracle):
    """The instrumented mode (eager launches with an event pair around every
    coupling kernel) gives the graph result bit for bit and counts 13
    coupling launches per attempt that did work."""
    g, _ = build(mpci, oracle, [130, 90, 110], "l", seed=9)
    x0 = np.zeros(2 * g.N)
    x0[0] = 1.0
    pg = mpci.Pulse.sine(0.04 ** 2, 0.7, 3, "l")
    T = mpci.period(0.7, 3)
    out = []
    for instr in (False, True):
        prop = mpci.Propagator(g)
        prop.set_instrument(instr)
        prop.set_state(x0)
        prop.set_pulse(pg)
        st, _ = prop.prop_adaptive(0.0, T, 0.1, 1e-10, 1e-10)
        assert st[0].attempts > 64  # more than one instrumented chunk
        out.append(prop.get_state())
        if instr:
            ms, n, b, f = prop.kernel_time()
            assert n == 13 * st[0].attempts, (n, st[0].attempts)
            assert ms > 0 and b == 8 * g.dipole_elements
    assert np.array_equal(out[0], out[1])


def test_full_he_shape_two_steps(mpci, oracle):
    """Config-2 He shape (4 blocks of 4000): two reference-mode steps."""
    g, o = build(mpci, oracle, [4000] * 4, "v", seed=20261202)
    prop = mpci.Propagator(g)
    x0 = np.zeros(2 * g.N)
    x0[0] = 1.0
    pulse = oracle.Pulse.sine(0.03 ** 2, 0.7, 10)
    F = fixed_fields(pulse, 20.0, 0.1, 2)
    # start from a spread-out state so every block is exercised
    x0 = np.random.default_rng(6).standard_normal(2 * g.N)
    x0 /= np.linalg.norm(x0)
    prop.set_state(x0)
    prop.prop_fixed(20.0, 0.1, F)
    x = prop.get_state()
    xr = x0.copy()
    for k in range(2):
        xr = o.step(F[0, k], xr, 0.1)
        xr /= math.sqrt(np.dot(xr, xr))
    assert np.abs(x - xr).max() < 1e-12
    assert abs(np.dot(x, x) - 1.0) < 1e-13


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_reference_he_test_shape_two_steps(mpci, oracle, gauge):
    """The block sizes of the reference's own He input
    (inp/He_test.yaml: states_in_l [3030, 2799, 2808, 2808]), uneven and
    ragged against every GEMV unit: two reference-mode steps in both
    gauges, and the batched form (P = 2) on the same system."""
    s = [3030, 2799, 2808, 2808]
    g, o = build(mpci, oracle, s, gauge, seed=31)
    pulse = oracle.Pulse.sine(0.03 ** 2, 0.55, 12)
    F = fixed_fields(pulse, 5.0, 0.1, 2)
    x0 = np.random.default_rng(7).standard_normal(2 * g.N)
    x0 /= np.linalg.norm(x0)
    xr = x0.copy()
    for k in range(2):
        xr = o.step(F[0, k], xr, 0.1)
        xr /= math.sqrt(np.dot(xr, xr))
    prop = mpci.Propagator(g)
    prop.set_state(x0)
    prop.prop_fixed(5.0, 0.1, F)
    assert np.abs(prop.get_state() - xr).max() < 1e-12
    batch = mpci.Propagator(g, P=2)
    for p in range(2):
        batch.set_state(x0, p)
    batch.prop_fixed(5.0, 0.1, np.vstack([F, F]))
    for p in range(2):
        assert np.abs(batch.get_state(p) - xr).max() < 1e-12, p


@pytest.mark.parametrize("gauge", ["v", "l"])
@pytest.mark.parametrize("s", [[3, 2, 4], [1, 1], [300, 257, 513],
                               [198, 198, 198, 198], [1, 700, 2]])
def test_small_path_against_tile_path(mpci, oracle, gauge, s):
    """One pulse on a small system: the fused small-system stage kernel and
    the GEMV tile + stage kernels give the same RHS (to 1e-13 of max|k|,
    the coupling sums run in another order) and the same fixed-step
    propagation, both against the oracle; the automatic choice takes the
    fused path for these shapes.  The propagation bar is 4e-12 after 25
    steps: rounding of ~5e-14 per step (13 stages of sums over up to 513
    terms) accumulates, against the 1e-12 of the two-step tests."""
    from mpci import synth
    g, o = build(mpci, oracle, s, gauge, seed=sum(s))
    rng = np.random.default_rng(len(s))
    x = rng.standard_normal(2 * g.N)
    ref = o.rhs(-0.61, x)
    bar = 1e-13 * max(1.0, np.abs(ref).max())
    small = mpci.Propagator(g)
    tile = mpci.Propagator(g)
    assert small.set_small_path(-1) == (len(s) > 1)
    assert not tile.set_small_path(0)
    assert np.abs(small.rhs(-0.61, x) - ref).max() <= bar
    assert np.abs(tile.rhs(-0.61, x) - ref).max() <= bar
    pulse = oracle.Pulse.sine(0.05 ** 2, 0.5, 3, gauge)
    F = fixed_fields(pulse, 0.0, 0.1, 25)
    x0 = synth.ground_state(g.N)
    xr, _, _, _ = o.prop_fixed(pulse, 0.0, 0.1, 25, x0)
    got = []
    for prop in (small, tile):
        prop.set_state(x0)
        prop.prop_fixed(0.0, 0.1, F)
        got.append(prop.get_state())
        assert np.abs(got[-1] - xr).max() < 4e-12
    assert np.abs(got[0] - got[1]).max() < 4e-12


def test_small_path_adaptive_and_graph(mpci, oracle):
    """The fused path under adaptive control, captured and eager: the
    oracle's accepted/rejected counts and its state within 1e-10, and the
    captured and eager runs bitwise equal."""
    from mpci import synth
    g, o = build(mpci, oracle, [198] * 4, "v", seed=41)
    x0 = synth.ground_state(g.N)
    T = 2 * 2 * math.pi / 0.5
    po = oracle.Pulse.sine(0.05 ** 2, 0.5, 2)
    xr, sr = o.prop_adaptive(po, 0.0, T, 0.05, 1e-10, 1e-10, x0)[:2]
    out = []
    for graph in (True, False):
        prop = mpci.Propagator(g)
        assert prop.set_small_path(1)
        prop.set_graph(graph)
        prop.set_pulse(mpci.Pulse.sine(0.05 ** 2, 0.5, 2))
        prop.set_state(x0)
        st, _ = prop.prop_adaptive(0.0, T, 0.05, 1e-10, 1e-10)
        out.append(prop.get_state())
        assert st[0].accepted == sr.accepted and st[0].rejected == sr.rejected
        assert np.abs(out[-1] - xr).max() < 1e-10
    assert np.array_equal(out[0], out[1])


def test_measurement_probes(mpci):
    """The ceiling probes: a 1 GiB streaming read lands between a sanity
    floor and the 8 TB/s published HBM peak (with room for cache effects);
    back-to-back FP64 MFMA stays at or below the 78.6 TFLOP/s published
    peak; arguments outside the documented ranges are refused."""
    bw = mpci.probe_hbm_read(1 << 30, 3)
    assert 500.0 < bw < 8000.0 * 1.25, bw
    tf = mpci.probe_mfma_f64(1024)
    assert 5.0 < tf <= 78.6 * 1.01, tf
    for bad in ((100, 1), (1 << 20, 0)):
        with pytest.raises(mpci.MpciError) as e:
            mpci.probe_hbm_read(*bad)
        assert e.value.code == -1
    with pytest.raises(mpci.MpciError):
        mpci.probe_mfma_f64(8)


@pytest.mark.parametrize("electrons", [1, 2])
def test_observables_match_oracle(mpci, oracle, electrons):
    """pes's observables on the device (mpci_prop_observables) for three
    pulses propagated adaptively, against the oracle's or_observables /
    or_observables_1e (src/pes/pes.cpp:48-262): on the GPU's own states
    within 1e-13, and on the oracle's propagated states within the
    adaptive-mode bar."""
    from mpci import synth
    s = [30, 36, 28]
    lo, hi = (-0.5, 2.0) if electrons == 1 else (-2.9, 5.0)
    E, D = synth.system_arrays(s, seed=31 + electrons, e_lo=lo, e_hi=hi)
    g = mpci.System(E, D, gauge="v")
    o = oracle.System(E, D, gauge="v")
    P, w, cyc, thr = 3, 0.7, 2, -2.0
    amps = [0.02, 0.05, 0.08]
    T = mpci.period(w, cyc)
    x0 = synth.ground_state(g.N)
    prop = mpci.Propagator(g, P=P)
    for p in range(P):
        prop.set_state(x0, p)
        prop.set_pulse(mpci.Pulse.sine(amps[p] ** 2, w, cyc), p)
    prop.prop_adaptive(0.0, T, 0.05, 1e-10, 1e-10)
    got = prop.observables(electrons, thr)
    assert got.shape == (mpci.OBS_ROWS, P)
    for p in range(P):
        xg = prop.get_state(p)
        ref_g = (o.observables_1e(xg) if electrons == 1
                 else o.observables(xg, thr))
        assert np.abs(got[:, p] - ref_g).max() < 1e-13
        xo, _, _ = o.prop_adaptive(oracle.Pulse.sine(amps[p] ** 2, w, cyc),
                                   0.0, T, 0.05, 1e-10, 1e-10, x0)
        ref = (o.observables_1e(xo) if electrons == 1
               else o.observables(xo, thr))
        assert np.abs(got[:, p] - ref).max() < 1e-10
    # the yield grows with the field strength, the norm stays 1
    assert got[mpci.OBS_YIELD, 0] < got[mpci.OBS_YIELD, 2]
    assert np.abs(got[mpci.OBS_NORM] - 1.0).max() < 1e-12


def test_pwl_envelope_change_between_graph_runs(mpci, oracle):
    """set_pulse with a longer piecewise-linear envelope moves the device
    envelope store, which a captured step refers to: the graph is recaptured
    and graph runs stay bit-equal to eager runs (then a same-length envelope
    is rewritten in place), each run matching the oracle's map replayed on
    the run's own accepted steps to 1e-10.  (Two independent adaptive runs
    through these jagged envelopes differ by ~1e-6: at an envelope kink the
    order-7 and order-8 solutions share an O(h^2) error that their
    difference, the controller's estimate, cannot see.)"""
    from mpci import synth
    g, o = build(mpci, oracle, [40, 52, 33], "v", seed=23)
    w, T = 0.7, 20.0
    envs = [synth.uniform(201, 0, 6, 0.0, 0.05),
            synth.uniform(202, 0, 40, 0.0, 0.05),
            synth.uniform(203, 0, 40, 0.0, 0.05)]
    x0 = synth.ground_state(g.N)
    runs, logs = {}, {}
    for graph in (True, False):
        prop = mpci.Propagator(g)
        prop.set_graph(graph)
        out, hl = [], []
        for env in envs:
            prop.set_state(x0)
            prop.set_pulse(mpci.Pulse.pwl(env, w, T))
            st, log = prop.prop_adaptive(0.0, T, 0.05, 1e-10, 1e-10,
                                         log_cap=4096)
            assert st[0].t == T and st[0].accepted < 4096
            out.append(prop.get_state())
            hl.append(log[0, :st[0].accepted, 1].copy())
        runs[graph], logs[graph] = out, hl
        prop.close()
    for k, env in enumerate(envs):
        assert np.array_equal(runs[True][k], runs[False][k]), k
        assert np.array_equal(logs[True][k], logs[False][k]), k
        xr = o.prop_schedule(oracle.Pulse.pwl(env, w, T), 0.0, logs[True][k],
                             x0)
        assert np.abs(runs[True][k] - xr).max() < 1e-10, k
