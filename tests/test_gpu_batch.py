"""GPU parity of the batched form (P pulses on one system, coupling as FP64
MFMA GEMMs with autotuned tiles, per-pulse step control) against the CPU
oracle run pulse by pulse.

Bars as for one pulse: RHS within 1e-13 (relative to max|k|); reference
mode within 1e-12; adaptive mode with each pulse's accepted/rejected
counts equal to the oracle's and its final state within 1e-10 (the
populations bar is 1e-8).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mpci():
    import mpci as m
    assert m.device_count() >= 1
    return m


def build(mpci, oracle, s, gauge, seed):
    from mpci import synth
    E, D = synth.system_arrays(s, seed=seed)
    return mpci.System(E, D, gauge=gauge), oracle.System(E, D, gauge=gauge)


@pytest.mark.parametrize("P", [2, 3, 8, 64])
@pytest.mark.parametrize("gauge", ["v", "l"])
def test_batched_rhs_parity(mpci, oracle, P, gauge):
    g, o = build(mpci, oracle, [70, 130, 45, 90], gauge, seed=P)
    prop = mpci.Propagator(g, P=P)
    cfg = prop.gemm_config()
    assert cfg["wm"] in (1, 2, 4) and cfg["wn"] in (1, 2, 4) and cfg["ms"] > 0
    assert cfg["fm"] in (2, 4) and cfg["fn"] in (2, 4)
    assert cfg["kernel"] in ("k_gemm_coupling", "k_gemm_db", "k_gemm_lds")
    rng = np.random.default_rng(P)
    x = rng.standard_normal((P, 2 * g.N))
    F = rng.uniform(-1, 1, P)
    got = prop.rhs(F, x.reshape(-1)).reshape(P, -1)
    for p in range(P):
        ref = o.rhs(F[p], x[p])
        assert np.abs(got[p] - ref).max() <= 1e-13 * max(1, np.abs(ref).max()), p


def test_batched_fixed_mode(mpci, oracle):
    from mpci import synth
    g, o = build(mpci, oracle, [50, 64, 33], "v", seed=21)
    P = 5
    prop = mpci.Propagator(g, P=P)
    w, E0, cyc = synth.pulse_batch(P, seed=3)
    steps, dt = 60, 0.1
    F = np.empty((P, steps))
    x0 = synth.ground_state(g.N)
    pulses = []
    for p in range(P):
        pl = oracle.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p]))
        pulses.append(pl)
        t = 0.0
        for k in range(steps):
            F[p, k] = pl(t + dt)
            t += dt
        prop.set_state(x0, p)
    pop = prop.prop_fixed(0.0, dt, F, track=1)
    for p in range(P):
        xr, _, popr, _ = o.prop_fixed(pulses[p], 0.0, dt, steps, x0, track=1)
        assert np.abs(prop.get_state(p) - xr).max() < 1e-12, p
        assert np.abs(pop[p] - popr).max() < 1e-12, p


def test_batched_adaptive_matches_oracle(mpci, oracle):
    from mpci import synth
    g, o = build(mpci, oracle, [40, 72, 56, 30], "v", seed=5)
    P = 6
    prop = mpci.Propagator(g, P=P)
    w, E0, cyc = synth.pulse_batch(P, seed=11)
    cyc = np.minimum(cyc, 5)  # keep the oracle side short
    x0 = synth.ground_state(g.N)
    t1 = np.array([mpci.period(w[p], int(cyc[p])) for p in range(P)])
    for p in range(P):
        prop.set_state(x0, p)
        prop.set_pulse(mpci.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p])), p)
    st, _ = prop.prop_adaptive(0.0, t1, 0.05, 1e-10, 1e-10)
    for p in range(P):
        po = oracle.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p]))
        xr, sr, _ = o.prop_adaptive(po, 0.0, t1[p], 0.05, 1e-10, 1e-10, x0)
        assert st[p].accepted == sr.accepted, p
        assert st[p].rejected == sr.rejected, p
        assert st[p].t == t1[p]
        x = prop.get_state(p)
        assert np.abs(x - xr).max() < 1e-10, p


def test_batched_equals_single_pulse_runs(mpci):
    """Uneven pulse lengths in one batch give the single-pulse results."""
    from mpci import synth
    E, D = synth.system_arrays([96, 80, 112], seed=9)
    g = mpci.System(E, D, gauge="l")
    P = 4
    w, E0, cyc = synth.pulse_batch(P, seed=7)
    cyc = np.minimum(cyc, 6)
    x0 = synth.ground_state(g.N)
    t1 = np.array([mpci.period(w[p], int(cyc[p])) for p in range(P)])
    batch = mpci.Propagator(g, P=P)
    for p in range(P):
        batch.set_state(x0, p)
        batch.set_pulse(mpci.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p]), "l"), p)
    stb, _ = batch.prop_adaptive(0.0, t1, 0.05, 1e-10, 1e-10)
    for p in range(P):
        one = mpci.Propagator(g)
        one.set_state(x0)
        one.set_pulse(mpci.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p]), "l"))
        st1, _ = one.prop_adaptive(0.0, t1[p], 0.05, 1e-10, 1e-10)
        assert st1[0].accepted == stb[p].accepted
        assert np.abs(one.get_state() - batch.get_state(p)).max() < 1e-11


@pytest.mark.parametrize("s", [[1, 1], [5], [3, 2, 4], [1, 700, 2], [45, 31]])
def test_batched_rhs_degenerate_shapes(mpci, oracle, s):
    """Single-state blocks, one block (no coupling), odd sizes (the padded
    row stride of the row-major copy) in the batched form."""
    P = 5
    g, o = build(mpci, oracle, s, "l", seed=sum(s) + 1)
    prop = mpci.Propagator(g, P=P)
    rng = np.random.default_rng(len(s))
    x = rng.standard_normal((P, 2 * g.N))
    F = rng.uniform(-1, 1, P)
    got = prop.rhs(F, x.reshape(-1)).reshape(P, -1)
    for p in range(P):
        ref = o.rhs(F[p], x[p])
        assert np.abs(got[p] - ref).max() <= 1e-13 * max(1, np.abs(ref).max())


def test_batched_rhs_max_pulses(mpci, oracle):
    """P = 256, the largest batch the interface accepts; 0 and 257 are
    refused with MPCI_ERR_ARG."""
    g, o = build(mpci, oracle, [40, 33, 20], "v", seed=77)
    P = 256
    prop = mpci.Propagator(g, P=P)
    rng = np.random.default_rng(256)
    x = rng.standard_normal((P, 2 * g.N))
    F = rng.uniform(-1, 1, P)
    got = prop.rhs(F, x.reshape(-1)).reshape(P, -1)
    for p in (0, 1, 127, 255):
        ref = o.rhs(F[p], x[p])
        assert np.abs(got[p] - ref).max() <= 1e-13 * max(1, np.abs(ref).max())
    for bad in (0, 257):
        with pytest.raises(mpci.MpciError) as e:
            mpci.Propagator(g, P=bad)
        assert e.value.code == -1


def test_empty_runs_leave_state(mpci, oracle):
    """Zero fixed steps and an empty adaptive interval change nothing; a
    batch where some pulses have an empty interval still runs the others."""
    from mpci import synth
    g, o = build(mpci, oracle, [30, 20, 25], "v", seed=8)
    P = 3
    prop = mpci.Propagator(g, P=P)
    rng = np.random.default_rng(9)
    xs = [rng.standard_normal(2 * g.N) for _ in range(P)]
    xs = [x / np.linalg.norm(x) for x in xs]
    for p in range(P):
        prop.set_state(xs[p], p)
        prop.set_pulse(mpci.Pulse.sine(0.03 ** 2, 0.7, 2), p)
    assert prop.prop_fixed(0.0, 0.1, np.zeros((P, 0))).shape == (P, 0)
    st, _ = prop.prop_adaptive(1.0, 1.0, 0.05, 1e-10, 1e-10)
    assert all(s.accepted == 0 and s.attempts == 0 for s in st)
    for p in range(P):
        assert np.array_equal(prop.get_state(p), xs[p])
    t1 = np.array([1.0, 0.0, 2.0])  # pulse 1: empty interval
    st, _ = prop.prop_adaptive(0.0, t1, 0.05, 1e-10, 1e-10)
    assert st[1].accepted == 0 and np.array_equal(prop.get_state(1), xs[1])
    for p in (0, 2):
        po = oracle.Pulse.sine(0.03 ** 2, 0.7, 2)
        xr, sr, _ = o.prop_adaptive(po, 0.0, t1[p], 0.05, 1e-10, 1e-10, xs[p])
        assert st[p].accepted == sr.accepted
        assert np.abs(prop.get_state(p) - xr).max() < 1e-10


@pytest.mark.parametrize("P", [3, 40])
def test_every_gemm_candidate(mpci, oracle, P):
    """Each tile the autotuner may choose, pinned in turn, on blocks that
    are ragged against every tile height (70, 130, 45, 90 rows) and pulse
    counts that leave partial column panels (6 and 80 real columns)."""
    g, o = build(mpci, oracle, [70, 130, 45, 90], "l", seed=P + 100)
    prop = mpci.Propagator(g, P=P)
    rng = np.random.default_rng(P)
    x = rng.standard_normal((P, 2 * g.N))
    F = rng.uniform(-1, 1, P)
    ref = np.stack([o.rhs(F[p], x[p]) for p in range(P)])
    bar = 1e-13 * max(1, np.abs(ref).max())
    seen = set()
    for i in range(mpci.Propagator.gemm_candidates()):
        cfg = prop.set_gemm_candidate(i)
        seen.add((cfg["kernel"], cfg["wm"], cfg["wn"], cfg["fm"], cfg["fn"],
                  cfg["tk"], cfg["ksplit"]))
        got = prop.rhs(F, x.reshape(-1)).reshape(P, -1)
        assert np.abs(got - ref).max() <= bar, (i, cfg)
    assert len(seen) == mpci.Propagator.gemm_candidates()


@pytest.mark.parametrize("P", [3, 40])
def test_candidates_with_equal_k_split_agree_bitwise(mpci, oracle, P):
    """Every K-split kernel sums a panel's K range in the same order
    (increasing k, four per MFMA), so tiles and kernels that share a K
    split give the same bits.  A grouped launch (the direct-to-LDS kernel,
    the group's tile, each member's own K split) relies on this to
    reproduce each member's own run, whichever tile its autotuner chose;
    a member tuned to k_gemm_pair keeps that kernel's own launch."""
    g, o = build(mpci, oracle, [70, 130, 45, 90], "v", seed=P + 300)
    prop = mpci.Propagator(g, P=P)
    rng = np.random.default_rng(P + 7)
    x = rng.standard_normal((P, 2 * g.N)).reshape(-1)
    F = rng.uniform(-1, 1, P)
    by_split = {}
    for i in range(mpci.Propagator.gemm_candidates()):
        cfg = prop.set_gemm_candidate(i)
        if cfg["kernel"] == "k_gemm_pair":
            continue  # its own order: the strip sum is split over 8 waves
        by_split.setdefault(cfg["ksplit"], []).append(
            (cfg, prop.rhs(F, x).copy()))
    assert len(by_split) >= 2
    for ks, runs in by_split.items():
        cfg0, ref = runs[0]
        for cfg, got in runs[1:]:
            assert np.array_equal(got, ref), (ks, cfg0, cfg)


def test_pinned_candidate_in_captured_steps(mpci, oracle):
    """Pinning a tile after steps were captured rebuilds the captured step
    (the job list and partial buffers are reallocated) and gives the same
    fixed-step result as the oracle."""
    from mpci import synth
    g, o = build(mpci, oracle, [50, 64, 33], "v", seed=22)
    P, steps, dt = 4, 20, 0.1
    w, E0, cyc = synth.pulse_batch(P, seed=4)
    pulses = [oracle.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p]))
              for p in range(P)]
    F = np.array([[pl(dt * (k + 1)) for k in range(steps)] for pl in pulses])
    x0 = synth.ground_state(g.N)
    prop = mpci.Propagator(g, P=P)
    for p in range(P):
        prop.set_state(x0, p)
    prop.prop_fixed(0.0, dt, F[:, :steps // 2])
    prop.set_gemm_candidate(mpci.Propagator.gemm_candidates() - 1)
    prop.prop_fixed(steps // 2 * dt, dt, F[:, steps // 2:])
    for p in range(P):
        xr, _, _, _ = o.prop_fixed(pulses[p], 0.0, dt, steps, x0, track=0)
        assert np.abs(prop.get_state(p) - xr).max() < 1e-12, p


def test_full_he_shape_batch64(mpci, oracle):
    """The batch-64 workload at its real shape (4 blocks of 4000 states,
    velocity gauge, P = 64): the autotuned MFMA coupling against the
    oracle's RHS on a sample of pulses, then three adaptive attempts of the
    batch against single-pulse runs of the same pulses (same accepted
    counts, states within 1e-11)."""
    from mpci import synth
    g, o = build(mpci, oracle, [4000] * 4, "v", seed=20261202)
    P = 64
    prop = mpci.Propagator(g, P=P)
    rng = np.random.default_rng(64)
    x = rng.standard_normal((P, 2 * g.N))
    F = rng.uniform(-0.05, 0.05, P)
    got = prop.rhs(F, x.reshape(-1)).reshape(P, -1)
    for p in (0, 21, 42, 63):
        ref = o.rhs(F[p], x[p])
        assert np.abs(got[p] - ref).max() <= 1e-13 * max(1, np.abs(ref).max()), p
    w, E0, cyc = synth.pulse_batch(P, seed=3)
    x0 = synth.ground_state(g.N)
    t1 = np.array([mpci.period(w[p], int(cyc[p])) for p in range(P)])
    for p in range(P):
        prop.set_state(x0, p)
        prop.set_pulse(mpci.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p])), p)
    st, _ = prop.prop_adaptive(0.0, t1, 0.05, 1e-10, 1e-10, max_attempts=3,
                               allow_limit=True)
    for p in (0, 21, 42, 63):
        one = mpci.Propagator(g)
        one.set_state(x0)
        one.set_pulse(mpci.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p])))
        s1, _ = one.prop_adaptive(0.0, t1[p], 0.05, 1e-10, 1e-10,
                                  max_attempts=3, allow_limit=True)
        assert s1[0].attempts == st[p].attempts == 3
        assert s1[0].accepted == st[p].accepted
        assert abs(s1[0].t - st[p].t) < 1e-12
        assert np.abs(one.get_state() - prop.get_state(p)).max() < 1e-11, p
        one.close()
    prop.close()


@pytest.mark.parametrize("ks", [1, 2])
def test_full_he_shape_pair_kernel(mpci, oracle, ks):
    """k_gemm_pair pinned at the adjoint's batch size on the real He shape
    (4 blocks of 4000 states: 16 strips of 256 rows per pair, the last one
    160 rows, and 5 or 10 column ranges): its RHS against the oracle's on a
    sample of pulses, at one and at two 32-column panels of X."""
    g, o = build(mpci, oracle, [4000] * 4, "v", seed=20261202)
    for P in (16, 20):
        prop = mpci.Propagator(g, P=P)
        for i in range(mpci.Propagator.gemm_candidates()):
            cfg = prop.set_gemm_candidate(i)
            if cfg["kernel"] == "k_gemm_pair" and cfg["ksplit"] == ks:
                break
        assert prop.gemm_config()["kernel"] == "k_gemm_pair"
        rng = np.random.default_rng(P + ks)
        x = rng.standard_normal((P, 2 * g.N))
        F = rng.uniform(-0.05, 0.05, P)
        got = prop.rhs(F, x.reshape(-1)).reshape(P, -1)
        for p in (0, 7, P - 1):
            ref = o.rhs(F[p], x[p])
            assert np.abs(got[p] - ref).max() <= \
                1e-13 * max(1, np.abs(ref).max()), (P, p)
        prop.close()

