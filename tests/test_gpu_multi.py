"""Multi-instance batch on one device: an H-shaped system with many small
blocks and a He-shaped system with few large ones, each with its own batch
of sin^2 pulses of uneven lengths, propagated concurrently by
mpci_multi_adaptive.  Every member must give exactly what its own
mpci_prop_adaptive call gives (same kernels, same order of operations on
its stream), and the members' pulses must match the oracle run pulse by
pulse (accepted/rejected counts equal, final state within 1e-10)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def mpci():
    import mpci as m
    assert m.device_count() >= 1
    return m


def member(mpci, s, gauge, seed, P, pseed, max_cycles):
    from mpci import synth
    E, D = synth.system_arrays(s, seed=seed)
    g = mpci.System(E, D, gauge=gauge)
    w, E0, cyc = synth.pulse_batch(P, seed=pseed)
    cyc = np.minimum(cyc, max_cycles)
    t1 = np.array([mpci.period(w[p], int(cyc[p])) for p in range(P)])
    pulses = [(E0[p] ** 2, w[p], int(cyc[p])) for p in range(P)]
    return dict(E=E, D=D, g=g, P=P, t1=t1, pulses=pulses, gauge=gauge)


def fresh(mpci, m):
    from mpci import synth
    prop = mpci.Propagator(m["g"], P=m["P"])
    x0 = synth.ground_state(m["g"].N)
    for p, (I, w, c) in enumerate(m["pulses"]):
        prop.set_state(x0, p)
        prop.set_pulse(mpci.Pulse.sine(I, w, c, m["gauge"]), p)
    return prop


@pytest.fixture(scope="module")
def members(mpci):
    return [member(mpci, [24] * 9, "v", 31, 5, 41, 4),      # H-like: 9 blocks
            member(mpci, [60, 72, 48], "v", 32, 3, 42, 5),   # He-like
            member(mpci, [40, 36, 30, 28], "l", 33, 1, 43, 3)]  # GEMV path


def test_multi_equals_separate_runs(mpci, members):
    props = [fresh(mpci, m) for m in members]
    st = mpci.multi_adaptive(props, [0.0] * 3, [m["t1"] for m in members],
                             [0.05] * 3, TOL, TOL)
    for i, m in enumerate(members):
        one = fresh(mpci, m)
        s1, _ = one.prop_adaptive(0.0, m["t1"], 0.05, TOL, TOL)
        for p in range(m["P"]):
            assert st[i][p].accepted == s1[p].accepted, (i, p)
            assert st[i][p].rejected == s1[p].rejected, (i, p)
            assert st[i][p].t == m["t1"][p]
            assert np.array_equal(props[i].get_state(p), one.get_state(p)), \
                (i, p)


def test_multi_matches_oracle(mpci, oracle, members):
    from mpci import synth
    props = [fresh(mpci, m) for m in members]
    st = mpci.multi_adaptive(props, [0.0] * 3, [m["t1"] for m in members],
                             [0.05] * 3, TOL, TOL)
    for i, m in enumerate(members):
        o = oracle.System(m["E"], m["D"], gauge=m["gauge"])
        x0 = synth.ground_state(m["g"].N)
        for p in (0, m["P"] - 1):
            I, w, c = m["pulses"][p]
            xr, sr, _ = o.prop_adaptive(oracle.Pulse.sine(I, w, c, m["gauge"]),
                                        0.0, m["t1"][p], 0.05, TOL, TOL, x0)
            assert st[i][p].accepted == sr.accepted, (i, p)
            assert st[i][p].rejected == sr.rejected, (i, p)
            assert np.abs(props[i].get_state(p) - xr).max() < 1e-10, (i, p)


def test_multi_attempt_limit(mpci, members):
    """A limit hit by one member is reported; the others still finish."""
    props = [fresh(mpci, m) for m in members[:2]]
    st = mpci.multi_adaptive(props, [0.0] * 2, [members[0]["t1"], 1e-3],
                             [0.05] * 2, TOL, TOL, max_attempts=7,
                             allow_limit=True)
    assert all(s.status != 0 and s.attempts == 7 for s in st[0])
    assert all(s.status == 0 and s.t == 1e-3 for s in st[1])
    with pytest.raises(mpci.MpciError):
        mpci.multi_adaptive([fresh(mpci, members[0])], [0.0],
                            [members[0]["t1"]], [0.05], TOL, TOL,
                            max_attempts=7)


def test_multi_full_shapes(mpci):
    """The uneven multi-instance workload at its real shapes: 32 pulses on
    an H-shaped system of 21 blocks of 998 states and 32 on the He shape
    (4 blocks of 4000), three attempts each, run concurrently; every member
    equals its own separate run bit for bit."""
    ms = [member(mpci, [998] * 21, "v", 51, 32, 61, 20),
          member(mpci, [4000] * 4, "v", 20261202, 32, 62, 20)]
    props = [fresh(mpci, m) for m in ms]
    st = mpci.multi_adaptive(props, [0.0] * 2, [m["t1"] for m in ms],
                             [0.05] * 2, TOL, TOL, max_attempts=3,
                             allow_limit=True)
    for i, m in enumerate(ms):
        one = fresh(mpci, m)
        s1, _ = one.prop_adaptive(0.0, m["t1"], 0.05, TOL, TOL,
                                  max_attempts=3, allow_limit=True)
        for p in range(m["P"]):
            assert st[i][p].attempts == s1[p].attempts == 3, (i, p)
            assert st[i][p].accepted == s1[p].accepted, (i, p)
            assert st[i][p].t == s1[p].t, (i, p)
            assert np.array_equal(props[i].get_state(p), one.get_state(p)), \
                (i, p)
        one.close()
    for pr in props:
        pr.close()


def test_grouped_equals_separate_runs(mpci, members):
    """Every member's attempt in the same launches (one interleaved GEMM job
    list, one stage, finalize and apply launch per attempt): each member
    equals its own separate run bit for bit, whatever tile the group uses
    (each member keeps its own K split)."""
    ms = members[:2]  # the batched members
    props = [fresh(mpci, m) for m in ms]
    st = mpci.multi_adaptive(props, [0.0] * 2, [m["t1"] for m in ms],
                             [0.05] * 2, TOL, TOL, grouped=True)
    for i, m in enumerate(ms):
        one = fresh(mpci, m)
        s1, _ = one.prop_adaptive(0.0, m["t1"], 0.05, TOL, TOL)
        for p in range(m["P"]):
            assert st[i][p].accepted == s1[p].accepted, (i, p)
            assert st[i][p].rejected == s1[p].rejected, (i, p)
            assert st[i][p].t == m["t1"][p]
            assert np.array_equal(props[i].get_state(p), one.get_state(p)), \
                (i, p)


def test_grouped_with_pair_kernel_member(mpci, members):
    """A member tuned to k_gemm_pair (both coupling directions from one read
    of each dipole block) keeps its own coupling launch inside the grouped
    graph, the other member joining the grouped job list: both still equal
    their own separate runs bit for bit."""
    ms = members[:2]

    def pinned(m, i):
        prop = fresh(mpci, m)
        if i == 0:
            for k in range(mpci.Propagator.gemm_candidates()):
                cfg = prop.set_gemm_candidate(k)
                if cfg["kernel"] == "k_gemm_pair" and cfg["ksplit"] == 1:
                    break
            assert prop.gemm_config()["kernel"] == "k_gemm_pair"
        return prop

    props = [pinned(m, i) for i, m in enumerate(ms)]
    st = mpci.multi_adaptive(props, [0.0] * 2, [m["t1"] for m in ms],
                             [0.05] * 2, TOL, TOL, grouped=True)
    for i, m in enumerate(ms):
        one = pinned(m, i)
        s1, _ = one.prop_adaptive(0.0, m["t1"], 0.05, TOL, TOL)
        for p in range(m["P"]):
            assert st[i][p].accepted == s1[p].accepted, (i, p)
            assert st[i][p].rejected == s1[p].rejected, (i, p)
            assert np.array_equal(props[i].get_state(p), one.get_state(p)), \
                (i, p)


def test_grouped_refuses_single_pulse_member(mpci, members):
    props = [fresh(mpci, m) for m in members]  # the third has P = 1
    with pytest.raises(mpci.MpciError) as ei:
        mpci.multi_adaptive(props, [0.0] * 3, [m["t1"] for m in members],
                            [0.05] * 3, TOL, TOL, grouped=True)
    assert ei.value.code == -5


def test_grouped_full_shapes(mpci):
    """The uneven multi-instance workload at its real shapes (32 pulses on
    21 blocks of 998 states next to 32 on 4 blocks of 4000) in grouped
    launches: three attempts each, every member equal to its own separate
    run bit for bit."""
    ms = [member(mpci, [998] * 21, "v", 51, 32, 61, 20),
          member(mpci, [4000] * 4, "v", 20261202, 32, 62, 20)]
    props = [fresh(mpci, m) for m in ms]
    st = mpci.multi_adaptive(props, [0.0] * 2, [m["t1"] for m in ms],
                             [0.05] * 2, TOL, TOL, max_attempts=3,
                             allow_limit=True, grouped=True)
    for i, m in enumerate(ms):
        one = fresh(mpci, m)
        s1, _ = one.prop_adaptive(0.0, m["t1"], 0.05, TOL, TOL,
                                  max_attempts=3, allow_limit=True)
        for p in range(m["P"]):
            assert st[i][p].attempts == s1[p].attempts == 3, (i, p)
            assert st[i][p].accepted == s1[p].accepted, (i, p)
            assert st[i][p].t == s1[p].t, (i, p)
            assert np.array_equal(props[i].get_state(p), one.get_state(p)), \
                (i, p)
        one.close()
    for pr in props:
        pr.close()
