"""Discrete-adjoint gradients (extension: the reference has no gradients)
checked against central finite differences with h = 1e-6 of the same
discrete map on a frozen step schedule, both on the GPU (schedule replay
through the C ABI) and on the CPU oracle (or_prop_schedule).  Bar: relative
error of the gradient vector <= 1e-6."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FD_H = 1e-6


@pytest.fixture(scope="module")
def mpci():
    import mpci as m
    assert m.device_count() >= 1
    return m


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def setup(mpci, oracle, s, gauge, seed):
    from mpci import synth
    E, D = synth.system_arrays(s, seed=seed)
    return E, mpci.System(E, D, gauge=gauge), oracle.System(E, D, gauge=gauge)


def yield_mask(E, thr=-2.0):
    """pes::genPES2e yield mask (src/pes/pes.cpp:199-205)."""
    W = []
    for e in E:
        w = np.zeros(len(e))
        for i in range(1, len(e) - 1):
            w[i] = 1.0 if e[i] - thr > 0 else 0.0
        W.append(w)
    return np.concatenate(W)


def gpu_fd(mpci, g, env, w, T, gauge, hs, W, x0):
    """Central FD of J over every envelope sample, all perturbations in one
    batched schedule replay (2 K pulses)."""
    K = len(env)
    P = 2 * K
    prop = mpci.Propagator(g, P=P)
    keep = []
    for k in range(K):
        for sgn, p in ((1, 2 * k), (-1, 2 * k + 1)):
            e = env.copy()
            e[k] += sgn * FD_H
            pl = mpci.Pulse.pwl(e, w, T, gauge=gauge)
            keep.append(pl)
            prop.set_pulse(pl, p)
            prop.set_state(x0, p)
    prop.prop_schedule(0.0, np.tile(hs, (P, 1)))
    J = prop.populations(W)[0]
    return (J[0::2] - J[1::2]) / (2 * FD_H)


@pytest.mark.parametrize("gauge,objective", [("v", "ground"), ("l", "yield"),
                                             ("v", "yield")])
def test_adjoint_vs_finite_differences(mpci, oracle, gauge, objective):
    from mpci import synth
    E, g, o = setup(mpci, oracle, [24, 30, 18], gauge, seed=13)
    w, cyc = 0.7, 3
    T = cyc * 2 * math.pi / w
    env = synth.uniform(5, 0, 12, 0.0, 0.04)
    x0 = synth.ground_state(g.N)
    W = np.zeros(g.N)
    if objective == "ground":
        W[0] = 1.0
    else:
        W = yield_mask(E)
    prop = mpci.Propagator(g)
    prop.set_state(x0)
    prop.set_pulse(mpci.Pulse.pwl(env, w, T, gauge=gauge))
    J, grad, nst, hlog, st = prop.adjoint(0.0, T, 0.05, W, ckpt_every=7,
                                          hlog_cap=100000)
    hs = hlog[0, :nst[0]]
    assert st.segments == math.ceil(nst[0] / 7) and st.segments >= 3
    # J matches a schedule replay and the oracle
    xo = o.prop_schedule(oracle.Pulse.pwl(env, w, T, gauge), 0.0, hs, x0)
    Jo = float(W @ (xo[0::2] ** 2 + xo[1::2] ** 2))
    assert abs(J[0] - Jo) < 1e-12
    fd = gpu_fd(mpci, g, env, w, T, gauge, hs, W, x0)
    assert rel(grad[0], fd) < 1e-6, (grad[0], fd)
    # finite differences of the CPU restatement of the reference map
    fdo = np.empty(len(env))
    for k in range(len(env)):
        J2 = []
        for sgn in (1, -1):
            e = env.copy()
            e[k] += sgn * FD_H
            xk = o.prop_schedule(oracle.Pulse.pwl(e, w, T, gauge), 0.0, hs, x0)
            J2.append(float(W @ (xk[0::2] ** 2 + xk[1::2] ** 2)))
        fdo[k] = (J2[0] - J2[1]) / (2 * FD_H)
    assert rel(grad[0], fdo) < 1e-6


@pytest.mark.parametrize("pin", [None, "k_gemm_pair"])
def test_batched_adjoint_uneven_pulses(mpci, oracle, pin):
    """Four envelopes with different lengths in one batched adjoint (with
    the autotuned coupling, and with k_gemm_pair pinned: its forward and
    transposed products both run in the adjoint's recompute and sweep)."""
    from mpci import synth
    E, g, o = setup(mpci, oracle, [32, 40, 28, 20], "v", seed=17)
    P = 4
    w = 0.7
    Ts = np.array([2, 3, 2.5, 4]) * 2 * math.pi / w
    envs = [synth.uniform(100 + p, 0, 10, 0.0, 0.04) for p in range(P)]
    x0 = synth.ground_state(g.N)
    W = np.zeros(g.N)
    W[0] = 1.0
    prop = mpci.Propagator(g, P=P)
    if pin:
        for i in range(mpci.Propagator.gemm_candidates()):
            cfg = prop.set_gemm_candidate(i)
            if cfg["kernel"] == pin and cfg["ksplit"] == 1:
                break
        assert prop.gemm_config()["kernel"] == pin
    for p in range(P):
        prop.set_state(x0, p)
        prop.set_pulse(mpci.Pulse.pwl(envs[p], w, Ts[p]), p)
    J, grad, nst, hlog, st = prop.adjoint(0.0, Ts, 0.05, W, ckpt_every=9,
                                          hlog_cap=100000)
    assert len(set(nst.tolist())) > 1
    for p in range(P):
        single = mpci.Propagator(g)
        single.set_state(x0)
        single.set_pulse(mpci.Pulse.pwl(envs[p], w, Ts[p]))
        J1, g1, n1, _, _ = single.adjoint(0.0, Ts[p], 0.05, W, ckpt_every=9)
        assert n1[0] == nst[p]
        # the batch couples on MFMA, the single pulse on the GEMV kernel:
        # their sums differ in the last bits, the accepted step sizes then
        # differ by ~1e-7 relative (the embedded error cancels ~10 digits),
        # the final states by up to ~1e-10 (the adaptive-mode bar), and the
        # gradients of the two discrete maps (different schedules) by up to
        # that relative step difference; each is held to its own finite
        # differences at 1e-6 below
        assert abs(J1[0] - J[p]) < 1e-10
        assert rel(grad[p], g1[0]) < 1e-6
        fd = gpu_fd(mpci, g, envs[p], w, Ts[p], "v", hlog[p, :nst[p]], W, x0)
        assert rel(grad[p], fd) < 1e-6


def test_torch_autograd_wrapper(mpci):
    """torch.autograd.Function over the C ABI: backward returns the adjoint
    gradient scaled by the incoming gradient."""
    import torch
    from mpci import synth
    from mpci.autograd import ObjectiveSpec, PulseObjective
    E, D = synth.system_arrays([20, 24, 16], seed=3)
    g = mpci.System(E, D, gauge="v")
    P, K, w = 2, 8, 0.7
    T = 2 * 2 * math.pi / w
    prop = mpci.Propagator(g, P=P)
    W = np.zeros(g.N)
    W[0] = 1.0
    spec = ObjectiveSpec(prop=prop, x0=synth.ground_state(g.N), w=w, T=T, W=W,
                         ckpt_every=5)
    env = torch.tensor(np.stack([synth.uniform(40 + p, 0, K, 0.0, 0.04)
                                 for p in range(P)]), requires_grad=True)
    J = PulseObjective.apply(env, spec)
    (J * torch.tensor([1.0, -2.0], dtype=J.dtype)).sum().backward()
    ref = mpci.Propagator(g, P=P)
    for p in range(P):
        ref.set_state(spec.x0, p)
        ref.set_pulse(mpci.Pulse.pwl(env[p].detach().numpy(), w, T), p)
    Jr, gr, _, _, _ = ref.adjoint(0.0, T, 0.05, W, ckpt_every=5)
    assert np.allclose(J.detach().numpy(), Jr, rtol=0, atol=1e-14)
    assert np.allclose(env.grad.numpy()[0], gr[0], rtol=1e-12, atol=0)
    assert np.allclose(env.grad.numpy()[1], -2.0 * gr[1], rtol=1e-12, atol=0)


def test_checkpoint_store_sized_by_the_run(mpci):
    """The checkpoint store grows with the forward pass: its slots follow
    the accepted steps taken, not max_steps (whose default would ask for
    100000 / K slots)."""
    from mpci import synth
    E, D = synth.system_arrays([30, 24, 20], seed=19)
    g = mpci.System(E, D, gauge="v")
    prop = mpci.Propagator(g)
    w, T, K = 0.7, 30.0, 8
    prop.set_state(synth.ground_state(g.N))
    prop.set_pulse(mpci.Pulse.pwl(synth.uniform(9, 0, 16, 0.0, 0.04), w, T))
    W = np.zeros(g.N)
    W[0] = 1.0
    J, grad, nst, _, st = prop.adjoint(0.0, T, 0.05, W, ckpt_every=K,
                                       max_steps=100000)
    assert st.segments == math.ceil(nst[0] / K)
    assert nst[0] // K + 2 <= st.ckpt_slots < 100000 // K
    # the slots a chunk can fill, at most one chunk of attempts ahead
    assert st.ckpt_slots <= (nst[0] + 4096) // K + 2
    assert st.peak_bytes >= st.ckpt_slots * g.N * 16
