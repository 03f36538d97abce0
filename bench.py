"""Benchmark of the tdse propagation hot path on one MI355X.

Workload (the headline configuration): synthetic He CI system, L <= 3 with
4000 states per L (N = 16000 complex128 coefficients, dense FP64 dipole
blocks L <-> L+1 of 4000 x 4000), velocity gauge, one sin^2 pulse of
10 cycles, w = 0.7 a.u., E0 = 0.03 a.u., adaptive RKF78 at atol = rtol =
1e-10 from the lowest L = 0 level.  A step is one RKF78 attempt; `value`
is accepted steps per second with all inputs resident in HBM.

The timed region replays the captured step graph once per attempt (the
production mode).  The same attempts are then run once more with a HIP
event pair attached to each coupling-kernel dispatch on the propagator's
stream (hipExtLaunchKernelGGL), which gives the average duration of the
dominant kernel (k_gemv_tile) for the roofline figure (events do not time
kernels inside graph replays).

Extra figures on the same line: the reference mode (fixed dt = 0.1, field
frozen per step, as src/tdse/tdse.cpp:186-204), the batch of 64 pulses on
FP64 MFMA, the uneven multi-instance batch (32 H-shaped + 32 He pulses),
the discrete adjoint, the CI basis producer (w2e stage), the H-shaped
reference-mode case, and the CPU baseline (the oracle's restatement of the
reference with MKL cblas_dgemv, timed on the host cores over a bounded
sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mp-citdse_amd"))
sys.path.insert(0, ROOT)

METRIC = ("RKF78 steps/s, 1 pulse vs batch-64, He L≤3 at tol 1e-10; "
          "adjoint/forward time")
HBM_PEAK_GBS = 8000.0
S_HE = [4000, 4000, 4000, 4000]
W_AU, E0_AU, CYCLES = 0.7, 0.03, 10
TOL = 1e-10


GRAPH = True
PIN_GEMM = None  # "kernel:WMxWN:kKS" pins that autotuner candidate


def make_prop(mpci, sysg, P=1):
    prop = mpci.Propagator(sysg, P=P)
    if not GRAPH:
        prop.set_graph(False)
    if PIN_GEMM and P > 1:
        for i in range(mpci.Propagator.gemm_candidates()):
            cfg = prop.set_gemm_candidate(i)
            if "%s:%dx%d:k%d" % (cfg["kernel"], cfg["wm"], cfg["wn"],
                                 cfg["ksplit"]) == PIN_GEMM:
                break
        else:
            raise SystemExit("--pin-gemm %s: no such candidate" % PIN_GEMM)
    return prop


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def he_system(mpci, synth):
    E, D = synth.system_arrays(S_HE, seed=synth.HE_SEED)
    return E, D, mpci.System(E, D, gauge="v")


def run_gpu(args, mpci, synth, sysg):
    prop = make_prop(mpci, sysg)
    pulse = mpci.Pulse.sine(E0_AU ** 2, W_AU, CYCLES, gauge="v")
    prop.set_pulse(pulse)
    tau = mpci.period(W_AU, CYCLES)
    x0 = synth.ground_state(sysg.N)
    # warmup
    prop.set_state(x0)
    prop.prop_adaptive(0.0, tau, 0.05, TOL, TOL, max_attempts=max(args.warmup, 1),
                       allow_limit=True)
    # timed region: exactly `steps` attempts from t = 0, each attempt one
    # replay of the captured step graph
    prop.set_state(x0)
    t0 = time.perf_counter()
    st, _ = prop.prop_adaptive(0.0, tau, 0.05, TOL, TOL,
                               max_attempts=args.steps, allow_limit=True)
    wall = time.perf_counter() - t0
    # the same attempts again, launched individually, every coupling kernel
    # carrying a HIP event pair on its own dispatch on the propagator's
    # stream (events do not time kernels inside graph replays): the dominant
    # kernel's duration
    prop.set_state(x0)
    prop.set_instrument(True)
    st_i, _ = prop.prop_adaptive(0.0, tau, 0.05, TOL, TOL,
                                 max_attempts=args.steps, allow_limit=True)
    kms, klaunch, kbytes, _ = prop.kernel_time()
    prop.set_instrument(False)
    if st_i[0].accepted != st[0].accepted:
        raise RuntimeError("instrumented pass took a different step sequence")
    # the whole pulse, 0 -> tau, graph-replayed: the controller's accept
    # ratio through the strong-field middle of the pulse, where the timed
    # region above (the first attempts from t = 0) never rejects
    prop.set_state(x0)
    t0 = time.perf_counter()
    sw, _ = prop.prop_adaptive(0.0, tau, 0.05, TOL, TOL)
    wall_w = time.perf_counter() - t0
    whole = {"t0": 0.0, "t1": round(tau, 6), "accepted": sw[0].accepted,
             "rejected": sw[0].rejected, "attempts": sw[0].attempts,
             "wall_s": round(wall_w, 4),
             "steps_per_s": round(sw[0].accepted / wall_w, 3),
             "attempts_per_s": round(sw[0].attempts / wall_w, 3),
             "ms_per_attempt": round(wall_w / sw[0].attempts * 1e3, 4)}
    s = st[0]
    return {"wall": wall, "accepted": s.accepted, "rejected": s.rejected,
            "attempts": s.attempts, "t_reached": s.t, "kernel_ms": kms,
            "kernel_launches": klaunch, "kernel_bytes": kbytes, "prop": prop,
            "tau": tau, "whole_pulse": whole}


S_HE_TEST = [3030, 2799, 2808, 2808]  # inp/He_test.yaml states_in_l


def run_he_test_shape(mpci, synth, attempts=300):
    """The block sizes of the reference's own He input (uneven, ragged
    against the GEMV units), synthetic E and D of that shape, the same
    single pulse and tolerance as the headline: graph-replayed attempts,
    then the same attempts instrumented for the coupling kernel's time."""
    E, D = synth.system_arrays(S_HE_TEST, seed=synth.HE_SEED + 7)
    sysg = mpci.System(E, D, gauge="v")
    prop = make_prop(mpci, sysg)
    prop.set_pulse(mpci.Pulse.sine(E0_AU ** 2, W_AU, CYCLES, gauge="v"))
    tau = mpci.period(W_AU, CYCLES)
    x0 = synth.ground_state(sysg.N)
    prop.set_state(x0)
    prop.prop_adaptive(0.0, tau, 0.05, TOL, TOL, max_attempts=10,
                       allow_limit=True)
    prop.set_state(x0)
    t0 = time.perf_counter()
    st, _ = prop.prop_adaptive(0.0, tau, 0.05, TOL, TOL,
                               max_attempts=attempts, allow_limit=True)
    wall = time.perf_counter() - t0
    prop.set_state(x0)
    prop.set_instrument(True)
    prop.prop_adaptive(0.0, tau, 0.05, TOL, TOL, max_attempts=attempts,
                       allow_limit=True)
    kms, kl, kbytes, _ = prop.kernel_time()
    prop.set_instrument(False)
    avg_us = kms / max(kl, 1) * 1e3
    gbs = kbytes / (avg_us * 1e-6) / 1e9
    return {"states_in_l": S_HE_TEST, "N": sysg.N,
            "dipole_elements": sysg.dipole_elements,
            "accepted": st[0].accepted, "steps_per_s": st[0].accepted / wall,
            "k_gemv_tile_avg_us": avg_us, "achieved_gbs": gbs,
            "frac_of_8tbs": gbs / HBM_PEAK_GBS}


def run_reference_mode(mpci, synth, sysg, steps=200, dt=0.1):
    """Fixed dt, field frozen at t + dt per step (src/tdse/tdse.cpp:187)."""
    prop = make_prop(mpci, sysg)
    pulse = mpci.Pulse.sine(E0_AU ** 2, W_AU, CYCLES, gauge="v")
    import numpy as np
    F, t = np.empty((1, steps)), 0.0
    for k in range(steps):
        F[0, k] = pulse(t + dt)
        t += dt
    prop.set_state(synth.ground_state(sysg.N))
    prop.prop_fixed(0.0, dt, F[:, :10])
    prop.set_state(synth.ground_state(sysg.N))
    t0 = time.perf_counter()
    prop.prop_fixed(0.0, dt, F)
    wall = time.perf_counter() - t0
    return {"steps_per_s": steps / wall, "steps": steps, "dt": dt}


def run_batch(mpci, synth, sysg, P=64, attempts=200):
    """Batched form: P pulses (w ~ U[0.5, 1], E0 log-uniform in [0.005,
    0.05], 4-20 cycles) on the He system, per-pulse adaptive control; a
    bounded number of lockstep attempts from t = 0."""
    import numpy as np
    w, E0, cyc = synth.pulse_batch(P)
    prop = make_prop(mpci, sysg, P=P)
    x0 = synth.ground_state(sysg.N)
    t1 = np.array([mpci.period(w[p], int(cyc[p])) for p in range(P)])
    for p in range(P):
        prop.set_pulse(mpci.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p])), p)
    for p in range(P):
        prop.set_state(x0, p)
    prop.prop_adaptive(0.0, t1, 0.05, TOL, TOL, max_attempts=10,
                       allow_limit=True)
    for p in range(P):
        prop.set_state(x0, p)
    t0 = time.perf_counter()
    st, _ = prop.prop_adaptive(0.0, t1, 0.05, TOL, TOL, max_attempts=attempts,
                               allow_limit=True)
    wall = time.perf_counter() - t0
    # same attempts, instrumented: duration of the MFMA coupling kernel
    for p in range(P):
        prop.set_state(x0, p)
    prop.set_instrument(True)
    prop.prop_adaptive(0.0, t1, 0.05, TOL, TOL, max_attempts=attempts,
                       allow_limit=True)
    kms, kl, _, kflops = prop.kernel_time()
    prop.set_instrument(False)
    acc = sum(x.accepted for x in st)
    avg_ms = kms / max(kl, 1)
    return {"pulse_steps_per_s": acc / wall, "P": P, "attempts": attempts,
            "accepted_total": acc, "wall_s": wall,
            "gemm": prop.gemm_config(),
            "mfma_tflops": kflops / (avg_ms * 1e-3) / 1e12 if kl else None,
            "avg_launch_us": avg_ms * 1e3}


S_H = [998] * 21
H_SEED_OFFSET = 1


def run_multi(mpci, synth, sysg, P=32, attempts=120):
    """Uneven multi-instance batch: P pulses on an H-shaped synthetic system
    (21 blocks of 998 states, E ~ U[-0.5, 5]) and P pulses on the He
    system, pulses drawn as for the batch line, propagated concurrently by
    mpci_multi_adaptive for a bounded number of attempts; the same two
    batches are then run one after the other for comparison."""
    import numpy as np
    E, D = synth.system_arrays(S_H, seed=synth.HE_SEED + H_SEED_OFFSET,
                               e_lo=-0.5, e_hi=5.0)
    sysh = mpci.System(E, D, gauge="v")
    del E, D
    members = []
    for g, seed in ((sysh, synth.HE_SEED + 5), (sysg, synth.HE_SEED + 6)):
        w, E0, cyc = synth.pulse_batch(P, seed=seed)
        t1 = np.array([mpci.period(w[p], int(cyc[p])) for p in range(P)])
        prop = make_prop(mpci, g, P=P)
        for p in range(P):
            prop.set_pulse(mpci.Pulse.sine(E0[p] ** 2, w[p], int(cyc[p])), p)
        members.append((g, prop, t1))

    def reset():
        for g, prop, _ in members:
            x0 = synth.ground_state(g.N)
            for p in range(P):
                prop.set_state(x0, p)

    props = [m[1] for m in members]
    t1s = [m[2] for m in members]
    reset()
    mpci.multi_adaptive(props, [0.0, 0.0], t1s, [0.05, 0.05], TOL, TOL,
                        max_attempts=5, allow_limit=True)
    reset()
    t0 = time.perf_counter()
    st = mpci.multi_adaptive(props, [0.0, 0.0], t1s, [0.05, 0.05], TOL, TOL,
                             max_attempts=attempts, allow_limit=True)
    wall = time.perf_counter() - t0
    reset()
    seq = []
    for prop, t1 in zip(props, t1s):
        t0 = time.perf_counter()
        prop.prop_adaptive(0.0, t1, 0.05, TOL, TOL, max_attempts=attempts,
                           allow_limit=True)
        seq.append(time.perf_counter() - t0)
    # every member's attempt in the same launches (one interleaved GEMM job
    # list, one stage / finalize / apply launch per attempt)
    reset()
    mpci.multi_adaptive(props, [0.0, 0.0], t1s, [0.05, 0.05], TOL, TOL,
                        max_attempts=5, allow_limit=True, grouped=True)
    reset()
    t0 = time.perf_counter()
    stg = mpci.multi_adaptive(props, [0.0, 0.0], t1s, [0.05, 0.05], TOL, TOL,
                              max_attempts=attempts, allow_limit=True,
                              grouped=True)
    wall_g = time.perf_counter() - t0
    flops = 0.0
    for (g, _, _), sm in zip(members, st):
        flops += sum(x.attempts for x in sm) * 13 * 8.0 * g.dipole_elements
    acc = [sum(x.accepted for x in sm) for sm in st]
    if [sum(x.accepted for x in sm) for sm in stg] != acc:
        raise RuntimeError("grouped run took a different step sequence")
    tf = flops / wall / 1e12
    tfg = flops / wall_g / 1e12
    return {"P_h": P, "P_he": P, "attempts": attempts,
            "h_blocks": len(S_H), "h_states_per_block": S_H[0],
            "accepted_h": acc[0], "accepted_he": acc[1],
            "pulse_steps_per_s": sum(acc) / wall, "wall_s": wall,
            "sequential_wall_s": sum(seq),
            "overlap_speedup": sum(seq) / wall,
            "mfma_tflops_wall": tf, "mfma_frac_wall": tf / 78.6,
            "mode": "per-stream members (each on its own stream and graph)",
            "grouped_wall_s": wall_g,
            "grouped_overlap_speedup": sum(seq) / wall_g,
            "grouped_mfma_frac_wall": tfg / 78.6,
            "grouped_mode": "one launch per kernel of the attempt over both "
                            "members (one GEMM job list, shortest K first)"}


S_C1 = [198] * 4


def c1_inputs(mpci, synth):
    """C1: H-shaped synthetic system (4 blocks of 198 states, E ~ U[-0.5,
    5]), velocity gauge, w = 0.5, E0 = 0.05, 10 cycles, reference mode at
    dt = 0.1 over the whole pulse (steps = int(tau / dt) as
    src/tdse/tdse_main.cpp:164)."""
    E, D = synth.system_arrays(S_C1, seed=synth.HE_SEED + 2, e_lo=-0.5,
                               e_hi=5.0)
    w, e0, cyc, dt = 0.5, 0.05, 10, 0.1
    steps = int(mpci.period(w, cyc) / dt)
    return E, D, w, e0, cyc, dt, steps


def run_c1(mpci, synth, c1):
    """C1 on the GPU: the full reference-mode pulse (fixed dt, field frozen
    per step), launch-bound at this size (every D_L is cache-resident)."""
    import numpy as np
    E, D, w, e0, cyc, dt, steps = c1
    g = mpci.System(E, D, gauge="v")
    pulse = mpci.Pulse.sine(e0 ** 2, w, cyc, gauge="v")
    F, t = np.empty((1, steps)), 0.0
    for k in range(steps):
        F[0, k] = pulse(t + dt)
        t += dt
    x0 = synth.ground_state(g.N)
    rates = {}
    for mode in (-1, 0):  # the default (fused small-system path), the tiles
        prop = make_prop(mpci, g)
        on = prop.set_small_path(mode)
        prop.set_state(x0)
        prop.prop_fixed(0.0, dt, F[:, :20])
        prop.set_state(x0)
        t0 = time.perf_counter()
        prop.prop_fixed(0.0, dt, F)
        rates["small_path" if on else "tile_path"] = \
            steps / (time.perf_counter() - t0)
        prop.close()
    return {"steps": steps,
            "steps_per_s": round(rates["small_path"], 1),
            "tile_path_steps_per_s": round(rates["tile_path"], 1),
            "path": "fused small-system stage kernel (one launch per stage)",
            "bound": "launch/latency: 0.94 MB of dipoles stay in cache, so "
                     "no HBM roofline applies"}


S_W2E = [3000, 3000]


def run_w2e(mpci, synth, arrays):
    """CI basis producer (the w2e stage) on the GPU: two L blocks of 3000
    product states (dsyevd each) and the dipole pair between them, from host
    arrays to host arrays (the stage reads and writes files)."""
    from mpci import ci
    ens, v12, dip = arrays
    ci.ci_basis(ens[0][:64], v12[0][:64 * 65 // 2])  # warm the libraries
    t0 = time.perf_counter()
    V = []
    for L in range(len(ens)):
        V.append(ci.ci_basis(ens[L], v12[L])[1])
    t_basis = time.perf_counter() - t0
    t1 = time.perf_counter()
    ci.ci_dipoles(dip[0], V[0], V[1])
    t_dip = time.perf_counter() - t1
    n = S_W2E[0]
    return {"sizes": S_W2E, "basis_s": round(t_basis, 3),
            "dipoles_s": round(t_dip, 4),
            "total_s": round(t_basis + t_dip, 3),
            "dipole_gemm_tflops": round(
                2 * 2.0 * n * n * n / t_dip / 1e12, 2)}


D1E_NEN, D1E_Q, D1E_NL = 400, 1000, 3


def run_d1e(mpci, arrays):
    """One-electron dipole producer (the d1e stage) on the GPU: 400 radial
    functions per l on 1000 quadrature nodes, l = 0..2, velocity gauge, host
    arrays in and out."""
    from mpci import ci
    qx, qw, po, pp, _ = arrays
    ci.d1e_dipoles(0, "v", qx, qw, po[0][:8], po[1][:8], pp[1][:8])  # warm
    t0 = time.perf_counter()
    for l in range(D1E_NL - 1):
        ci.d1e_dipoles(l, "v", qx, qw, po[l], po[l + 1], pp[l + 1])
    return {"nen": D1E_NEN, "quadrature": D1E_Q, "l_blocks": D1E_NL,
            "gauge": "v", "total_s": round(time.perf_counter() - t0, 4)}


def run_adjoint(mpci, synth, sysg, B=16, K=256, cycles=20, w=0.7):
    """Discrete-adjoint gradients of the final ground-state population with
    respect to K piecewise-linear envelope samples (uniform in [0, 0.04]
    a.u.) for a batch of B envelopes, checkpoints every 32 steps."""
    import numpy as np
    T = mpci.period(w, cycles)
    prop = make_prop(mpci, sysg, P=B)
    x0 = synth.ground_state(sysg.N)
    W = np.zeros(sysg.N)
    W[0] = 1.0
    for p in range(B):
        prop.set_state(x0, p)
        prop.set_pulse(mpci.Pulse.pwl(synth.uniform(77 + p, 0, K, 0.0, 0.04),
                                      w, T), p)
    J, grad, nst, _, st = prop.adjoint(0.0, T, 0.05, W, ckpt_every=32)
    return {"B": B, "samples": K, "cycles": cycles, "steps_max": st.steps_max,
            "fwd_ms": st.fwd_ms, "bwd_ms": st.bwd_ms,
            "adjoint_over_forward": st.bwd_ms / st.fwd_ms}


def host_cores():
    """(logical CPUs, physical cores) of the host, the latter from lscpu."""
    import subprocess
    logical = os.cpu_count() or 1
    try:
        out = subprocess.run(["lscpu", "-p=Core,Socket"], capture_output=True,
                             text=True, timeout=10).stdout
        pairs = {ln for ln in out.splitlines() if ln and not ln.startswith("#")}
        physical = len(pairs) or None
    except (OSError, subprocess.SubprocessError):
        physical = None
    return logical, physical


def mkl_set_threads(path, n):
    import ctypes
    try:
        ctypes.CDLL(path).MKL_Set_Num_Threads(ctypes.c_int(n))
        return True
    except (OSError, AttributeError):
        return False


def cpu_baseline(E, D, c1, w2e=None, d1e=None, budget_s=12.0):
    """The oracle (reference call pattern: 4 cblas_dgemv per block per RHS,
    stride 2, boost do_step stage algebra) with MKL on the host cores, then
    on one thread; and the C1 H-shaped reference-mode pulse (c1 = its system
    arrays, pulse and step table) at both thread counts."""
    import numpy as np
    from oracle import oracle as orc
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") \
        else (os.cpu_count() or 1)
    cores = min(cores, 16)
    os.environ.setdefault("MKL_NUM_THREADS", str(cores))
    blas, mkl = "builtin loops (1 thread)", None
    for path in ("/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"):
        if os.path.exists(path) and orc.use_blas(path):
            blas, mkl = f"MKL cblas_dgemv ({path})", path
            break
    else:
        cores = 1
    if mkl:
        mkl_set_threads(mkl, cores)
    osys = orc.System(E, D, gauge="v")
    pulse = orc.Pulse.sine(E0_AU ** 2, W_AU, CYCLES, gauge="v")
    tau = orc.period(W_AU, CYCLES)
    x = np.zeros(2 * osys.N)
    x[0] = 1.0

    def sample(budget, cap):
        # one step to size the sample, then as many attempts as fit
        t0 = time.perf_counter()
        osys.step_stages(np.zeros(13), x, 0.05)
        one = time.perf_counter() - t0
        n = max(2, min(cap, int(budget / max(one, 1e-3))))
        t0 = time.perf_counter()
        _, sto, _ = osys.prop_adaptive(pulse, 0.0, tau, 0.05, TOL, TOL, x,
                                       max_attempts=n, allow_limit=True)
        return n, sto.accepted, time.perf_counter() - t0

    n, acc, wall = sample(budget_s, 50)
    out = {"value": round(acc / wall, 3), "unit": "steps/s", "cores": cores,
           "kind": "port",
           "sample": f"first {n} adaptive RKF78 attempts of the same He pulse "
                     f"(tol 1e-10) with the oracle, {blas}, {wall:.1f} s"}
    logical, physical = host_cores()
    out["host_logical_cpus"] = logical
    out["host_physical_cores"] = physical
    if mkl and mkl_set_threads(mkl, 1):
        n1, acc1, wall1 = sample(budget_s / 2, 10)
        out["value_1thread"] = round(acc1 / wall1, 3)
        out["sample_1thread"] = f"first {n1} attempts, {wall1:.1f} s"
        if acc1 / wall1 >= 0.9 * out["value"]:
            # the reference's stride-2 cblas_dgemv calls do not use MKL's
            # threads: the sample ran on one core whatever was allowed
            out["sample"] += (f"; MKL allowed {cores} threads, but the "
                              "stride-2 dgemv of the reference's call pattern "
                              "runs on one (the 1-thread rate is the same)")
            out["cores"] = 1
        mkl_set_threads(mkl, cores)
    # C1: the H-shaped reference-mode pulse, cache-resident (0.94 MB of D)
    Eh, Dh, w, e0, cyc, dt, steps = c1
    oh = orc.System(Eh, Dh, gauge="v")
    ph = orc.Pulse.sine(e0 ** 2, w, cyc, gauge="v")
    xh = np.zeros(2 * oh.N)
    xh[0] = 1.0
    for thr in ((cores, 1) if mkl else (1,)):
        if mkl:
            mkl_set_threads(mkl, thr)
        t0 = time.perf_counter()
        oh.prop_fixed(ph, 0.0, dt, steps, xh)
        out[f"c1_h4x198_steps_per_s_{thr}thr"] = round(
            steps / (time.perf_counter() - t0), 1)
    if mkl:
        mkl_set_threads(mkl, cores)
    orc.use_builtin_dgemv()
    if w2e is not None:
        # the w2e stage on the CPU restatement (numpy LAPACK/BLAS, all host
        # threads numpy uses)
        from oracle import ci_oracle as co
        ens, v12, dip = w2e
        t0 = time.perf_counter()
        V = [co.ci_basis(ens[L], v12[L])[1] for L in range(len(ens))]
        co.ci_dipoles(dip[0], V[0], V[1])
        out["w2e_3000x2_s"] = round(time.perf_counter() - t0, 3)
    if d1e is not None:
        from oracle import ci_oracle as co
        qx, qw, po, pp, _ = d1e
        t0 = time.perf_counter()
        for l in range(D1E_NL - 1):
            co.d1e_dipoles(l, "v", qx, qw, po[l], po[l + 1], pp[l + 1])
        out["d1e_400x1000_s"] = round(time.perf_counter() - t0, 3)
    return out


def copy_calibration(n_bytes=1 << 30, reps=3):
    """A plain elementwise copy (torch.mul by 1.0, out=) of n_bytes: each
    launch reads and writes exactly n_bytes, which scales the hardware byte
    counters for the kernels measured in the same profiled run."""
    import torch
    x = torch.ones(n_bytes // 8, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    return {"bytes_read": n_bytes, "bytes_written": n_bytes, "launches": reps,
            "kernel": "torch elementwise mul (copy)"}


def measured_ceilings(mpci, stream_bytes):
    """The ceilings this device reaches in practice, beside the published
    peaks: a non-temporal streaming read over a buffer the size of the
    dipole store, and back-to-back FP64 MFMA with no memory traffic."""
    return {"hbm_read_gbs": mpci.probe_hbm_read(stream_bytes, 20),
            "hbm_probe": f"non-temporal 16-byte read of {stream_bytes} bytes "
                         "(the dipole store's size), best of 2/4/8 "
                         "workgroups per CU",
            "mfma_f64_tflops": mpci.probe_mfma_f64(16384),
            "mfma_probe": "v_mfma_f64_16x16x4f64, 4 independent chains per "
                          "wave, 2 waves per SIMD, no memory traffic"}


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "k_gemv_tile_he4x4000_traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-multi", action="store_true")
    ap.add_argument("--no-reference-mode", action="store_true")
    ap.add_argument("--no-he-test-shape", action="store_true",
                    help="skip the leg on the block sizes of the "
                         "reference's He input")
    ap.add_argument("--no-w2e", action="store_true",
                    help="skip the w2e and d1e producer legs")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every attempt eagerly instead of replaying "
                         "the captured step graph (for tools that cannot "
                         "follow graph launches)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed run, launch a plain copy kernel "
                         "of 1 GiB read + 1 GiB write (scale reference for "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--pin-gemm", default=None, metavar="KERNEL:WMxWN:kKS",
                    help="batched legs: pin this candidate of the GEMM "
                         "autotuner (e.g. k_gemm_pair:16x1:k1) instead of its "
                         "choice, for profiling one kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this implementation runs on a "
              "single MI355X (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    global GRAPH, PIN_GEMM
    GRAPH = not args.no_graph
    PIN_GEMM = args.pin_gemm
    import mpci
    from mpci import synth
    log("bench: building He system (4 x 4000) ...")
    E, D, sysg = he_system(mpci, synth)
    log("bench: adaptive single-pulse run ...")
    r = run_gpu(args, mpci, synth, sysg)
    value = r["accepted"] / r["wall"]
    log(f"bench: {r['accepted']} accepted / {r['attempts']} attempts in "
        f"{r['wall']:.3f} s")
    avg_us = r["kernel_ms"] / max(r["kernel_launches"], 1) * 1e3
    achieved = r["kernel_bytes"] / (avg_us * 1e-6) / 1e9
    log("bench: measured ceilings (HBM read, FP64 MFMA) ...")
    ceil = measured_ceilings(mpci, r["kernel_bytes"])
    ref = None
    if not args.no_reference_mode:
        log("bench: reference mode ...")
        ref = run_reference_mode(mpci, synth, sysg)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / max(r["attempts"], 1) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded He-shaped CI system; no reference outputs "
                "ship)",
        "config": {"workload": "He L<=3 synthetic CI, 4 x 4000 states, dense "
                               "FP64 dipoles, velocity gauge, 1 sin^2 pulse "
                               "(w=0.7, E0=0.03, 10 cycles), adaptive RKF78 "
                               "tol 1e-10",
                   "N": sysg.N, "dipole_elements": sysg.dipole_elements,
                   "accepted": r["accepted"], "rejected": r["rejected"]},
        "roofline": {"bound": "hbm", "kernel": "k_gemv_tile",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(),
                     "measured_peak": round(ceil["hbm_read_gbs"], 1),
                     "frac_of_measured": round(
                         achieved / ceil["hbm_read_gbs"], 4),
                     "measured_peak_probe": ceil["hbm_probe"],
                     "avg_launch_us": round(avg_us, 3),
                     "bytes_per_launch": r["kernel_bytes"],
                     "launches": r["kernel_launches"]},
    }
    out["whole_pulse"] = r["whole_pulse"]
    if ref is not None:
        out["reference_mode"] = {"steps_per_s": round(ref["steps_per_s"], 3),
                                 "dt": ref["dt"], "steps": ref["steps"]}
    if not args.no_he_test_shape:
        log("bench: He_test.yaml block sizes (3030, 2799, 2808, 2808) ...")
        h = run_he_test_shape(mpci, synth)
        out["he_test_shape"] = {k: (round(v, 4) if isinstance(v, float)
                                    else v) for k, v in h.items()}
    if not args.no_batch:
        log("bench: batch-64 ...")
        b = run_batch(mpci, synth, sysg)
        b["per_pulse_speedup_vs_single"] = b["pulse_steps_per_s"] / value
        if b["mfma_tflops"] is not None:
            b["mfma_frac"] = b["mfma_tflops"] / 78.6
            b["mfma_measured_peak_tflops"] = ceil["mfma_f64_tflops"]
            b["mfma_frac_of_measured"] = (b["mfma_tflops"] /
                                          ceil["mfma_f64_tflops"])
            b["mfma_probe"] = ceil["mfma_probe"]
        out["batch64"] = {k: (round(v, 4) if isinstance(v, float) else v)
                          for k, v in b.items()}
        # 16 pulses: the adjoint workload's batch size, where the dipole
        # stream rather than the MFMA work bounds the coupling
        log("bench: batch-16 ...")
        b = run_batch(mpci, synth, sysg, P=16)
        if b["mfma_tflops"] is not None:
            b["mfma_frac"] = b["mfma_tflops"] / 78.6
        out["batch16"] = {k: (round(v, 4) if isinstance(v, float) else v)
                          for k, v in b.items()}
    if not args.no_multi:
        log("bench: multi-instance 32 H (21 x 998) + 32 He ...")
        mres = run_multi(mpci, synth, sysg)
        out["multi_h32_he32"] = {k: (round(v, 4) if isinstance(v, float)
                                     else v) for k, v in mres.items()}
    if not args.no_adjoint:
        log("bench: adjoint (16 envelopes x 256 samples) ...")
        a = run_adjoint(mpci, synth, sysg)
        out["adjoint"] = {k: (round(v, 4) if isinstance(v, float) else v)
                          for k, v in a.items()}
        out["adjoint_over_forward"] = round(a["adjoint_over_forward"], 4)
    c1 = c1_inputs(mpci, synth)
    w2e = None
    if not args.no_w2e:
        log("bench: w2e CI basis (2 x 3000) ...")
        w2e = synth.w2e_arrays(S_W2E)
        out["w2e_ci_basis"] = run_w2e(mpci, synth, w2e)
    d1e = None
    if not args.no_w2e:
        log("bench: d1e dipoles (400 x 1000, 3 l) ...")
        d1e = synth.w1e_arrays(D1E_NEN, D1E_Q, D1E_NL)
        out["d1e_dipoles"] = run_d1e(mpci, d1e)
    if not args.no_reference_mode:
        log("bench: C1 (H 4 x 198, reference mode, full pulse) ...")
        out["c1_h4x198"] = run_c1(mpci, synth, c1)
    if args.calibrate_copy:
        out["copy_calibration"] = copy_calibration()
    if not args.no_cpu_baseline:
        log("bench: CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(E, D, c1, w2e, d1e)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
