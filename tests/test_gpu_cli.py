"""End-to-end GPU runs of the `tdse` drop-in (src/tdse/tdse_main.cpp:9-201)
on reference-format HDF5 inputs written by tdse_gen, against the oracle on
the same leading blocks: coefficient-file names and cadence (t0, after step
0, every steps/10, final), the field file (default precision) and the
population file (16 digits), full-precision final states, restart with
-s, the adaptive extension, and the 1e file prefixes with a gaussian
length-gauge pulse."""
import math
import os
import subprocess
import textwrap

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mp-citdse_amd", "bin")
EV = 1.0 / 27.211396181
IO = (1.0 / 3.509338) * 1.0e-16


def yaml_2e(states, gauge="v", shape="sine", cycles=2, extra=""):
    return textwrap.dedent(f"""\
        Global_Settings:
          potential: "he"
          L_max: {len(states) - 1}
          gauge: '{gauge}'
        Basis_Settings:
          l_max: 7
        Propagator_Settings:
          states_in_l: [{",".join(map(str, states))}]
          dt: 0.1
          full_precision: true
        {extra}Field_Parameters:
          shape: "{shape}"
          w: 15
          Io: 1.0e14
          cepd: 0.0
          cycles: {cycles}
        """)


def read_state(path, n):
    v = np.loadtxt(path, comments="#")
    assert v.shape == (n, 3)
    x = np.empty(2 * n)
    x[0::2], x[1::2] = v[:, 1], v[:, 2]
    return x


def run(args, cwd):
    r = subprocess.run([os.path.join(BIN, "tdse")] + args, cwd=cwd,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def leading(full, states, seed=20261202):
    from mpci import synth
    E, D = synth.system_arrays(full, seed=seed)
    E = [E[L][:n] for L, n in enumerate(states)]
    D = [D[L][:states[L], :states[L + 1]] for L in range(len(states) - 1)]
    return E, D


def test_cli_2e_reference_mode(tmp_path, oracle):
    full, states = [40, 35, 30], [30, 35, 25]
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", ",".join(map(str, full)), "-g", "v"], cwd=tmp_path,
                   check=True)
    (tmp_path / "He.yaml").write_text(yaml_2e(states))
    r = run(["-e", "2", "-f", "He.yaml", "-o", "out"], tmp_path)
    E, D = leading(full, states)
    o = oracle.System(E, D, gauge="v")
    w = 15 * EV
    tau = 2 * 2.0 * math.pi / w
    steps = int(tau / 0.1)
    pulse = oracle.Pulse.sine(1.0e14 * IO, w, 2, gauge="v")
    x0 = np.zeros(2 * o.N)
    x0[0] = 1.0
    xr, tr, popr, fldr = o.prop_fixed(pulse, 0.0, 0.1, steps, x0, track=0)
    out = tmp_path / "out"
    # cadence: t0, after step 0, every steps/10, final
    pr, t, want = steps // 10, 0.0, {"%f" % 0.0}
    for st in range(steps):
        t += 0.1
        if st % pr == 0:
            want.add("%f" % t)
    want.add("%f" % t)
    got = {f[len("he_ct_"):-4] for f in os.listdir(out)
           if f.startswith("he_ct_")}
    assert got == want
    # final state at full precision
    x = read_state(out / ("he_ct17_%f.dat" % t), o.N)
    assert np.abs(x - xr).max() < 1e-12
    # population file: 16 digits, one line per step plus t0
    pop = np.loadtxt(out / "he_pop10.dat", comments="#")
    assert pop.shape == (steps + 1, 2)
    assert np.abs(pop[1:, 1] - popr).max() < 1e-12
    # field file: default ostream precision, F(t + dt) as the reference
    lines = (out / "he_field.dat").read_text().splitlines()
    assert lines[0] == "#time (a.u.), field (A(t) or E(t)) (a.u.)"
    tt = 0.0
    for k in range(steps):
        assert lines[k + 1] == f"{tt + 0.1:.6g} {fldr[k]:.6g}"
        tt += 0.1
    # printed fields at the coefficient-file steps
    printed = [float(v) for v in r.stdout.split("number of cycles:")[1]
               .split()[1:]]
    assert len(printed) == len(range(0, steps, pr))


def test_cli_restart(tmp_path, oracle):
    full, states = [30, 30, 30], [30, 30, 30]
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", "30,30,30", "-g", "l"], cwd=tmp_path, check=True)
    (tmp_path / "He.yaml").write_text(yaml_2e(states, gauge="l"))
    run(["-e", "2", "-f", "He.yaml", "-o", "out"], tmp_path)
    ct = sorted(f for f in os.listdir(tmp_path / "out")
                if f.startswith("he_ct_"))[3]
    t_start = float(ct[len("he_ct_"):-4])
    r = run(["-e", "2", "-f", "He.yaml", "-o", "out2", "-s", f"out/{ct}"],
            tmp_path)
    assert f"Start time (a.u.): {t_start:g}" in r.stdout
    E, D = leading(full, states)
    o = oracle.System(E, D, gauge="l")
    w = 15 * EV
    tau = 2 * 2.0 * math.pi / w
    steps = int((tau - t_start) / 0.1)
    x0 = read_state(tmp_path / "out" / ct, o.N)  # 6-digit restart values
    pulse = oracle.Pulse.sine(1.0e14 * IO, w, 2, gauge="l")
    xr, tr, _, _ = o.prop_fixed(pulse, t_start, 0.1, steps, x0)
    x = read_state(tmp_path / "out2" / ("he_ct17_%f.dat" % tr), o.N)
    assert np.abs(x - xr).max() < 1e-12


def test_cli_adaptive_extension(tmp_path, oracle):
    full = states = [26, 31, 22]
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", "26,31,22", "-g", "v"], cwd=tmp_path, check=True)
    (tmp_path / "He.yaml").write_text(
        yaml_2e(states, extra="").replace(
            "  full_precision: true\n",
            "  full_precision: true\n  mode: \"adaptive\"\n  tol: 1.0e-10\n"))
    run(["-e", "2", "-f", "He.yaml", "-o", "out"], tmp_path)
    E, D = leading(full, states)
    o = oracle.System(E, D, gauge="v")
    w = 15 * EV
    tau = 2 * 2.0 * math.pi / w
    pulse = oracle.Pulse.sine(1.0e14 * IO, w, 2, gauge="v")
    x = np.zeros(2 * o.N)
    x[0] = 1.0
    t, h = 0.0, 0.1
    for k in range(1, 11):  # ten segments, as the drop-in runs them
        t1 = tau if k == 10 else 0.0 + k * (tau - 0.0) / 10.0
        x, st, _ = o.prop_adaptive(pulse, t, t1, h, 1e-10, 1e-10, x)
        t, h = st.t, st.h
    xg = read_state(tmp_path / "out" / ("he_ct17_%f.dat" % tau), o.N)
    pops = lambda v: v[0::2] ** 2 + v[1::2] ** 2
    assert np.abs(pops(xg) - pops(x)).max() < 1e-8
    assert np.abs(xg - x).max() < 1e-10


def test_cli_1e_gaussian_length(tmp_path, oracle):
    states = [20, 18, 16, 14]
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "1", "-p", "h",
                    "-s", "20,18,16,14", "-g", "l"], cwd=tmp_path, check=True)
    y = textwrap.dedent("""\
        Global_Settings:
          potential: "h"
          gauge: 'l'
        Basis_Settings:
          l_max: 3
        Propagator_Settings:
          states_in_l: [20,18,16,14]
          dt: 0.05
          full_precision: true
        Field_Parameters:
          shape: "gaussian"
          w: 8.5
          Io: 1.0e13
          cepd: 0.0
          cycles: 3
        """)
    (tmp_path / "H.yaml").write_text(y)
    run(["-e", "1", "-f", "H.yaml", "-o", "out", "-n", "2", "-l", "1"],
        tmp_path)
    E, D = leading(states, states)
    o = oracle.System(E, D, gauge="l")
    w = 8.5 * EV
    tau = 3 * 2.0 * math.pi / w
    steps = int(tau / 0.05)
    pulse = oracle.Pulse.gauss(1.0e13 * IO, w, tau, 3, gauge="l")
    x0 = np.zeros(2 * o.N)
    x0[0] = 1.0
    track = 20 + 1  # -n 2 -l 1: offs[1] + (2 - 1)
    xr, tr, popr, _ = o.prop_fixed(pulse, 0.0, 0.05, steps, x0, track=track)
    x = read_state(tmp_path / "out" / ("h_ct17_%f.dat" % tr), o.N)
    assert np.abs(x - xr).max() < 1e-12
    pop = np.loadtxt(tmp_path / "out" / "h_pop21.dat", comments="#")
    assert np.abs(pop[1:, 1] - popr).max() < 1e-12


def test_cli_golden_he_input(tmp_path, oracle):
    """The reference's own inp/He_test.yaml (tests/golden), unchanged,
    through the drop-in on tdse_gen inputs of its block sizes (3030, 2799,
    2808, 2808; velocity gauge; w = 15 eV, 1e14 W/cm^2, 12 cycles: 1367
    reference-mode steps).  Checked: the coefficient-file cadence; the
    state after the first step against the oracle to the files' six
    significant digits; the 16-digit population file against the library
    path driven with the same frozen fields; the final norm."""
    import shutil

    import mpci
    states = [3030, 2799, 2808, 2808]
    shutil.copy(os.path.join(ROOT, "tests", "golden", "He_test.yaml"),
                tmp_path / "He_test.yaml")
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", ",".join(map(str, states)), "-g", "v"],
                   cwd=tmp_path, check=True)
    run(["-e", "2", "-f", "He_test.yaml", "-o", "out"], tmp_path)
    out = tmp_path / "out"
    w = 15 * EV
    steps = int(12 * 2.0 * math.pi / w / 0.1)
    assert steps == 1367
    pr, t, want = steps // 10, 0.0, {"%f" % 0.0}
    F = np.empty(steps)
    pulse = oracle.Pulse.sine(1.0e14 * IO, w, 12, gauge="v")
    for st in range(steps):
        F[st] = pulse(t + 0.1)
        t += 0.1
        if st % pr == 0:
            want.add("%f" % t)
    want.add("%f" % t)
    got = {f[len("he_ct_"):-4] for f in os.listdir(out)
           if f.startswith("he_ct_")}
    assert got == want
    E, D = leading(states, states)
    o = oracle.System(E, D, gauge="v")
    x0 = np.zeros(2 * o.N)
    x0[0] = 1.0
    x1 = o.step(F[0], x0, 0.1)
    x1 /= math.sqrt(np.dot(x1, x1))
    x = read_state(out / ("he_ct_%f.dat" % 0.1), o.N)
    assert np.all(np.abs(x - x1) <= 1e-5 * np.abs(x1) + 1e-14)
    prop = mpci.Propagator(mpci.System(E, D, gauge="v"))
    prop.set_state(x0)
    popl = prop.prop_fixed(0.0, 0.1, F[None, :], track=0)[0]
    pop = np.loadtxt(out / "he_pop10.dat", comments="#")
    assert pop.shape == (steps + 1, 2)
    assert np.abs(pop[1:, 1] - popl).max() < 1e-15
    xf = read_state(out / ("he_ct_%f.dat" % t), o.N)
    assert abs(np.dot(xf, xf) - 1.0) < 1e-4
    prop.close()


def test_cli_golden_h_input(tmp_path, oracle):
    """The reference's own inp/H_test.yaml (tests/golden), unchanged,
    through the 1e drop-in (`tdse -e 1`) on tdse_gen inputs of its block
    sizes (l <= 5, 300 states each; velocity gauge; w = 8.5 eV,
    1e13 W/cm^2, 12 cycles: 2413 reference-mode steps).  Checked: the
    coefficient-file cadence, the 16-digit ground-population file against
    the oracle over the first 250 steps (within 1e-12), and the final
    norm."""
    import shutil
    states = [300] * 6
    shutil.copy(os.path.join(ROOT, "tests", "golden", "H_test.yaml"),
                tmp_path / "H_test.yaml")
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "1", "-p", "h",
                    "-s", ",".join(map(str, states)), "-g", "v"],
                   cwd=tmp_path, check=True)
    run(["-e", "1", "-f", "H_test.yaml", "-o", "out"], tmp_path)
    out = tmp_path / "out"
    w = 8.5 * EV
    steps = int(12 * 2.0 * math.pi / w / 0.1)
    assert steps == 2413
    pr, t, want = steps // 10, 0.0, {"%f" % 0.0}
    for st in range(steps):
        t += 0.1
        if st % pr == 0:
            want.add("%f" % t)
    want.add("%f" % t)
    got = {f[len("h_ct_"):-4] for f in os.listdir(out)
           if f.startswith("h_ct_")}
    assert got == want
    E, D = leading(states, states)
    o = oracle.System(E, D, gauge="v")
    x0 = np.zeros(2 * o.N)
    x0[0] = 1.0
    pulse = oracle.Pulse.sine(1.0e13 * IO, w, 12, gauge="v")
    _, _, popr, _ = o.prop_fixed(pulse, 0.0, 0.1, 250, x0, track=0)
    pop = np.loadtxt(out / "h_pop10.dat", comments="#")
    assert pop.shape == (steps + 1, 2)
    assert np.abs(pop[1:251, 1] - popr).max() < 1e-12
    xf = read_state(out / ("h_ct_%f.dat" % t), o.N)
    assert abs(np.dot(xf, xf) - 1.0) < 1e-4


def test_cli_restart_full_precision(tmp_path, oracle):
    """Restart from a 17-digit side file (<pot>_ct17_<t>.dat): the start
    time is read after the tag (not the 17 of the tag), the run continues
    from the full-precision state, and its final state matches the oracle
    started from the same state within 1e-12."""
    full = states = [30, 30, 30]
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", "30,30,30", "-g", "v"], cwd=tmp_path, check=True)
    (tmp_path / "He.yaml").write_text(yaml_2e(states, gauge="v"))
    run(["-e", "2", "-f", "He.yaml", "-o", "out"], tmp_path)
    names = sorted(f for f in os.listdir(tmp_path / "out")
                   if f.startswith("he_ct17_"))
    ct = names[3]
    t_start = float(ct[len("he_ct17_"):-4])
    assert 0.0 < t_start < 17.0
    r = run(["-e", "2", "-f", "He.yaml", "-o", "out2", "-s", f"out/{ct}"],
            tmp_path)
    assert f"Start time (a.u.): {t_start:g}" in r.stdout
    E, D = leading(full, states)
    o = oracle.System(E, D, gauge="v")
    w = 15 * EV
    tau = 2 * 2.0 * math.pi / w
    steps = int((tau - t_start) / 0.1)
    x0 = read_state(tmp_path / "out" / ct, o.N)  # 17-digit restart values
    pulse = oracle.Pulse.sine(1.0e14 * IO, w, 2, gauge="v")
    xr, tr, popr, _ = o.prop_fixed(pulse, t_start, 0.1, steps, x0)
    x = read_state(tmp_path / "out2" / ("he_ct17_%f.dat" % tr), o.N)
    assert np.abs(x - xr).max() < 1e-12
    pop = np.loadtxt(tmp_path / "out2" / "he_pop10.dat", comments="#")
    assert pop.shape == (steps + 1, 2)
    assert np.abs(pop[1:, 1] - popr).max() < 1e-12


def _obs_lines(path):
    lines = path.read_text().splitlines()
    assert lines[0] == ("#time (a.u.), ground state pop, yield, excited "
                        "population, norm")
    return np.array([[float(v) for v in ln.split()] for ln in lines[1:]])


@pytest.mark.parametrize("electrons", [2, 1])
def test_cli_observables_side_file(tmp_path, oracle, electrons):
    """Propagator_Settings.observables: true adds <pot>_obs.dat with pes's
    ground population, yield, excited population and norm (16 digits) at
    every coefficient-file time, evaluated on the device; the last line
    matches the oracle's observables (or_observables with the threshold of
    dat/<pot>0.h5:/En[0], or or_observables_1e) of the oracle's final
    state within 1e-12."""
    pot = "he" if electrons == 2 else "h"
    full = states = [30, 36, 28]
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", str(electrons),
                    "-p", pot, "-s", "30,36,28", "-g", "v"], cwd=tmp_path,
                   check=True)
    y = yaml_2e(states).replace("  full_precision: true\n",
                                "  full_precision: true\n"
                                "  observables: true\n")
    y = y.replace('potential: "he"', f'potential: "{pot}"')
    if electrons == 1:
        y = y.replace("l_max: 7", "l_max: 2")
    (tmp_path / "in.yaml").write_text(y)
    run(["-e", str(electrons), "-f", "in.yaml", "-o", "out"], tmp_path)
    out = tmp_path / "out"
    obs = _obs_lines(out / f"{pot}_obs.dat")
    cts = sorted(float(f[len(pot) + 4:-4]) for f in os.listdir(out)
                 if f.startswith(f"{pot}_ct_"))
    # one line per coefficient file (the final time appears twice when the
    # last step also ends a tenth, as for the ct files written there)
    assert sorted(set(np.round(obs[:, 0], 6))) == sorted(set(cts))
    assert obs[0, 1] == 1.0 and obs[0, 4] == 1.0
    E, D = leading(full, states)
    o = oracle.System(E, D, gauge="v")
    w = 15 * EV
    tau = 2 * 2.0 * math.pi / w
    steps = int(tau / 0.1)
    pulse = oracle.Pulse.sine(1.0e14 * IO, w, 2, gauge="v")
    x0 = np.zeros(2 * o.N)
    x0[0] = 1.0
    xr, _, _, _ = o.prop_fixed(pulse, 0.0, 0.1, steps, x0)
    ref = (o.observables(xr, -2.0) if electrons == 2
           else o.observables_1e(xr))
    assert np.abs(obs[-1, 1:] - ref).max() < 1e-12
    assert abs(obs[-1, 4] - 1.0) < 1e-12
