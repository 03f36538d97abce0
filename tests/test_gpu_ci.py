"""GPU parity of the producers of the tdse dipole inputs against the CPU
restatement (oracle/ci_oracle.py): the CI basis (the w2e stage,
src/w2e/w2e.cpp:17-125) and the one-electron dipoles (the d1e stage,
src/d1e/dmx1e.cpp:24-139), through the C ABI and through the `w2e` and `d1e`
drop-ins on reference-layout HDF5 files; then `tdse` on the files they
wrote.

Bars: eigenvalues within 1e-12 of the oracle (relative to the spectrum's
scale); eigenvectors, with the sign convention applied on both sides,
within 1e-9 (bounded by eps / smallest eigenvalue gap of these random
matrices); CI dipoles within 1e-10; the propagated state within 1e-12."""
import math
import os
import subprocess
import textwrap

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mp-citdse_amd", "bin")


# The CI library links ROCm's rocSOLVER and rocBLAS.  A process that has
# imported torch already holds torch's own copies of those libraries (same
# sonames, another ROCm release), so the library calls run in a child
# process that imports only the mpci package.
CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
from mpci import ci
inp = np.load(sys.argv[2])
out = {}
for key in sorted({k.split(":")[0] for k in inp.files}):
    if key.startswith("basis"):
        w, V = ci.ci_basis(inp[key + ":ens"], inp[key + ":v12"])
        out[key + ":w"], out[key + ":V"] = w, V
    elif key.startswith("d1e"):
        g = "v" if key.endswith("v") else "l"
        wp = inp[key + ":wp"] if g == "v" else None
        out[key + ":out"] = ci.d1e_dipoles(int(inp[key + ":l"]), g,
                                           inp[key + ":qx"], inp[key + ":qw"],
                                           inp[key + ":w0"], inp[key + ":w1"],
                                           wp)
    else:
        out[key + ":out"] = ci.ci_dipoles(inp[key + ":D"], inp[key + ":V0"],
                                          inp[key + ":V1"])
np.savez(sys.argv[3], **out)
"""


def gpu_ci(tmp_path, inputs):
    """Runs ci_basis / ci_dipoles on the GPU in a child process."""
    import sys
    fi, fo = tmp_path / "ci_in.npz", tmp_path / "ci_out.npz"
    np.savez(fi, **inputs)
    r = subprocess.run([sys.executable, "-c", CHILD,
                        os.path.join(ROOT, "mp-citdse_amd"), str(fi), str(fo)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return dict(np.load(fo))


def pack_upper(H):
    return H[np.triu_indices(H.shape[0])]


def gap(w):
    return np.diff(w).min() if len(w) > 1 else 1.0


def test_ci_basis_vs_oracle(tmp_path):
    from oracle import ci_oracle as co
    inputs, cases = {}, {}
    for n in (1, 2, 7, 64, 301):
        rng = np.random.default_rng(n)
        ens = np.sort(rng.uniform(-2.9, 5, n))
        A = rng.standard_normal((n, n)) * (0.2 / math.sqrt(n))
        v12 = pack_upper(A + A.T)
        inputs[f"basis{n}:ens"], inputs[f"basis{n}:v12"] = ens, v12
        cases[n] = co.ci_basis(ens, v12)
    out = gpu_ci(tmp_path, inputs)
    for n, (wo, Vo) in cases.items():
        w, V = out[f"basis{n}:w"], out[f"basis{n}:V"]
        scale = max(1.0, np.abs(wo).max())
        assert np.abs(w - wo).max() < 1e-12 * scale, n
        assert np.abs(V - Vo).max() < max(1e-9, 1e-14 / gap(wo)), n
        assert np.abs(V @ V.T - np.eye(n)).max() < 1e-12, n


def sym(rng, n, scale):
    A = rng.standard_normal((n, n)) * scale
    return A + A.T


def test_ci_dipoles_vs_oracle(tmp_path):
    from oracle import ci_oracle as co
    rng = np.random.default_rng(7)
    n0, n1 = 57, 43
    V0 = co.ci_basis(np.sort(rng.uniform(-1, 1, n0)),
                     pack_upper(sym(rng, n0, 0.05)))[1]
    V1 = co.ci_basis(np.sort(rng.uniform(-1, 1, n1)),
                     pack_upper(sym(rng, n1, 0.05)))[1]
    D = rng.standard_normal((n0, n1)) / math.sqrt(n1)
    out = gpu_ci(tmp_path, {"dip:D": D, "dip:V0": V0, "dip:V1": V1})
    got = out["dip:out"]
    ref = co.ci_dipoles(D, V0, V1)
    assert np.abs(got - ref).max() < 1e-12 * max(1, np.abs(ref).max())


def _io():
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "mp-citdse_amd", "lib",
                              "libmpci_hostio.so"))
    lib.mpci_io_last_error.restype = C.c_char_p
    return lib


def _read(io, path, name, shape):
    import ctypes as C
    out = np.empty(shape)
    dp = out.ctypes.data_as(C.POINTER(C.c_double))
    if len(shape) == 1:
        rc = io.mpci_io_read_1d(str(path).encode(), name.encode(),
                                C.c_long(shape[0]), dp)
    else:
        rc = io.mpci_io_read_2d(str(path).encode(), name.encode(),
                                C.c_long(shape[0]), C.c_long(shape[1]), dp)
    assert rc == 0, io.mpci_io_last_error()
    return out


def test_w2e_cli_then_tdse(tmp_path, oracle):
    """tdse_gen -w (product-basis inputs) -> w2e (GPU) -> tdse -e 2 (GPU),
    every file in the reference layout, checked against the oracle."""
    from mpci import synth
    from oracle import ci_oracle as co
    s = [40, 33, 25]
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", ",".join(map(str, s)), "-g", "l", "-w"],
                   cwd=tmp_path, check=True)
    (tmp_path / "W.yaml").write_text(textwrap.dedent("""\
        Global_Settings:
          potential: "he"
          L_max: 2
          gauge: 'l'
        """))
    r = subprocess.run([os.path.join(BIN, "w2e"), "-f", "W.yaml"],
                       cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ens, v12, dip = synth.w2e_arrays(s)
    io = _io()
    E, V = [], []
    for L, n in enumerate(s):
        wo, Vo = co.ci_basis(ens[L], v12[L])
        w = _read(io, tmp_path / "dat" / f"heCI{L}.h5", "En_CI", (n,))
        Vg = _read(io, tmp_path / "dat" / f"heCI{L}.h5", "CI_vecs", (n, n))
        assert np.abs(w - wo).max() < 1e-12 * max(1, np.abs(wo).max())
        assert np.abs(Vg - Vo).max() < max(1e-9, 1e-14 / gap(wo))
        E.append(wo)
        V.append(Vo)
    assert f"Ground state energy (a.u.): {E[0][0]:g}" in r.stdout
    D = []
    for L in range(len(s) - 1):
        ref = co.ci_dipoles(dip[L], V[L], V[L + 1])
        got = _read(io, tmp_path / "dat" / f"heCI_{L}{L + 1}l.h5", "CI_dip",
                    (s[L], s[L + 1]))
        assert np.abs(got - ref).max() < 1e-10 * max(1, np.abs(ref).max())
        D.append(ref)
    # the CI files feed tdse -e 2 unchanged (reference mode, full precision)
    (tmp_path / "He.yaml").write_text(textwrap.dedent(f"""\
        Global_Settings:
          potential: "he"
          L_max: 2
          gauge: 'l'
        Basis_Settings:
          l_max: 3
        Propagator_Settings:
          states_in_l: [{",".join(map(str, s))}]
          dt: 0.1
          full_precision: true
        Field_Parameters:
          shape: "sine"
          w: 15
          Io: 1.0e14
          cepd: 0.0
          cycles: 1
        """))
    r = subprocess.run([os.path.join(BIN, "tdse"), "-e", "2", "-f", "He.yaml",
                        "-o", "out"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # the oracle on the oracle's own CI system
    o = oracle.System(E, D, gauge="l")
    w = 15 / 27.211396181
    steps = int(2 * math.pi / w / 0.1)
    pulse = oracle.Pulse.sine(1.0e14 * (1.0 / 3.509338) * 1e-16, w, 1,
                              gauge="l")
    x0 = np.zeros(2 * o.N)
    x0[0] = 1.0
    xr, tr, _, _ = o.prop_fixed(pulse, 0.0, 0.1, steps, x0)
    got = np.loadtxt(tmp_path / "out" / ("he_ct17_%f.dat" % tr), comments="#")
    x = np.empty(2 * o.N)
    x[0::2], x[1::2] = got[:, 1], got[:, 2]
    # eigenvector sign conventions agree, so the states agree coefficientwise
    assert np.abs(x - xr).max() < 1e-9


def test_d1e_vs_oracle_and_hydrogen(tmp_path):
    """Random samples in both gauges against the oracle (1e-12), and the
    hydrogen 1s-2p radial dipole in both gauges against its closed form."""
    from oracle import ci_oracle as co
    rng = np.random.default_rng(11)
    nen, q = 37, 129
    qx = np.sort(rng.uniform(0.05, 60, q))
    qw = rng.uniform(0, 0.2, q)
    w0, w1, wp = (rng.standard_normal((nen, q)) * 0.1 for _ in range(3))
    inputs = {}
    for l, g in ((0, "l"), (2, "l"), (0, "v"), (3, "v")):
        k = f"d1e{l}{g}"
        inputs.update({k + ":l": np.array(l), k + ":qx": qx, k + ":qw": qw,
                       k + ":w0": w0, k + ":w1": w1, k + ":wp": wp})
    x, w = np.polynomial.legendre.leggauss(200)
    r, wr = 30.0 * (x + 1), 30.0 * w
    p10 = (2 * r * np.exp(-r))[None, :]
    p21 = (r * r * np.exp(-r / 2) / (2 * math.sqrt(6)))[None, :]
    dp21 = ((2 * r - r * r / 2) * np.exp(-r / 2) / (2 * math.sqrt(6)))[None, :]
    for g in "lv":
        k = f"d1eH{g}"
        inputs.update({k + ":l": np.array(0), k + ":qx": r, k + ":qw": wr,
                       k + ":w0": p10, k + ":w1": p21, k + ":wp": dp21})
    out = gpu_ci(tmp_path, inputs)
    for l, g in ((0, "l"), (2, "l"), (0, "v"), (3, "v")):
        ref = co.d1e_dipoles(l, g, qx, qw, w0, w1, wp)
        got = out[f"d1e{l}{g}:out"]
        assert np.abs(got - ref).max() < 1e-12 * max(1, np.abs(ref).max())
    radial = 128 * math.sqrt(6) / 243
    assert abs(out["d1eHl:out"][0, 0] - radial / math.sqrt(3)) < 1e-12
    assert abs(out["d1eHv:out"][0, 0] - 0.375 * radial / math.sqrt(3)) < 1e-12


@pytest.mark.parametrize("gauge", ["v", "l"])
def test_d1e_cli_then_tdse(tmp_path, oracle, gauge):
    """tdse_gen -e 1 -w (radial samples, quadrature, energies) -> d1e (GPU)
    -> tdse -e 1 (GPU) against the oracle."""
    from mpci import synth
    from oracle import ci_oracle as co
    nen, q, nl = 24, 90, 3
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "1", "-p", "h",
                    "-s", ",".join([str(nen)] * nl), "-g", gauge, "-w",
                    "-q", str(q)], cwd=tmp_path, check=True)
    (tmp_path / "D.yaml").write_text(textwrap.dedent(f"""\
        Global_Settings:
          potential: "h"
          Outer_quadrature_size: {q}
          gauge: '{gauge}'
        Basis_Settings:
          l_max: {nl - 1}
        """))
    r = subprocess.run([os.path.join(BIN, "d1e"), "-f", "D.yaml"],
                       cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"No. of outer quadrature points:              {q}" in r.stdout
    qx, qw, po, pp, en = synth.w1e_arrays(nen, q, nl)
    io = _io()
    D = []
    for l in range(nl - 1):
        ref = co.d1e_dipoles(l, gauge, qx, qw, po[l], po[l + 1], pp[l + 1])
        got = _read(io, tmp_path / "dat" / f"h{l}{l + 1}{gauge}.h5", "d_if",
                    (nen, nen))
        assert np.abs(got - ref).max() < 1e-12 * max(1, np.abs(ref).max())
        D.append(ref)
    (tmp_path / "H.yaml").write_text(textwrap.dedent(f"""\
        Global_Settings:
          potential: "h"
          gauge: '{gauge}'
        Basis_Settings:
          l_max: {nl - 1}
        Propagator_Settings:
          states_in_l: [{",".join([str(nen)] * nl)}]
          dt: 0.05
          full_precision: true
        Field_Parameters:
          shape: "sine"
          w: 8.5
          Io: 1.0e13
          cepd: 0.0
          cycles: 1
        """))
    r = subprocess.run([os.path.join(BIN, "tdse"), "-e", "1", "-f", "H.yaml",
                        "-o", "out", "-n", "1", "-l", "0"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    o = oracle.System(en, D, gauge=gauge)
    w = 8.5 / 27.211396181
    steps = int(2 * math.pi / w / 0.05)
    pulse = oracle.Pulse.sine(1.0e13 * (1.0 / 3.509338) * 1e-16, w, 1,
                              gauge=gauge)
    x0 = np.zeros(2 * o.N)
    x0[0] = 1.0
    xr, tr, _, _ = o.prop_fixed(pulse, 0.0, 0.05, steps, x0)
    got = np.loadtxt(tmp_path / "out" / ("h_ct17_%f.dat" % tr), comments="#")
    x = np.empty(2 * o.N)
    x[0::2], x[1::2] = got[:, 1], got[:, 2]
    assert np.abs(x - xr).max() < 1e-10
