"""CPU tests of the drop-in's host side: the YAML subset reader
(td_read.cpp:9-43 key set), HDF5 leading-block reads (td_read.cpp:50-120),
coefficient-file format (tdse.cpp:152-158, default ostream precision),
restart parsing (tdse_main.cpp:152-160, td_read.cpp:122-144) and the CLI's
behaviour up to the GPU boundary (settings echo, copies, t0 file, prompt,
help, and a clean failure without a GPU)."""
import ctypes as C
import os
import subprocess
import textwrap

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mp-citdse_amd")
BIN = os.path.join(PKG, "bin")

YAML = textwrap.dedent("""\
    # Settings consistent among all programs
    Global_Settings:
      potential: "he"           # potential name
      Quadrature_type: "gauss-legendre" # unused by tdse
      L_max: 2                  # total L <= L_max
      gauge: 'v'               # 'l' length or 'v' velocity gauge
    Basis_Settings:
      l_max: 7                  # l <= l_max
    Propagator_Settings:
      states_in_l: [20,25,18]
      dt: 0.1 # time step size
    Field_Parameters:
      shape: "sine" # envelope shape "sine" or "gaussian"
      w: 15 # photon energy
      Io: 1.0e14 # peak intensity
      cepd: 0.0  # phase delay
      cycles: 12 # number of field cycles
    """)

ECHO = textwrap.dedent("""\
    Global Settings:
      Core Potential:                       he
      max l/L:                              2
      Gauge type ('l' length/'v' velocity): v
    Propagator Settings:
      states per l/L:          20 25 18 @
      timestep dt:                          0.1
    Field Parameters:
      envelope shape:                       sine
      photon energy:                        15
      Peak intensity:                       1e+14
      phase shift:                          0
      number of cycles:                     12
    """).replace(" @\n", " \n")  # the reference prints a blank after each size


@pytest.fixture(scope="module")
def io():
    lib = C.CDLL(os.path.join(PKG, "lib", "libmpci_hostio.so"))
    lib.mpci_io_last_error.restype = C.c_char_p
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_yaml_subset(io, tmp_path):
    p = tmp_path / "s.yaml"
    p.write_text(YAML)
    buf = C.create_string_buffer(64)
    assert io.mpci_io_yaml_str(str(p).encode(), b"Global_Settings",
                               b"potential", buf, 64) == 0
    assert buf.value == b"he"
    assert io.mpci_io_yaml_str(str(p).encode(), b"Global_Settings", b"gauge",
                               buf, 64) == 0
    assert buf.value == b"v"
    arr = (C.c_int * 8)()
    n = C.c_int()
    assert io.mpci_io_yaml_int_list(str(p).encode(), b"Propagator_Settings",
                                    b"states_in_l", arr, 8, C.byref(n)) == 0
    assert list(arr[:n.value]) == [20, 25, 18]
    v = C.c_double()
    assert io.mpci_io_yaml_double(str(p).encode(), b"Field_Parameters", b"Io",
                                  C.byref(v)) == 0
    assert v.value == 1.0e14
    assert io.mpci_io_yaml_double(str(p).encode(), b"Field_Parameters",
                                  b"missing", C.byref(v)) == -1
    assert b"missing" in io.mpci_io_last_error()


def test_hdf5_leading_block(io, tmp_path):
    from mpci import synth
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", "30,25,28", "-g", "l", "-d", str(tmp_path / "dat"),
                    "-r", "77"], check=True)
    E, D = synth.system_arrays([30, 25, 28], seed=77)
    out = np.empty(12)
    f = str(tmp_path / "dat" / "heCI1.h5").encode()
    assert io.mpci_io_read_1d(f, b"En_CI", 12, _dp(out)) == 0
    assert np.array_equal(out, E[1][:12])
    blk = np.empty((20, 11))
    f = str(tmp_path / "dat" / "heCI_12l.h5").encode()
    assert io.mpci_io_read_2d(f, b"CI_dip", 20, 11, _dp(blk)) == 0
    assert np.array_equal(blk, D[1][:20, :11])
    big = np.empty(31)
    f = str(tmp_path / "dat" / "heCI0.h5").encode()
    assert io.mpci_io_read_1d(f, b"En_CI", 31, _dp(big)) == -1
    thr = np.empty(1)
    assert io.mpci_io_read_1d(str(tmp_path / "dat" / "he0.h5").encode(),
                              b"En", 1, _dp(thr)) == 0 and thr[0] == -2.0


def test_ct_file_format_and_read_back(io, tmp_path):
    ct = np.array([1.0, 0.0, 0.123456789, -2.5e-10, 1e-7, 3.0, -0.5, 12345678.0])
    prefix = str(tmp_path / "he")
    assert io.mpci_io_write_ct(prefix.encode(), C.c_double(12.3), _dp(ct), 4,
                               0) == 0
    txt = (tmp_path / "he_ct_12.300000.dat").read_text()
    lines = ["#Index, Re(c(t)), Im(c(t))"] + [
        f"{k}  {ct[2 * k]:.6g}  {ct[2 * k + 1]:.6g}" for k in range(4)]
    assert txt == "\n".join(lines) + "\n"
    assert "0  1  0" in txt and "1e-07" in txt and "1.23457e+07" in txt
    # readInitCt: comments and leading blanks skipped (td_read.cpp:127-133)
    (tmp_path / "in.dat").write_text("#hdr\n   0  0.5  -0.25\n1 1e-3 2\n")
    back = np.empty(4)
    assert io.mpci_io_read_ct(str(tmp_path / "in.dat").encode(), 2,
                              _dp(back)) == 0
    assert list(back) == [0.5, -0.25, 1e-3, 2.0]
    assert io.mpci_io_read_ct(str(tmp_path / "in.dat").encode(), 3,
                              _dp(np.empty(6))) == -1


@pytest.mark.parametrize("name,t", [
    ("he_ct_12.300000.dat", 12.3),
    ("/x/y/he_ct_0.000000.dat", 0.0),
    ("h_ct_125.600000.dat", 125.6),
    # full-precision side files: the time follows the _ct17_ tag
    ("he_ct17_12.300000.dat", 12.3),
    ("out/he_ct17_0.000000.dat", 0.0),
    ("/a/b1/h_ct17_125.600000.dat", 125.6),
    # plain names keep the reference's rule, the number at the first digit
    # of the file name (tdse_main.cpp:153-157), even a digit of the name
    ("h2_ct_5.000000.dat", 2.0),
])
def test_restart_time_from_name(io, name, t):
    v = C.c_double()
    assert io.mpci_io_time_from_name(name.encode(), C.byref(v)) == 0
    assert v.value == t


@pytest.mark.parametrize("name", ["he_ct_.dat", "out/he_ct17_.dat"])
def test_restart_time_missing(io, name):
    v = C.c_double()
    assert io.mpci_io_time_from_name(name.encode(), C.byref(v)) == -1
    assert b"start time" in io.mpci_io_last_error()


def _cli(args, cwd, stdin=None):
    return subprocess.run([os.path.join(BIN, "tdse")] + args, cwd=cwd,
                          input=stdin, capture_output=True, text=True,
                          timeout=60)


def test_cli_cpu_side(tmp_path):
    import mpci
    if mpci.device_count() > 0:
        pytest.skip("checks the no-GPU path")
    subprocess.run([os.path.join(BIN, "tdse_gen"), "-e", "2", "-p", "he",
                    "-s", "30,25,28", "-g", "v"], cwd=tmp_path, check=True)
    (tmp_path / "He.yaml").write_text(YAML)
    r = _cli(["-e", "2", "-f", "He.yaml", "-o", "out"], tmp_path)
    assert r.returncode == 1
    assert r.stdout == ECHO
    assert "GPU" in r.stderr
    out = tmp_path / "out"
    assert (out / "settings.yaml").read_text() == YAML
    for L in range(3):
        assert (out / f"cfg-{L}.inp").exists()
    t0 = (out / "he_ct_0.000000.dat").read_text().splitlines()
    assert t0[0] == "#Index, Re(c(t)), Im(c(t))" and len(t0) == 1 + 63
    assert t0[1] == "0  1  0" and t0[2] == "1  0  0"
    # settings.yaml exists: the prompt (tdse_main.cpp:78-92)
    r = _cli(["-e", "2", "-f", "He.yaml", "-o", "out"], tmp_path, stdin="n\n")
    assert r.returncode == 255
    assert "contians old data" in r.stdout
    r = _cli(["-h"], tmp_path)
    assert r.returncode == 255 and "-e <1 or 2>" in r.stdout
    r = _cli(["-e", "3"], tmp_path)
    assert r.returncode == 255 and "Invalid number of electrons" in r.stdout


def test_cli_missing_inputs_fail_cleanly(tmp_path):
    (tmp_path / "He.yaml").write_text(YAML)
    r = _cli(["-e", "2", "-f", "He.yaml", "-o", "out"], tmp_path)
    assert r.returncode == 1 and "tdse:" in r.stderr


def test_cli_option_checks(tmp_path):
    """Usage errors end with the reference's messages and its exit status
    (-1); -l is checked itself (the reference tests -n there,
    tdse_main.cpp:59-64), and a missing -e is refused before anything is
    written."""
    (tmp_path / "He.yaml").write_text(YAML)
    cases = [(["-e", "2", "-n", "0"], "Invalid primary quantum number -n"),
             (["-e", "2", "-l", "-1"], "Invalid angular momentum number -l"),
             (["-f", "He.yaml", "-o", "out"], "Invalid number of electrons")]
    for args, msg in cases:
        r = _cli(args, tmp_path)
        assert r.returncode == 255 and msg in r.stdout, (args, r.stdout)
    assert not (tmp_path / "out").exists()
