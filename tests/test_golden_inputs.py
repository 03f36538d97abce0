"""The reference's own example inputs (tests/golden, copied from inp/)
through the drop-in's input side, on the CPU: the settings echo of
tdrd::readConfig (src/tdse/td_read.cpp:3-48) for a two-electron (-e 2,
L_max from Global_Settings) and a one-electron (-e 1, l_max from
Basis_Settings; tdse_main.cpp:104-121) run, the settings and cfg-L.inp
copies (tdse_main.cpp:94-100, 122-126), and a clean stop at the missing
HDF5 inputs.  The expected echoes are written out by hand from the YAML
files and the reference's print statements."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PKG = os.path.join(ROOT, "mp-citdse_amd")
TDSE = os.path.join(PKG, "bin", "tdse")

ECHO_HE = (
    "Global Settings:\n"
    "  Core Potential:                       he\n"
    "  max l/L:                              3\n"
    "  Gauge type ('l' length/'v' velocity): v\n"
    "Propagator Settings:\n"
    "  states per l/L:          3030 2799 2808 2808 \n"
    "  timestep dt:                          0.1\n"
    "Field Parameters:\n"
    "  envelope shape:                       sine\n"
    "  photon energy:                        15\n"
    "  Peak intensity:                       1e+14\n"
    "  phase shift:                          0\n"
    "  number of cycles:                     12\n")

ECHO_H = (
    "Global Settings:\n"
    "  Core Potential:                       h\n"
    "  max l/L:                              5\n"
    "  Gauge type ('l' length/'v' velocity): v\n"
    "Propagator Settings:\n"
    "  states per l/L:          300 300 300 300 300 300 \n"
    "  timestep dt:                          0.1\n"
    "Field Parameters:\n"
    "  envelope shape:                       sine\n"
    "  photon energy:                        8.5\n"
    "  Peak intensity:                       1e+13\n"
    "  phase shift:                          0\n"
    "  number of cycles:                     12\n")


def _read(p):
    with open(p, "rb") as f:
        return f.read()


@pytest.mark.parametrize("yaml,e,echo,first_h5", [
    ("He_test.yaml", "2", ECHO_HE, "dat/heCI0.h5"),
    ("H_test.yaml", "1", ECHO_H, "dat/h0.h5"),
])
def test_reference_example_inputs(tmp_path, yaml, e, echo, first_h5):
    shutil.copy(os.path.join(GOLDEN, yaml), tmp_path / yaml)
    if e == "2":
        (tmp_path / "dat").mkdir()
        for L in range(4):
            shutil.copy(os.path.join(GOLDEN, f"cfg-{L}.inp"),
                        tmp_path / "dat" / f"cfg-{L}.inp")
    r = subprocess.run([TDSE, "-e", e, "-f", yaml, "-o", "out"],
                       cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.stdout == echo
    assert r.returncode == 1
    assert f"cannot open HDF5 file {first_h5}" in r.stderr
    out = tmp_path / "out"
    assert _read(out / "settings.yaml") == _read(os.path.join(GOLDEN, yaml))
    if e == "2":
        for L in range(4):
            assert _read(out / f"cfg-{L}.inp") == _read(
                os.path.join(GOLDEN, f"cfg-{L}.inp"))


def test_reference_example_values():
    """The keys tdse and the structure programs share, read through the
    drop-in's YAML subset as typed values."""
    io = C.CDLL(os.path.join(PKG, "lib", "libmpci_hostio.so"))
    io.mpci_io_yaml_double.restype = C.c_int
    he = os.path.join(GOLDEN, "He_test.yaml").encode()
    h = os.path.join(GOLDEN, "H_test.yaml").encode()
    buf = C.create_string_buffer(64)
    assert io.mpci_io_yaml_str(he, b"Global_Settings", b"potential", buf,
                               64) == 0 and buf.value == b"he"
    assert io.mpci_io_yaml_str(h, b"Global_Settings", b"gauge", buf,
                               64) == 0 and buf.value == b"v"
    sizes = (C.c_int * 8)()
    n = C.c_int()
    assert io.mpci_io_yaml_int_list(he, b"Propagator_Settings",
                                    b"states_in_l", sizes, 8,
                                    C.byref(n)) == 0
    assert list(sizes[:n.value]) == [3030, 2799, 2808, 2808]
    v = C.c_double()
    assert io.mpci_io_yaml_double(he, b"Field_Parameters", b"Io",
                                  C.byref(v)) == 0 and v.value == 1.0e14
    assert io.mpci_io_yaml_double(h, b"Field_Parameters", b"w",
                                  C.byref(v)) == 0 and v.value == 8.5
    assert io.mpci_io_yaml_double(h, b"Basis_Settings", b"l_max",
                                  C.byref(v)) == 0 and v.value == 5.0
