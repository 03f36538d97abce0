"""CPU check of the gfx950 assembly of every kernel source: device code
writes memory through vector instructions only (no scalar-unit stores,
atomics or scalar-cache write-backs).  It compiles the sources on the CPU
and runs nothing on a device, so it never needs to travel with the GPU
tests."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mp-citdse_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
SCALAR_WRITES = ("s_store", "s_atomic", "s_buffer_store", "s_buffer_atomic",
                 "s_dcache_wb", "s_dcache_discard", "s_scratch_store")


@pytest.mark.parametrize("src", ["kernels.hip", "ci.hip"])
def test_device_code_has_no_scalar_writes(tmp_path, src):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path / (src + ".s")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17",
                    "--cuda-device-only", "-S", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "-mllvm",
                    "-amdgpu-atomic-optimizer-strategy=None",
                    os.path.join(CSRC, src), "-o", str(out)],
                   check=True, capture_output=True, timeout=300)
    asm = out.read_text()
    for op in SCALAR_WRITES:
        assert not re.search(r"^\s+" + op, asm, flags=re.M), op
