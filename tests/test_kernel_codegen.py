"""Code-generation guards for the HIP kernels (CPU only: hipcc cross-compiles
gfx950 device assembly here).

Every kernel must be a single call-free function with no private (scratch)
segment.  hipcc (ROCm 7.2) compiled an out-of-line device-function call
inside a divergent branch so that lanes outside the branch lost threadIdx.x
across the call; the block reduction in the adjoint stage kernel then wrote
to the wrong LDS slots.  Device helpers are __forceinline__ for that reason,
and this test keeps any new call or spill from slipping in."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "mp-citdse_amd", "csrc", "kernels.hip")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def device_asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "kernels.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17",
                    "--cuda-device-only", "-S", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", SRC, "-o", str(out)],
                   check=True, capture_output=True, timeout=300)
    return out.read_text()


def kernel_meta(asm):
    """(name, private_segment_fixed_size) for every kernel in the metadata."""
    names = re.findall(r"^\s+\.name:\s+(\S+)", asm, flags=re.M)
    priv = re.findall(r"^\s+\.private_segment_fixed_size:\s+(\d+)", asm,
                      flags=re.M)
    return [n for n in names if n.startswith("_Z")], [int(p) for p in priv]


def test_no_device_calls(device_asm):
    assert "s_swappc_b64" not in device_asm
    assert "s_setpc_b64" not in device_asm


def test_no_scratch(device_asm):
    names, priv = kernel_meta(device_asm)
    assert len(priv) >= 10
    assert all(p == 0 for p in priv), list(zip(names, priv))


def test_expected_kernels_present(device_asm):
    for k in ("k_gemv_tile", "k_stage", "k_stage1", "k_finalize", "k_apply",
              "k_adj_stage", "k_adj_init", "k_dot_partial",
              "k_gemm_coupling", "k_gemm_pair", "k_copy_strided"):
        assert k in device_asm, k


def test_stage_need_masks_match_tableau(oracle):
    """k_stage1 loads only the k_j whose tableau weight is nonzero, chosen
    by a per-stage bit mask written into the kernel source: the masks must
    equal the zero pattern of the oracle's Fehlberg 7(8) tableau (which the
    oracle tests pin by the exact order conditions)."""
    src = open(SRC).read()
    m = re.search(r"cNeed\[kStages\]\s*=\s*\{([^}]*)\}", src)
    assert m, "cNeed table not found"
    masks = [int(v, 16) for v in re.findall(r"0x[0-9A-Fa-f]+", m.group(1))]
    a, b, db, _ = oracle.tableau()
    want = []
    for s in range(12):  # stage s forms stage s + 1's input from k_j, j < s
        want.append(sum(1 << j for j in range(s) if a[s + 1][j] != 0.0))
    want.append(sum(1 << j for j in range(12) if b[j] != 0.0 or db[j] != 0.0))
    assert masks == want, [hex(v) for v in want]
