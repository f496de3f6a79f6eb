"""The radix scatter's destination arithmetic (gs3d_internal.h scatter_delta /
scatter_index, called by every k_radix_pass) at key counts of 2^31 and more.

The helpers are __host__ __device__; this compiles them into a host program
with hipcc and checks every index against the exact 64-bit destination,
around the int32 and uint32 limits. No GPU is used: a >= 2^31-key frame is
never run on the device.
"""
import subprocess
from pathlib import Path

REPO = Path(__file__).resolve().parent.parent
HIPCC = "/opt/rocm/bin/hipcc"


def test_scatter_index_exact_above_2_31(tmp_path):
    exe = tmp_path / "scatter_index_check"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "--no-gpu-bundle-output",
                    "--cuda-host-only", f"-I{REPO / 'include'}", f"-I{REPO / 'webgpu-3dgs_amd' / 'csrc'}",
                    str(REPO / "tests" / "cpp" / "scatter_index_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
    assert int(r.stdout.split()[1]) > 10_000
