"""Code generation of the HIP kernels (CPU only: hipcc cross-compiles for gfx950).

The device records' pointers carry the global address space in the kernels' device pass, so every
load and store of the kernels is a global instruction: a flat one would also count on the LDS
counter, and the first use of any flat-loaded value would wait for every flat access in flight."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not available")
def test_kernels_use_no_flat_memory_instructions(tmp_path):
    out = tmp_path / "md_kernels.s"
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "mesh-deform_amd", "csrc"),
           "--cuda-device-only", "-S", os.path.join(ROOT, "mesh-deform_amd", "csrc", "md_kernels.hip"),
           "-o", str(out)]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
    asm = out.read_text()
    kernels = re.findall(r"^(_Z\w+):", asm, re.M)
    assert any("k_fwd" in k for k in kernels) and any("k_sub_bwd" in k for k in kernels)
    flat = re.findall(r"^\s+(flat_(?:load|store|atomic)\w*)", asm, re.M)
    assert not flat, "flat memory instructions in the kernels: %s" % sorted(set(flat))
    assert len(re.findall(r"^\s+global_load", asm, re.M)) > 100
