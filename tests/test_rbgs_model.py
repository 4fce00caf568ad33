"""The fused red-black sweep's phase functions (vg_rbgs.h) run on the host,
workgroup by workgroup, against a plain two-colour red-black Gauss-Seidel
sweep: bit-identical results, no out-of-bounds global or LDS access
(AddressSanitizer), no ring-slot aliasing between unbarriered phases."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "rbgs_model.hip")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_fused_sweep_host_model(tmp_path):
    exe = str(tmp_path / "rbgs_model")
    cmd = [HIPCC, "-O1", "-g", "--offload-arch=gfx950", "-Xarch_host", "-fsanitize=address",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "ALL MATCH" in r.stdout
