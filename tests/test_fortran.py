"""The Fortran side of the boundary: ISO_C_BINDING module and driver build
with the reference's toolchain (amdflang) and report library errors."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
DRIVE = ROOT / "damian.par_amd" / "lib" / "dmp_drive"


def test_driver_built():
    if not Path("/opt/rocm/bin/amdflang").exists():
        pytest.skip("amdflang not available")
    assert DRIVE.exists(), "run python damian.par_amd/build.py"
    assert (ROOT / "damian.par_amd" / "lib" / "damian_step.mod").exists()


def test_driver_reports_missing_device(tmp_path):
    if not DRIVE.exists():
        pytest.skip("driver not built")
    p = subprocess.run([str(DRIVE), str(tmp_path), "0"], capture_output=True, text=True, timeout=120)
    if p.returncode == 0:
        pytest.skip("a HIP device is present")
    assert p.returncode == 3
    assert "error -3" in p.stdout
