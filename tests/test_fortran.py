"""The Fortran side of the boundary: ISO_C_BINDING module, driver and
drop-in harness build with the reference's toolchain (amdflang); every
exported entry point has a Fortran interface; INTEGRATION.md shows the
drop-in bodies that are compiled; library errors are reported."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "damian.par_amd" / "lib"
DRIVE = LIBDIR / "dmp_drive"
DROPIN = LIBDIR / "dmp_dropin"
FORTRAN = ROOT / "damian.par_amd" / "fortran"


def test_driver_built():
    if not Path("/opt/rocm/bin/amdflang").exists():
        pytest.skip("amdflang not available")
    assert DRIVE.exists() and DROPIN.exists(), "run python damian.par_amd/build.py"
    assert (LIBDIR / "damian_step.mod").exists()


def test_every_export_has_a_fortran_interface():
    header = (ROOT / "include" / "damian_step.h").read_text()
    exports = set(re.findall(r"^int (dmp_\w+)\(", header, re.M))
    module = (FORTRAN / "damian_step.f90").read_text()
    bound = set(re.findall(r'bind\(C, name="(dmp_\w+)"\)', module))
    assert len(exports) >= 30
    assert exports == bound, (exports - bound, bound - exports)


def test_integration_md_shows_the_compiled_dropin():
    """The replacement bodies in INTEGRATION.md are the text of
    explicit_dmp.inc, which the drop-in harness compiles and the GPU tests run."""
    md = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```fortran\n(.*?)```", md, re.S)
    inc = (FORTRAN / "explicit_dmp.inc").read_text()
    assert inc in blocks


def test_dropin_compiles_against_reference_declarations(tmp_path):
    """explicit_dmp.inc in a module that uses the harness's stand-ins for
    the reference modules compiles with implicit none (no undeclared name,
    no interface mismatch)."""
    fc = Path("/opt/rocm/bin/amdflang")
    if not fc.exists() or not (LIBDIR / "damian_step.mod").exists():
        pytest.skip("amdflang or the built module missing")
    src = tmp_path / "check.f90"
    harness = (FORTRAN / "dropin_harness.f90").read_text()
    src.write_text(harness.split("program dmp_dropin")[0])
    p = subprocess.run([str(fc), "-c", "-I", str(LIBDIR), "-I", str(FORTRAN), "-J", str(tmp_path),
                        str(src), "-o", str(tmp_path / "check.o")], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_driver_reports_missing_device(tmp_path):
    if not DRIVE.exists():
        pytest.skip("driver not built")
    p = subprocess.run([str(DRIVE), str(tmp_path), "0"], capture_output=True, text=True, timeout=120)
    if p.returncode == 0:
        pytest.skip("a HIP device is present")
    assert p.returncode == 3
    assert "error -3" in p.stdout
