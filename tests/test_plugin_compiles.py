"""The reference-side plug-in (sqcalc_amd/fortran/sqc_hip_sf.f90) compiles
against the reference's own modules, where they lie: it extends
structure_factor_t (src/sqc_structure_factor.f90:71-192) with the deferred
method_setup / accumulate_frame and the prepare_output hook, reads the
reference's weight scheme and IT92 tables, and calls the C ABI through
sqc_hip.f90.  Compile only (amdflang -c): the reference's driver links
FINUFFT, which is not buildable here.  Runs where the reference and amdflang
are present."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
MODULES = ["sqc_kinds", "sqc_cell", "sqc_dump", "sqc_element_data", "sqc_elements",
           "sqc_weights", "sqc_finufft", "sqc_structure_factor"]


def _compile(src, tmp, cpp=False):
    obj = os.path.join(tmp, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [FC] + (["-cpp"] if cpp else []) + ["-c", src, "-module-dir", tmp, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, "%s\n%s" % (" ".join(cmd), r.stderr)


@pytest.mark.skipif(not os.path.isdir(REF) or not os.path.exists(FC),
                    reason="needs the reference sources and amdflang")
def test_plugin_compiles_against_reference_modules(tmp_path):
    tmp = str(tmp_path)
    for m in MODULES:
        _compile(os.path.join(REF, m + ".f90"), tmp, cpp=True)
    _compile(os.path.join(ROOT, "sqcalc_amd", "fortran", "sqc_hip.f90"), tmp)
    _compile(os.path.join(ROOT, "sqcalc_amd", "fortran", "sqc_hip_sf.f90"), tmp)
    assert os.path.exists(os.path.join(tmp, "sqc_hip_sf.mod"))
