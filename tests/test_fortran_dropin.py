"""The same-name drop-in (dirac_caspt2_amd/fortran/dcaspt2_amd_dropin.F90): subroutines
r4dcasci and r4dcaspt2_tra without arguments, compiled against the reference's own
module_global_variables (and the modules it uses) from the reference tree, in a scratch
directory; nothing of the reference is written into the repository.  The objects define the
external names the reference's driver calls (dcaspt2_run_subprograms.F90:57,62), so linking
them in place of r4dcasci.F90 and r4dcaspt2_tra.F90 replaces those two files."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"


def _flang():
    f = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    return f if os.path.exists(f) else None


def test_dropin_compiles_against_reference_modules(tmp_path):
    flang = _flang()
    if flang is None or not os.path.isdir(REF):
        pytest.skip("needs amdflang and the reference sources")
    out = str(tmp_path)

    def fc(src, obj, pre=True):
        cmd = [flang] + (["-cpp"] if pre else []) + ["-O1", "-I", out, "-module-dir", out, "-c", src, "-o",
                                                     os.path.join(out, obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout + r.stderr

    for m in ("module_error", "module_dict", "module_global_variables"):
        fc(os.path.join(REF, m + ".F90"), m + ".o")
    fc(os.path.join(ROOT, "dirac_caspt2_amd", "fortran", "dcaspt2_amd_interface.F90"), "iface.o")
    fc(os.path.join(ROOT, "dirac_caspt2_amd", "fortran", "dcaspt2_amd_dropin.F90"), "dropin.o")
    syms = subprocess.run(["nm", os.path.join(out, "dropin.o")], capture_output=True, text=True).stdout
    defined = set(re.findall(r"\sT\s(\S+)", syms))
    assert {"r4dcasci_", "r4dcaspt2_tra_"} <= defined, defined
    # the reference's routines of those names take no arguments, as the drop-in's
    for name, src in (("r4dcasci", "r4dcasci.F90"), ("r4dcaspt2_tra", "r4dcaspt2_tra.F90")):
        text = open(os.path.join(REF, src)).read()
        m = re.search(r"^\s*subroutine\s+%s\b([^\n!]*)" % name, text, re.I | re.M)
        assert m and "(" not in m.group(1), (name, m and m.group(1))
