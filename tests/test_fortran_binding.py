"""A Fortran host (tests/fortran/drive_caspt2.F90 through the iso_c_binding module
dirac_caspt2_amd/fortran/dcaspt2_amd_interface.F90) calling the C ABI.

CPU: the driver links and reaches the library (version string, status and message of
a call on a null context).  GPU: a full CASCI + CASPT2 run driven from Fortran matches
the oracle within 1e-9 Eh per subspace and in total.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dirac_caspt2_amd", "_build", "drive_caspt2")


def _driver():
    if not os.path.exists(DRIVER):
        if not (shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang")):
            pytest.skip("no Fortran compiler and no prebuilt driver")
        from dirac_caspt2_amd.build import build_fortran_interface
        build_fortran_interface()
    return DRIVER


def test_fortran_driver_reaches_library():
    exe = _driver()
    out = subprocess.run([exe, "version"], capture_output=True, text=True, check=True).stdout
    assert "version dcaspt2_amd" in out
    out = subprocess.run([exe, "badarg"], capture_output=True, text=True, check=True).stdout
    assert "status 1 null context" in out


def write_input(path, sp, nroot=2):
    s = sp.spec
    with open(path, "wb") as f:
        np.array([s.ninact, s.nact, s.nsec, s.nelec, s.selectroot, nroot], dtype=np.int32).tofile(f)
        np.array([sp.B.shape[0]], dtype=np.int64).tofile(f)
        np.array([sp.ecore, s.eshift], dtype=np.float64).tofile(f)
        np.asarray(sp.h, dtype=np.complex128).ravel(order="F").tofile(f)
        np.asarray(sp.B, dtype=np.complex128).ravel(order="F").tofile(f)


def parse(out):
    res = {"casci": [], "e2": {}, "sumc2": {}}
    for line in out.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "casci":
            res["casci"].append(float(w[1]))
        elif w[0] == "e2":
            res["e2"][w[1]] = float(w[2])
            res["sumc2"][w[1]] = float(w[3])
        elif w[0] == "total":
            res["total"] = float(w[1])
    return res


@pytest.mark.gpu
def test_fortran_host_caspt2(tmp_path):
    from dirac_caspt2_amd import synth
    from gpu_helpers import oracle_problem, say
    from oracle import caspt2_ref as R
    exe = _driver()
    sp = synth.generate(synth.CONFIG_SMALL)
    p = oracle_problem(sp)
    cas = R.casci(p)
    ref = R.run_caspt2(p, cas)
    inp = tmp_path / "small.bin"
    write_input(inp, sp)
    say("fortran driver: small synthetic")
    run = subprocess.run([exe, "run", str(inp)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    got = parse(run.stdout)
    np.testing.assert_allclose(got["casci"], cas[3][:2], atol=1e-10, rtol=0)
    for X in "ABCDEFGH":
        assert abs(got["e2"][X] - ref.e2[X]) < 1e-9, X
        assert abs(got["sumc2"][X] - ref.sumc2[X]) < 1e-9, X
    assert abs(got["total"] - ref.total) < 1e-9


def test_every_fortran_interface_names_a_declared_entry_point():
    """Each bind(C) name of the Fortran module is a function of include/dcaspt2_amd.h, and the
    module covers the driver path, the CI readers and the request server."""
    import re
    f90 = open(os.path.join(ROOT, "dirac_caspt2_amd", "fortran", "dcaspt2_amd_interface.F90")).read()
    header = open(os.path.join(ROOT, "include", "dcaspt2_amd.h")).read()
    bound = set(re.findall(r'bind\(C,\s*name="([a-z0-9_]+)"\)', f90))
    declared = set(re.findall(r"\b(dc2_[a-z0-9_]+)\s*\(", header))
    assert bound and bound <= declared, sorted(bound - declared)
    for name in ("dc2_casci", "dc2_caspt2", "dc2_caspt2_restart", "dc2_set_integral_records", "dc2_get_dets",
                 "dc2_get_ci_vector", "dc2_server_create_batched", "dc2_server_submit", "dc2_server_wait",
                 "dc2_server_destroy", "dc2_request_footprint"):
        assert name in bound, name


def test_module_exports_readers_and_server(tmp_path):
    """A Fortran program that `use`s the module can call a CI reader and the server entry
    points (compiled and linked against the library; nothing is run)."""
    flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    moddir = os.path.join(ROOT, "dirac_caspt2_amd", "_build")
    if not os.path.exists(flang) or not os.path.exists(os.path.join(moddir, "dcaspt2_amd.mod")):
        pytest.skip("no Fortran compiler or module not built")
    src = tmp_path / "uses_module.F90"
    src.write_text("""program uses_module
    use, intrinsic :: iso_c_binding
    use dcaspt2_amd
    implicit none
    type(c_ptr) :: ctx, srv
    integer(c_int64_t) :: bytes, dets(70), rid, st(7)
    integer(c_int32_t) :: mol, rstat
    complex(c_double_complex) :: c(70), h(4), b(8)
    real(c_double) :: out(26)
    integer(c_int) :: s
    ctx = c_null_ptr
    s = dc2_request_footprint(4_c_int32_t, 8_c_int32_t, 40_c_int32_t, 4_c_int32_t, bytes)
    s = dc2_get_dets(ctx, dets)
    s = dc2_get_ci_vector(ctx, 1_c_int32_t, c)
    s = dc2_server_create_batched(0_c_int32_t, 4_c_int32_t, 1_c_int64_t, 1_c_int32_t, 8_c_int32_t, 0.1d0, srv)
    s = dc2_server_add_molecule(srv, 0_c_int32_t, 2_c_int32_t, 0_c_int32_t, 1_c_int32_t, b, 2_c_int64_t, mol)
    s = dc2_server_submit(srv, mol, h, 0d0, 1_c_int32_t, 0d0, c_null_ptr, 0_c_int32_t, rid)
    s = dc2_server_wait(srv, 1d0, rid, rstat, out)
    s = dc2_server_stats(srv, st)
    s = dc2_server_destroy(srv)
end program
""")
    exe = tmp_path / "uses_module"
    lib = os.path.join(ROOT, "dirac_caspt2_amd")
    run = subprocess.run([flang, "-O0", "-I", moddir, str(src), os.path.join(moddir, "dcaspt2_amd_interface.o"),
                          "-L", lib, "-ldcaspt2_amd", "-o", str(exe)], capture_output=True, text=True)
    assert run.returncode == 0, run.stdout + run.stderr
