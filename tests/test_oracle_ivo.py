"""IVO (fockivo.F90 / r4divo.F90) on the CPU: the oracle restatement and the library's host code
against the reference's own expected files.

* oracle (oracle/ivo_ref.py + the LAPACK ZHEEV restatement oracle/lapack_zheev.py) on the three
  IVO fixtures: the DFPCMONEW it writes equals reference.DFPCMONEW within 1e-13 per number, the
  reference test's criterion (test_ivo_c2_h2o_dev.py:35-46);
* the sign convention pins: with dlartg in its LAPACK >= 3.10 form (r takes the sign of f) the
  eigenvector signs, and so the file, differ;
* the library's host ZHEEV sequence (dc2_host_heev method 3) returns the oracle's eigenvectors,
  signs included, on random complex and real symmetric matrices;
* the library's DFPCMO reader/writer (dirac_caspt2_amd/ivo.py) reproduces the reference's file
  token for token, and its coefficient rotation with the oracle's eigenvectors gives the
  reference's DFPCMONEW.
"""
import numpy as np
import pytest

from ivo_helpers import IVO_NAMES, assert_same_file, load, paths


@pytest.mark.parametrize("name", IVO_NAMES)
def test_oracle_ivo_equals_reference_file(name, tmp_path):
    from oracle import ivo_ref as I
    p = load(name)
    src, ref = paths(name)
    f = I.read_dfpcmo(src, p.inp.diracver)
    g = I.apply_ivo(p, f, I.ivo_rotations(p, f.nfsym))
    out = tmp_path / "DFPCMONEW"
    I.write_dfpcmo(str(out), g, p.inp.diracver)
    worst = assert_same_file(str(out), ref)
    assert worst < 1e-13


def test_eigenvector_sign_convention_is_pinned():
    from oracle import ivo_ref as I, lapack_zheev as Z
    name = IVO_NAMES[0]
    p = load(name)
    src, ref = paths(name)
    f = I.read_dfpcmo(src, p.inp.diracver)
    r = I.read_dfpcmo(ref, p.inp.diracver)
    rot = I.ivo_rotations(p, f.nfsym)
    F = I.fockivo_matrix(p)
    new = {k: (m, *Z.zheev_vu(F[np.ix_(m, m)], new_lartg=True)) for k, (m, _, _) in rot.items()}
    assert np.abs(I.apply_ivo(p, f, rot).buf - r.buf).max() < 1e-13
    assert np.abs(I.apply_ivo(p, f, new).buf - r.buf).max() > 1e-3


@pytest.mark.parametrize("n", [1, 2, 3, 5, 9, 24, 46])
@pytest.mark.parametrize("real", [True, False])
def test_library_host_zheev_matches_oracle_signs(n, real):
    from dirac_caspt2_amd.api import host_heev
    from oracle import lapack_zheev as Z
    rng = np.random.default_rng(17 * n + real)
    X = rng.normal(size=(n, n)) + (0 if real else 1j) * rng.normal(size=(n, n))
    H = X + X.conj().T
    w0, V0 = Z.zheev_vu(H)
    w, V = host_heev(H, "zheev")
    assert np.abs(w - w0).max() < 1e-12 * max(1.0, np.abs(w0).max())
    assert np.abs(V - V0).max() < 1e-10
    assert np.abs(H @ V - V * w).max() < 1e-12 * max(1.0, np.abs(w).max())


@pytest.mark.parametrize("name", IVO_NAMES)
def test_library_dfpcmo_io_and_rotation(name, tmp_path):
    from dirac_caspt2_amd import ivo
    from oracle import ivo_ref as I
    p = load(name)
    src, ref = paths(name)
    # writer round trip on the reference's own output file
    d = ivo.read_dfpcmo(ref, p.inp.diracver)
    ivo.write_dfpcmo(str(tmp_path / "copy"), d, p.inp.diracver)
    assert assert_same_file(str(tmp_path / "copy"), ref, tol=0.0) == 0.0
    # the library's rotation with the oracle's eigenvectors placed as dc2_fockivo returns them
    f = ivo.read_dfpcmo(src, p.inp.diracver)
    ns = p.inp.nsec
    U = np.zeros((ns, ns), dtype=np.complex128)
    for isym, (mos, w, V) in I.ivo_rotations(p, f.nfsym).items():
        U[np.ix_(mos, mos)] = V
    nmoc = p.inp.ninact + p.inp.nact
    g = ivo.rotate_virtuals(f, U, p.irpamo[nmoc:nmoc + ns], p.nsymrpa, p.inp.occ, p.inp.vcut)
    ivo.write_dfpcmo(str(tmp_path / "DFPCMONEW"), g, p.inp.diracver)
    assert assert_same_file(str(tmp_path / "DFPCMONEW"), ref) < 1e-13
