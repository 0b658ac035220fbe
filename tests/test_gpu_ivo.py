"""IVO on the GPU (dc2_fockivo + dirac_caspt2_amd/ivo.py) on the three IVO fixtures: the IVO Fock
matrix built on the device from the F1/F2 blocks equals the oracle's within 1e-13, the per-irrep
eigenvectors equal the oracle's ZHEEV restatement (signs included) within 1e-12, and the
DFPCMONEW written from them equals the reference's reference.DFPCMONEW within 1e-13 per number
(the reference test's criterion, test_ivo_c2_h2o_dev.py:35-46)."""
import numpy as np
import pytest

from ivo_helpers import IVO_NAMES, assert_same_file, load, paths

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", IVO_NAMES)
def test_fockivo_and_dfpcmonew(name, tmp_path):
    from dirac_caspt2_amd import ivo
    from dirac_caspt2_amd.api import Caspt2Context
    from gpu_helpers import say
    from oracle import ivo_ref as I
    say("ivo", name)
    p = load(name)
    inp = p.inp
    ni, na, ns = inp.ninact, inp.nact, inp.nsec
    nmoc, nmom = ni + na, ni + na + inp.nsec
    F1 = np.zeros((ns, ns, nmoc, nmoc), dtype=np.complex128)
    F2 = np.zeros((ns, nmoc, nmoc, ns), dtype=np.complex128)
    for t, k in enumerate(p.kset):
        F1[:, :, k, k] = p.f1h[:, :, t]
        F2[:, k, k, :] = p.f2h[:, t, :]
    src, ref = paths(name)
    with Caspt2Context(0) as ctx:
        ctx.set_spaces(ni, na, ns, inp.nelec)
        ctx.set_symmetry(p.nsymrpa, p.multb, p.irpamo[:nmom], 1)
        ctx.set_integral_block("F1", F1)
        ctx.set_integral_block("F2", F2)
        out = tmp_path / "DFPCMONEW"
        got = ivo.fockivo(ctx, p.eps, src, str(out), inp.diracver, inp.occ, inp.vcut, p.nsymrpa, p.irpamo[:nmom],
                          nhomo=inp.nhomo)
    F = I.fockivo_matrix(p)
    assert np.abs(got["fock"] - F).max() < 1e-13
    f = I.read_dfpcmo(src, inp.diracver)
    for isym, (mos, w, V) in I.ivo_rotations(p, f.nfsym).items():
        assert np.abs(got["w"][mos] - w).max() < 1e-12
        assert np.abs(got["U"][np.ix_(mos, mos)] - V).max() < 1e-12
    worst = assert_same_file(str(out), ref)
    say("ivo", name, "max |d| against reference.DFPCMONEW %.2e" % worst)
