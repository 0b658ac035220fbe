"""The reference's CIDATA and caspt2_restart files (dirac_caspt2_amd/formats.py).

The restart file is the reference's own fixture (test/dev/restart_file_c32_co_dev/
caspt2_restart_33_1, kept as tests/golden/restart/caspt2_restart_33_1); the expected values
are the ones the reference's run on it printed (reference.restart_file_c32_co_dev.out:519-617):
the skipped subspaces A-C carry the file's values, and the printed c^2 and total second-order
energy are the sums over the file's A-C and the computed D-H."""
import os

import numpy as np
import pytest

from dirac_caspt2_amd import formats as F

HERE = os.path.dirname(os.path.abspath(__file__))
RESTART = os.path.join(HERE, "golden", "restart", "caspt2_restart_33_1")

# reference.restart_file_c32_co_dev.out: e2x / sumc2,x lines, c^2 and the total second-order energy
REF_E2 = {"A": -0.417201652514757e-03, "B": -0.719587322489538e-02, "C": -0.107864437530619e-02,
          "D": -0.331454181076296e-01, "E": -0.184428451675404e-01, "F": -0.334692219646566e-01,
          "G": -0.596919362573088e-01, "H": -0.774639811350566e-01}
REF_SUMC2 = {"A": 0.221008073109341e-03, "B": 0.379829088564885e-02, "C": 0.443950630863423e-03,
             "D": 0.156610171086863e-01, "E": 0.511669356797333e-02, "F": 0.126127576092678e-01,
             "G": 0.134470836409199e-01, "H": 0.870490711169087e-02}
REF_C2 = 0.06000570862815982065
REF_E2ALL = -0.23090512188490835976


def test_reference_restart_file():
    rs = F.read_restart(RESTART, totsym=33, selectroot=1)
    assert (rs.totsym, rs.selectroot, rs.next_subspace) == (33, 1, "D")
    assert sorted(rs.e2) == ["A", "B", "C"]
    for X in "ABC":
        assert abs(rs.e2[X] - REF_E2[X]) < 1e-17
        assert abs(rs.sumc2[X] - REF_SUMC2[X]) < 1e-17
    # the reference's totals: the file's A-C plus the computed D-H (r4dcaspt2_tra.F90:168)
    e2all = sum(rs.e2.values()) + sum(REF_E2[X] for X in "DEFGH")
    c2 = sum(rs.sumc2.values()) + sum(REF_SUMC2[X] for X in "DEFGH")
    assert abs(e2all - REF_E2ALL) < 1e-15
    assert abs(c2 - REF_C2) < 1e-15


def test_restart_checks(tmp_path):
    assert F.read_restart(str(tmp_path / "missing")) is None    # no file: start from scratch
    with pytest.raises(ValueError, match="totsym"):
        F.read_restart(RESTART, totsym=1)
    with pytest.raises(ValueError, match="selectroot"):
        F.read_restart(RESTART, totsym=33, selectroot=2)
    bad = tmp_path / "bad"
    bad.write_text("totsym: 1\nselectroot: 1\nNext subspace: Q\n")
    with pytest.raises(ValueError, match="Invalid subspace"):
        F.read_restart(str(bad))


def test_restart_round_trip(tmp_path):
    rs = F.Restart(7, 2, "F", e2={X: -1e-3 * (k + 1) / 3 for k, X in enumerate("ABCDE")},
                   sumc2={X: 1e-4 * (k + 1) / 7 for k, X in enumerate("ABCDE")})
    p = str(tmp_path / F.restart_filename(7, 2))
    F.write_restart(p, rs)
    back = F.read_restart(p, 7, 2)
    assert back == rs


@pytest.mark.parametrize("int_bytes", [4, 8])
@pytest.mark.parametrize("complex_build", [True, False])
def test_cidata_round_trip(tmp_path, int_bytes, complex_build):
    rng = np.random.default_rng(int_bytes)
    ndet, nroot = 70, 3
    ci = rng.normal(size=(ndet, nroot)) + (1j * rng.normal(size=(ndet, nroot)) if complex_build else 0)
    cd = F.Cidata(4, 8, 40, 4, 1, rng.normal(size=nroot), np.arange(ndet, dtype=np.int64) * 3 + 15, ci)
    p = str(tmp_path / "CIDATA_sym1")
    F.write_cidata(p, cd, int_bytes, complex_build)
    back = F.read_cidata(p, complex_build)
    assert (back.ninact, back.nact, back.nsec, back.nelec, back.totsym, back.ndet, back.nroot) == (4, 8, 40, 4, 1, ndet, nroot)
    assert np.array_equal(back.dets, cd.dets) and np.array_equal(back.ecas, cd.ecas)
    assert np.array_equal(back.ci, cd.ci)
    # the record layout of casci.F90:117-160: key records of 30 characters with 4-byte markers
    raw = open(p, "rb").read()
    assert raw[:4] == (30).to_bytes(4, "little") and raw[4:34] == b"ninact".ljust(30) and raw[34:38] == raw[:4]
    assert raw[38:42] == int_bytes.to_bytes(4, "little")


def test_cidata_checks(tmp_path):
    cd = F.Cidata(1, 2, 3, 2, 1, np.zeros(1), np.array([3], dtype=np.int64), np.ones((1, 1), dtype=complex))
    p = str(tmp_path / "c")
    F.write_cidata(p, cd)
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-38])                        # drop the key record "end" (4 + 30 + 4 bytes)
    with pytest.raises(ValueError, match="corrupted"):
        F.read_cidata(p)
    open(p, "wb").write(raw[:-1])                        # a cut record marker
    with pytest.raises(ValueError):
        F.read_cidata(p)


def test_library_restart_writer_reads_back(tmp_path):
    """dc2_write_restart (what the request server writes per subspace) gives the file the
    reference's reader layout describes: read back by formats.read_restart, every value exact,
    only the subspaces before the next one listed."""
    from dirac_caspt2_amd.api import write_restart
    from dirac_caspt2_amd.formats import read_restart
    e2 = {X: -0.1 * (k + 1) / 3.0 for k, X in enumerate("ABCDEFGH")}
    s2 = {X: 1e-3 * (k + 1) / 7.0 for k, X in enumerate("ABCDEFGH")}
    for nxt in (0, 3, 8):
        path = str(tmp_path / f"caspt2_restart_1_{nxt}")
        write_restart(path, 1, 2, nxt, e2, s2)
        rs = read_restart(path, 1, 2)
        assert rs.next_subspace == "ABCDEFGHI"[nxt]
        assert sorted(rs.e2) == list("ABCDEFGH"[:nxt])
        for X in rs.e2:
            assert rs.e2[X] == e2[X] and rs.sumc2[X] == s2[X]
