"""The oracle's MDCINT -> MDCINTNEW conversion (oracle.dirac_files.mdcintnew_from_set,
create_newmdcint.F90:86-177) on the stored MDCINT sets (tests/golden/mdcint) reproduces the
records of the fixture files (tests/golden/fixtures), whose energies match the reference's
printed goldens (test_oracle_fixtures): the whole set for the two small fixtures, the first
records for the two large ones (the records are written in input order)."""
import os

import numpy as np
import pytest

from conftest import FIXTURE_DIR, FIXTURES

HERE = os.path.dirname(os.path.abspath(__file__))
MDCINT_DIR = os.path.join(HERE, "golden", "mdcint")


def load_set(name):
    from oracle.dirac_files import MdcintSet
    z = np.load(os.path.join(MDCINT_DIR, name + ".npz"), allow_pickle=False)
    ms = MdcintSet(z["ikr"], z["jkr"], z["nz"], z["indk"].astype(np.int64), z["indl"].astype(np.int64),
                   z["rklr"], None, int(z["nkr"]), z["kr"])
    return ms, z["d2c"], int(z["scheme"]), int(z["nmom"]), bool(z["whole"])


@pytest.mark.parametrize("name", FIXTURES)
def test_conversion_reproduces_fixture_records(name):
    from oracle.dirac_files import mdcintnew_from_set
    ms, d2c, scheme, nmom, whole = load_set(name)
    recs = [r for r in mdcintnew_from_set(ms, d2c, scheme) if max(r[:4]) <= nmom]
    fx = np.load(os.path.join(FIXTURE_DIR, name + ".npz"), allow_pickle=False)
    idx = np.array([r[:4] for r in recs], dtype=np.int64).reshape(-1, 4)
    val = np.array([complex(r[4], r[5]) for r in recs], dtype=np.complex128)
    if whole:
        assert len(recs) == fx["idx"].shape[0]
    assert len(recs) > 0
    np.testing.assert_array_equal(idx, fx["idx"][:len(recs)].astype(np.int64))
    np.testing.assert_array_equal(val, fx["val"][:len(recs)])
