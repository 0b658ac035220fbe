"""Pin the oracle against the reference's own regression outputs.

The four fixtures are the only CASPT2 test directories of the reference whose
integral files are complete (real integrals, double-group symmetry, RAS, MDCINT
schemes 4 and 6).  Their reference.<name>.out files print every e2X, sumc2,X,
e2X(isym), the CASCI energies and the total energy; tests/golden/make_fixtures.py
copied those numbers and the inputs into tests/golden/fixtures/*.npz.
The reference's own tests assert the total energy to 1e-10 Eh
(e.g. test/dev/c2_h2o_dev/test_c2_h2o_dev.py:23); we assert every printed
per-subspace number to 1e-12 Eh (they are printed with 15 significant digits).
"""
import numpy as np
import pytest

from conftest import FIXTURES
from oracle import caspt2_ref as R


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_energies(name, fixture_cache):
    fx = fixture_cache(name)
    assert fx.conflicts == 0, "symmetry images of the integral records disagree"
    res = R.run_caspt2(fx.problem)
    g = fx.gold
    ncas = len(g["casci"])
    np.testing.assert_allclose(res.casci_energies[:ncas], g["casci"], rtol=0, atol=1e-10)
    for k, X in enumerate("ABCDEFGH"):
        assert abs(res.e2[X] - g["e2"][k]) < 1e-12 + 1e-13 * abs(g["e2"][k]), X
        assert abs(res.sumc2[X] - g["sumc2"][k]) < 1e-12 + 1e-13 * abs(g["sumc2"][k]), X
        blocks = [b.e2 for _, b in res.e2_blocks[X]]
        gb = g["blocks_" + X]
        if len(gb):
            np.testing.assert_allclose(blocks, gb, rtol=1e-12, atol=1e-14, err_msg=X)
    assert abs(res.sumc2_all - g["c2"]) < 1e-12
    assert abs(res.total_second_order - g["total2"]) < 1e-10
    assert abs(res.total - g["total"]) < 1e-10


def test_subspace_skipping_rules(fixture_cache):
    # ninact = 0: only C and F are defined (r4dcaspt2_tra.F90:174-431)
    fx = fixture_cache("ninact_0_h2_dev")
    assert R.active_subspaces(fx.problem) == ["C", "F"]


def test_c1_treatment_equals_symmetry_blocks(fixture_cache, monkeypatch):
    """Solving each subspace as one block (the GPU path) equals the per-irrep blocks."""
    fx = fixture_cache("c2_h2o_dev")
    ref = R.run_caspt2(fx.problem)

    def one_block(p, Sf, Bf, Vt, alpha, tuples, perts, tok, pok):
        r = R.solve_block(Sf, Bf, Vt, np.asarray(alpha) if len(perts) else np.zeros(0))
        return r.e2, r.sumc2, [(1, r)]
    monkeypatch.setattr(R, "_run_blocks", one_block)
    one = R.run_caspt2(fx.problem)
    for X in "ABCDEFGH":
        assert abs(one.e2[X] - ref.e2[X]) < 1e-13, X
