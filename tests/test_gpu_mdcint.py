"""MDCINT -> MDCINTNEW records on the GPU (dc2_mdcint_records, create_newmdcint.F90:86-177)
against the oracle (oracle.dirac_files.mdcintnew_from_set): the same records in the same
order, bit for bit, on the reference's stored MDCINT sets (SCHEME 6 and 4, real integrals)
and on seeded random sets (complex and real, amplifying and not, every label pattern,
values under the cutoff, empty records); with the records ingested, the energies of the two
whole fixtures match the reference's goldens (1e-9 Eh)."""
import numpy as np
import pytest

from conftest import FIXTURES
from test_oracle_mdcint import load_set

pytestmark = pytest.mark.gpu


def _oracle(ms, d2c, scheme):
    from oracle.dirac_files import mdcintnew_from_set
    recs = list(mdcintnew_from_set(ms, d2c, scheme))
    idx = np.array([r[:4] for r in recs], dtype=np.int32).reshape(-1, 4)
    val = np.array([complex(r[4], r[5]) for r in recs], dtype=np.complex128)
    return idx, val


def _gpu(ctx, ms, d2c, scheme, ingest=False):
    return ctx.mdcint_records(ms.ikr, ms.jkr, ms.nz, ms.indk, ms.indl, ms.rklr, ms.rkli, ms.nkr, ms.kr, d2c, scheme,
                              ingest=ingest)


def _same(got, want, label):
    gi, gv = got
    wi, wv = want
    assert gi.shape == wi.shape, (label, gi.shape, wi.shape)
    assert np.array_equal(gi, wi), label
    # bit for bit, signed zeros included
    assert np.array_equal(gv.view(np.uint64), wv.view(np.uint64)), label


@pytest.fixture(scope="module")
def ctx():
    from dirac_caspt2_amd.api import Caspt2Context
    with Caspt2Context(0) as c:
        yield c


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_sets(ctx, name):
    from gpu_helpers import say
    ms, d2c, scheme, nmom, whole = load_set(name)
    say("mdcint fixture", name, "records", len(ms.ikr), "entries", len(ms.indk))
    _same(_gpu(ctx, ms, d2c, scheme), _oracle(ms, d2c, scheme), name)


def random_set(seed, nkr, nrec, complex_, scheme, zero_frac=0.1):
    from oracle.dirac_files import MdcintSet
    rng = np.random.default_rng(seed)
    nmo = 2 * nkr
    lab = lambda n: rng.integers(1, nkr + 1, n) * rng.choice([-1, 1], n)
    ikr, jkr = lab(nrec), lab(nrec)
    ikr[: nrec // 4] = np.abs(jkr[: nrec // 4])          # some |ikr| = |jkr| records (never amplified)
    nz = rng.integers(0, 40, nrec)                      # empty records too
    ne = int(nz.sum())
    indk, indl = lab(ne), lab(ne)
    rr = rng.standard_normal(ne)
    rr[rng.random(ne) < zero_frac] = 1e-13              # under the cutoff
    ri = rng.standard_normal(ne) if complex_ else None
    if complex_:
        ri[rng.random(ne) < zero_frac] = -2e-13
    # kr: Kramers label m -> DIRAC spinor (a permutation of 1..nmo); d2c: DIRAC -> CAS order
    perm = rng.permutation(nmo) + 1
    kr = np.zeros(2 * nkr + 1, np.int64)
    for m in range(1, nkr + 1):
        kr[nkr + m] = perm[2 * (m - 1)]
        kr[nkr - m] = perm[2 * (m - 1) + 1]
    d2c = np.zeros(nmo + 1, np.int64)
    d2c[1:] = rng.permutation(nmo) + 1
    return MdcintSet(ikr, jkr, nz, indk, indl, rr, ri, nkr, kr), d2c


@pytest.mark.parametrize("seed,complex_,scheme", [(1, True, 4), (2, True, 6), (3, False, 4), (4, False, 6),
                                                  (5, True, 1)])
def test_random_sets(ctx, seed, complex_, scheme):
    from gpu_helpers import say
    ms, d2c = random_set(seed, nkr=9, nrec=300, complex_=complex_, scheme=scheme)
    say("mdcint random", seed, complex_, scheme, "entries", len(ms.indk))
    _same(_gpu(ctx, ms, d2c, scheme), _oracle(ms, d2c, scheme), f"seed {seed}")


def test_empty_and_errors(ctx):
    from oracle.dirac_files import MdcintSet
    ms, d2c = random_set(7, nkr=4, nrec=0, complex_=False, scheme=4)
    gi, gv = _gpu(ctx, ms, d2c, 4)
    assert gi.shape == (0, 4) and gv.shape == (0,)
    ms, d2c = random_set(8, nkr=4, nrec=20, complex_=False, scheme=4)
    bad = MdcintSet(ms.ikr, ms.jkr, ms.nz, ms.indk.copy(), ms.indl, ms.rklr, None, ms.nkr, ms.kr)
    bad.indk[len(bad.indk) // 2] = ms.nkr + 1          # a label outside kr
    with pytest.raises(Exception):
        _gpu(ctx, bad, d2c, 4)


@pytest.mark.parametrize("name", [n for n in FIXTURES if load_set(n)[4]])
def test_ingested_energies(name, fixture_cache):
    """The whole stored sets, converted and ingested on the device, give the reference's
    CASCI and CASPT2 energies."""
    from gpu_helpers import say
    from dirac_caspt2_amd.api import Caspt2Context
    say("mdcint ingest", name)
    ms, d2c, scheme, nmom, whole = load_set(name)
    fx = fixture_cache(name)
    p = fx.problem
    g = fx.gold
    dims = (p.ninact, p.nact, p.nsec, p.nelec)
    with Caspt2Context(0) as c:
        c.set_spaces(*dims)
        if p.sym.n != 1:
            c.set_symmetry(p.sym.n, p.sym.multb, p.sym.irpamo, p.totsym)
        if p.ras is not None:
            r = p.ras
            c.set_ras(r.ras1_size, r.ras2_size, r.ras3_size, r.ras1_max_hole, r.ras3_max_elec, r.min_hole_ras1,
                      r.min_elec_ras3)
        c.set_one_electron(p.h, p.ecore)
        n = c.mdcint_records(ms.ikr, ms.jkr, ms.nz, ms.indk, ms.indl, ms.rklr, None, ms.nkr, ms.kr, d2c, scheme,
                             ingest=True, want_records=False)
        assert n > 0
        eig = c.casci(len(g["casci"]))
        np.testing.assert_allclose(eig, g["casci"][:len(eig)], atol=1e-10, rtol=0)
        out = c.caspt2(p.selectroot, p.eshift)
        for k, X in enumerate("ABCDEFGH"):
            assert abs(out.e2[X] - g["e2"][k]) < 1e-9, (X, out.e2[X], g["e2"][k])
        assert abs(out.total - g["total"]) < 1e-9
