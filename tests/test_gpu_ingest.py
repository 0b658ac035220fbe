"""GPU integral ingest (dc2_set_integral_records) against the oracle's dense fill.

The oracle expands every record into its 16 images in record order
(oracle/dirac_files.py dense_from_records, module_2integrals.F90:395-434); the
library scatters the same images on the device with a deterministic
"later record wins" rule.  Blocks must agree bit for bit, and the energies
computed from the ingested blocks must match the reference goldens.
"""
import numpy as np
import pytest

from conftest import FIXTURES

pytestmark = pytest.mark.gpu


def _all_blocks(dims):
    from dirac_caspt2_amd.api import BLOCK_SHAPES, block_shape
    return [n for n in BLOCK_SHAPES if all(v > 0 for v in block_shape(dims, n))]


def _check_blocks(ctx, dims, eri, label):
    from dirac_caspt2_amd.api import BLOCK_SHAPES, space_slices
    sls = space_slices(dims)
    for name in _all_blocks(dims):
        want = eri[tuple(sls[ch] for ch in BLOCK_SHAPES[name])]
        got = ctx.integral_block(name)
        assert got.shape == want.shape, (label, name)
        bad = np.count_nonzero(got != want)
        assert bad == 0, f"{label} block {name}: {bad} of {got.size} elements differ"


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_records(name, fixture_cache):
    from gpu_helpers import say
    from dirac_caspt2_amd.api import Caspt2Context
    from oracle.dirac_files import dense_from_records
    import os
    from conftest import FIXTURE_DIR
    say("ingest fixture", name)
    fx = fixture_cache(name)
    p = fx.problem
    z = np.load(os.path.join(FIXTURE_DIR, name + ".npz"), allow_pickle=False)
    idx = z["idx"].astype(np.int32)
    val = z["val"].astype(np.complex128)
    dims = (p.ninact, p.nact, p.nsec, p.nelec)
    nmom = sum(dims[:3])
    recs = ((int(a), int(b), int(c), int(d), float(v.real), float(v.imag)) for (a, b, c, d), v in zip(idx, val))
    eri, _ = dense_from_records(nmom, recs, nmom)
    g = fx.gold
    with Caspt2Context(0) as ctx:
        ctx.set_spaces(*dims)
        if p.sym.n != 1:
            ctx.set_symmetry(p.sym.n, p.sym.multb, p.sym.irpamo, p.totsym)
        if p.ras is not None:
            r = p.ras
            ctx.set_ras(r.ras1_size, r.ras2_size, r.ras3_size, r.ras1_max_hole, r.ras3_max_elec, r.min_hole_ras1,
                        r.min_elec_ras3)
        ctx.set_one_electron(p.h, p.ecore)
        ctx.set_integral_records(idx, val)
        _check_blocks(ctx, dims, eri, name)
        eig = ctx.casci(len(g["casci"]))
        np.testing.assert_allclose(eig, g["casci"][:len(eig)], atol=1e-10, rtol=0)
        out = ctx.caspt2(p.selectroot, p.eshift)
        for k, X in enumerate("ABCDEFGH"):
            assert abs(out.e2[X] - g["e2"][k]) < 1e-9, (X, out.e2[X], g["e2"][k])
        assert abs(out.total - g["total"]) < 1e-9


@pytest.mark.parametrize("seed,nrec", [(1, 0), (2, 1), (3, 500), (4, 20000)])
def test_records_collisions_and_range(seed, nrec):
    """Random records with repeated labels (conflicting values), labels beyond the
    spinor set (dropped), and the empty list (all blocks zero)."""
    from gpu_helpers import say
    from dirac_caspt2_amd.api import Caspt2Context
    from oracle.dirac_files import dense_from_records
    say("ingest random records seed", seed, "n", nrec)
    rng = np.random.default_rng(seed)
    dims = (2, 4, 4, 2)
    nmom = sum(dims[:3])
    idx = rng.integers(1, nmom + 3, size=(nrec, 4)).astype(np.int32)   # up to nmom + 2: some dropped
    if nrec > 10:
        idx[nrec // 2:nrec // 2 + 5] = idx[:5]                         # explicit repeats
    val = rng.standard_normal(nrec) + 1j * rng.standard_normal(nrec)
    recs = [(int(a), int(b), int(c), int(d), float(v.real), float(v.imag)) for (a, b, c, d), v in zip(idx, val)]
    eri, conflicts = dense_from_records(nmom, recs, nmom)
    if nrec >= 500:
        assert conflicts > 0
    with Caspt2Context(0) as ctx:
        ctx.set_spaces(*dims)
        ctx.set_one_electron(np.eye(nmom, dtype=np.complex128), 0.0)
        ctx.set_integral_records(idx, val)
        _check_blocks(ctx, dims, eri, f"seed {seed}")
        ctx.set_integral_records(idx[:nrec // 3], val[:nrec // 3])   # a second call replaces every block
        eri2, _ = dense_from_records(nmom, recs[:nrec // 3], nmom)
        _check_blocks(ctx, dims, eri2, f"seed {seed} second call")
