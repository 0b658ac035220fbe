"""Parity of the HIP CASCI/CASPT2 path with the oracle and the reference goldens.

Tolerance: 1e-9 Eh on every per-subspace e2 and sumc2, on e0 and on the total
energy (the plain-FP64 bar of the north star); CASCI energies 1e-10 Eh;
determinant strings bit-exact.
"""
import numpy as np
import pytest

from conftest import FIXTURES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from dirac_caspt2_amd import synth
    from gpu_helpers import oracle_problem
    from oracle import caspt2_ref as R
    sp = synth.generate(synth.CONFIG_SMALL)
    p = oracle_problem(sp)
    cas = R.casci(p)
    return sp, p, cas, R.run_caspt2(p, cas)


def test_small_casci_and_dets(small):
    from gpu_helpers import context_from_synth, say
    sp, p, cas, ref = small
    say("small synthetic: CASCI")
    with context_from_synth(sp) as ctx:
        eig = ctx.casci(5)
        assert list(ctx.dets()) == cas[0]
        np.testing.assert_allclose(eig, cas[3][:5], atol=1e-10, rtol=0)
        from oracle import caspt2_ref as R
        M = ctx.casmat()
        np.testing.assert_allclose(M, R.casmat(p, cas[0]), atol=1e-12, rtol=0)


METRICS = ["formula", "determinant", "antisymmetric", "antisymmetric_det"]


@pytest.mark.parametrize("metric", METRICS)
def test_small_caspt2(small, metric):
    """Both routes of the A/C metric (RDM formulas of size nact^3 and the determinant-space
    factorisation) against the oracle."""
    from gpu_helpers import assert_energies, context_from_synth, say
    sp, p, cas, ref = small
    say("small synthetic: CASPT2, A/C metric", metric)
    with context_from_synth(sp) as ctx:
        ctx.set_option("ac_metric", metric)
        ctx.casci(1)
        out = ctx.caspt2(1, 0.0)
        say("gpu", out.e2, out.total)
        eps, U = ctx.fock_rotation()
        np.testing.assert_allclose(eps, ref.eps, atol=1e-10)
        assert_energies(out, ref, 1e-9, "small")


def test_small_caspt2_eshift(small):
    from gpu_helpers import assert_energies, context_from_synth, say
    from oracle import caspt2_ref as R
    sp, p, cas, _ = small
    say("small synthetic: CASPT2 with eshift 0.1, root 2")
    p2 = R.Problem(p.ninact, p.nact, p.nsec, p.nelec, p.h, p.eri, p.ecore, selectroot=2, eshift=0.1)
    ref = R.run_caspt2(p2, cas)
    with context_from_synth(sp) as ctx:
        ctx.casci(2)
        out = ctx.caspt2(2, 0.1)
        assert_energies(out, ref, 1e-9, "eshift")


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_goldens(name, metric, fixture_cache):
    from gpu_helpers import context_from_problem, say
    say("fixture", name, "A/C metric", metric)
    fx = fixture_cache(name)
    g = fx.gold
    with context_from_problem(fx.problem) as ctx:
        ctx.set_option("ac_metric", metric)
        eig = ctx.casci(len(g["casci"]))
        np.testing.assert_allclose(eig, g["casci"][:len(eig)], atol=1e-10, rtol=0)
        out = ctx.caspt2(fx.problem.selectroot, fx.problem.eshift)
        for k, X in enumerate("ABCDEFGH"):
            assert abs(out.e2[X] - g["e2"][k]) < 1e-9, (X, out.e2[X], g["e2"][k])
            assert abs(out.sumc2[X] - g["sumc2"][k]) < 1e-9, X
        assert abs(out.total - g["total"]) < 1e-9


@pytest.fixture(scope="module")
def medium():
    from dirac_caspt2_amd import synth
    from gpu_helpers import oracle_problem, say
    from oracle import caspt2_ref as R
    sp = synth.generate(synth.CONFIG_FP64)
    say("medium: oracle")
    return sp, R.run_caspt2(oracle_problem(sp))


@pytest.mark.parametrize("metric,schur,lanes", [("formula", "on", 3), ("determinant", "on", 3),
                                                 ("determinant", "off", 1), ("determinant", "on", 1),
                                                 ("antisymmetric", "on", 3), ("antisymmetric_det", "on", 3),
                                                 ("antisymmetric_det", "on", 1), ("determinant", "on", 4),
                                                 ("determinant", "on", 2)])
def test_medium_caspt2(medium, metric, schur, lanes):
    """192 spinors (16/12/164, 6 electrons, 924 determinants); the retained dimensions of
    every subspace agree with the oracle for both A/C metric routes, with E2 of the
    determinant-space route from the compressed resolvent (schur on: 8 of 792 directions
    are cut for A and for C) or from the B~ eigendecomposition, on 1 to 4 lanes."""
    from gpu_helpers import assert_energies, context_from_synth, say
    sp, ref = medium
    say("medium: gpu, A/C metric", metric, "schur", schur, "lanes", lanes)
    with context_from_synth(sp) as ctx:
        ctx.set_option("ac_metric", metric)
        ctx.set_option("schur", schur)
        ctx.set_option("lanes", lanes)
        ctx.casci(1)
        out = ctx.caspt2(1, 0.0)
        say("gpu", out.e2, out.total, "timings", out.timings, "info", out.info)
        assert_energies(out, ref, 1e-9, "medium")
        for X in "ABCDEFGH":
            blocks = ref.e2_blocks[X]
            if blocks:
                assert out.info[X][1] == sum(b.dimm for _, b in blocks), X   # (dimn, dimm, npert, ran)


def test_medium_lane_plan(medium):
    """B and F on D's lane (the plan chosen from 256 secondary spinors) at 192 spinors."""
    from gpu_helpers import assert_energies, context_from_synth, say
    sp, ref = medium
    say("medium: gpu, lane_plan 1")
    with context_from_synth(sp) as ctx:
        ctx.set_option("lane_plan", 1)
        ctx.casci(1)
        out = ctx.caspt2(1, 0.0)
        assert_energies(out, ref, 1e-9, "medium lane_plan 1")


@pytest.mark.parametrize("solver", ["native", "qr"])
def test_medium_all_eigensolves_on_device(medium, solver):
    """Every eigenproblem of the path on the device (an explicit eigensolver: the S and B~
    matrices up to 32 and the Fock blocks up to 64 included, which the automatic choice
    solves on the host): the oracle's energies within 1e-9 Eh."""
    from gpu_helpers import assert_energies, context_from_synth, say
    sp, ref = medium
    say("medium: gpu, eigensolver", solver)
    with context_from_synth(sp) as ctx:
        ctx.set_option("eigensolver", solver)
        ctx.casci(1)
        out = ctx.caspt2(1, 0.0)
        assert_energies(out, ref, 1e-9, "medium eigensolver " + solver)


def test_medium_early_v(medium):
    """V of A and C formed on D's lane while the transforms run (option early_v, the default
    with 4 lanes) gives the energies of V formed on their own lanes, and both the oracle's."""
    from gpu_helpers import assert_energies, context_from_synth, say
    sp, ref = medium
    outs = {}
    for ev in ("on", "off"):
        say("medium: gpu, early_v", ev)
        with context_from_synth(sp) as ctx:
            ctx.set_option("early_v", ev)
            for _ in range(2):   # the second energy reuses the lanes' buffers
                ctx.casci(1)
                outs[ev] = ctx.caspt2(1, 0.0)
            assert_energies(outs[ev], ref, 1e-9, "medium early_v " + ev)
    for X in "ABCDEFGH":
        assert abs(outs["on"].e2[X] - outs["off"].e2[X]) < 1e-12, X
        assert abs(outs["on"].sumc2[X] - outs["off"].sumc2[X]) < 1e-12, X


def test_c_metric_thin_factor():
    """14 active spinors with 6 electrons: C's metric through its thin factor (an eigenproblem
    of 14*14*13/2 = 1274 instead of C(14, 5) = 2002) gives the energies of the full Gram
    route (itself checked against the oracle) to 1e-10 Eh, with the same retained count."""
    from dirac_caspt2_amd import synth
    from gpu_helpers import context_from_synth, say
    sp = synth.generate(synth.SynthSpec(ninact=4, nact=14, nsec=40, nelec=6, seed=7))
    outs = {}
    for thin in ("on", "off"):
        say("14 active, 6 electrons: c_thin", thin)
        with context_from_synth(sp) as ctx:
            ctx.set_option("c_thin", thin)
            ctx.casci(1)
            outs[thin] = ctx.caspt2(1, 0.0)
    a, b = outs["on"], outs["off"]
    say("thin", a.e2, a.info["C"], "full", b.e2, b.info["C"])
    assert a.info["C"][1] == b.info["C"][1]
    for X in "ABCDEFGH":
        assert abs(a.e2[X] - b.e2[X]) < 1e-10, X
        assert abs(a.sumc2[X] - b.sumc2[X]) < 1e-10, X
    assert abs(a.total - b.total) < 1e-10


def test_small_davidson(small):
    """Block Davidson CASCI (forced at 70 determinants) against the oracle: the 5 lowest
    CASCI energies within 1e-10 Eh, and CASPT2 for roots 1 and 2 within 1e-9 Eh."""
    from gpu_helpers import assert_energies, context_from_synth, say
    from oracle import caspt2_ref as R
    sp, p, cas, ref = small
    with context_from_synth(sp) as ctx:
        ctx.set_option("casci_solver", "davidson")
        eig = ctx.casci(5)
        nd, nr, iters = ctx.casci_info()
        say("davidson: ndet", nd, "roots", nr, "iterations", iters)
        assert nr == 5 and iters > 0
        np.testing.assert_allclose(eig, cas[3][:5], atol=1e-10, rtol=0)
        out = ctx.caspt2(1, 0.0)
        assert_energies(out, ref, 1e-9, "davidson root 1")
        p2 = R.Problem(p.ninact, p.nact, p.nsec, p.nelec, p.h, p.eri, p.ecore, selectroot=2, eshift=0.1)
        out2 = ctx.caspt2(2, 0.1)
        assert_energies(out2, R.run_caspt2(p2, cas), 1e-9, "davidson root 2")


def test_medium_davidson_vs_dense(medium):
    """192 spinors: Davidson and the dense solver give the same CASCI energies (1e-10 Eh)
    and CASPT2 energies (1e-9 Eh)."""
    from gpu_helpers import assert_energies, context_from_synth, say
    sp, ref = medium
    with context_from_synth(sp) as ctx:
        ctx.set_option("casci_solver", "davidson")
        e_auto = ctx.casci(3)
        assert ctx.casci_info()[2] > 0
        out_auto = ctx.caspt2(1, 0.0)
        ctx.set_option("casci_solver", "dense")
        e_dense = ctx.casci(3)
        assert ctx.casci_info()[2] == 0
        out_dense = ctx.caspt2(1, 0.0)
    say("davidson vs dense casci", np.abs(e_auto - e_dense).max())
    np.testing.assert_allclose(e_auto, e_dense, atol=1e-10, rtol=0)
    assert_energies(out_auto, ref, 1e-9, "davidson")
    assert_energies(out_dense, ref, 1e-9, "dense")


@pytest.mark.parametrize("lanes", [1, 2, 3, 4])
def test_ras_determinant_metric_on_lanes(lanes):
    """RASPT2 with the A/C determinant-space route on every lane layout against the oracle.
    With RAS limits tracic projects the rotated CI vector onto the restricted space, so the
    A/C metrics must come from the rotated state (with more than one lane the metrics of
    unrestricted problems are computed from the original orbitals ahead of the rotation)."""
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context
    from dirac_caspt2_amd.problem import load_factors
    from gpu_helpers import assert_energies, say
    from oracle import caspt2_ref as R
    say("RAS determinant metric, lanes", lanes)
    spec = synth.SynthSpec(ninact=4, nact=8, nsec=30, nelec=4, seed=11)
    ras = (2, 4, 2, 1, 1, 0, 0)
    sp = synth.generate(spec)
    h = synth.scan_point(sp, 0.3, synth.scan_displacements(spec, spec.seed + 1))
    p = R.Problem(spec.ninact, spec.nact, spec.nsec, spec.nelec, h, R.FactorIntegrals(sp.B), sp.ecore, 1,
                  ras=R.RasSpec(*ras))
    ref = R.run_caspt2(p)
    with Caspt2Context(0) as ctx:
        load_factors(ctx, (spec.ninact, spec.nact, spec.nsec, spec.nelec), h, sp.B, sp.ecore)
        ctx.set_ras(*ras)
        ctx.set_option("ac_metric", "determinant")
        ctx.set_option("lanes", lanes)
        ctx.casci(1)
        out = ctx.caspt2(1, 0.0)
        assert_energies(out, ref, 1e-9, f"RAS lanes {lanes}")
