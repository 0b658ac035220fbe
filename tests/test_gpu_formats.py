"""CASCI results through CIDATA files and CASPT2 resumed from caspt2_restart files, on the GPU.

A CASCI result written as CIDATA (casci.F90:117-160) and read back into a fresh context
(dc2_set_casci_result) must give the same CASPT2 energies as the context that computed it;
a CASPT2 resumed at any subspace with the earlier subspaces' values from a restart file
(dc2_caspt2_restart) must give the full run's energies, and, as in the reference
(r4dcaspt2_tra.F90:168), a file line for a subspace that is computed again is counted twice."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from dirac_caspt2_amd import synth
    return synth.generate(synth.CONFIG_SMALL)


def test_cidata_round_trip(small, tmp_path):
    from dirac_caspt2_amd import formats as F
    from gpu_helpers import context_from_synth, say
    say("CIDATA round trip")
    sp = small
    s = sp.spec
    with context_from_synth(sp) as ctx:
        eig = ctx.casci(2)
        ref = ctx.caspt2(2, 0.05)
        cd = F.Cidata(s.ninact, s.nact, s.nsec, s.nelec, 1, eig - sp.ecore, ctx.dets(),
                      np.stack([ctx.ci_vector(1), ctx.ci_vector(2)], axis=1))
    path = str(tmp_path / "CIDATA_sym1")
    F.write_cidata(path, cd)
    back = F.read_cidata(path)
    with context_from_synth(sp) as ctx:
        ctx.set_casci_result(back.dets, back.ci, back.ecas)
        out = ctx.caspt2(2, 0.05)
        for X in "ABCDEFGH":
            assert abs(out.e2[X] - ref.e2[X]) < 1e-12, X
            assert abs(out.sumc2[X] - ref.sumc2[X]) < 1e-12, X
        assert abs(out.total - ref.total) < 1e-12
        bad = back.dets.copy()
        bad[[0, 1]] = bad[[1, 0]]
        with pytest.raises(RuntimeError, match="determinant"):
            ctx.set_casci_result(bad, back.ci, back.ecas)


@pytest.mark.parametrize("nxt", ["A", "B", "D", "G", "H", "I"])
def test_restart_resume(small, nxt, tmp_path):
    from dirac_caspt2_amd import formats as F
    from gpu_helpers import context_from_synth, say
    say("restart at", nxt)
    with context_from_synth(small) as ctx:
        ctx.casci(1)
        full = ctx.caspt2(1, 0.1)
        done = "ABCDEFGH"[:"ABCDEFGHI".index(nxt)]
        rs = F.Restart(1, 1, nxt, e2={X: full.e2[X] for X in done}, sumc2={X: full.sumc2[X] for X in done})
        path = str(tmp_path / F.restart_filename(1, 1))
        F.write_restart(path, rs)
        back = F.read_restart(path, 1, 1)
        out = ctx.caspt2_restart(1, 0.1, back.next_subspace, back.e2, back.sumc2)
        for X in "ABCDEFGH":
            assert abs(out.e2[X] - full.e2[X]) < 1e-12, X
            assert abs(out.sumc2[X] - full.sumc2[X]) < 1e-12, X
        assert abs(out.total - full.total) < 1e-12
        assert abs(out.weight0 - full.weight0) < 1e-12
        if nxt <= "H":   # a line for a subspace that is computed again counts twice, as in the reference
            X = nxt
            twice = ctx.caspt2_restart(1, 0.1, nxt, {**back.e2, X: full.e2[X]}, {**back.sumc2, X: full.sumc2[X]})
            assert abs(twice.e2all - (full.e2all + full.e2[X])) < 1e-12
            assert abs(twice.e2[X] - full.e2[X]) < 1e-12
