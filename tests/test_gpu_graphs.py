"""Captured block-Davidson iterations (option "graphs"): the same launches replayed from
graphs give the same CASCI and CASPT2 energies as issued one by one, on every replay, with
the program's own eager work and synchronisations between the replays (each graph is
replayed in every later CASCI of the context).  Bit-identical results are expected (same
kernels, same order); the bound is 1e-12 Eh."""
import pytest

pytestmark = pytest.mark.gpu


def _energies(graphs, rounds):
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context
    from dirac_caspt2_amd.problem import load_factors
    spec = synth.SynthSpec(4, 12, 24, 6, seed=31)    # 924 determinants: the Davidson path
    sp = synth.generate(spec)
    out = []
    with Caspt2Context(0) as ctx:
        load_factors(ctx, (spec.ninact, spec.nact, spec.nsec, spec.nelec), sp.h, sp.B, sp.ecore, on_device=True)
        ctx.set_option("graphs", "on" if graphs else "off")
        for _ in range(rounds):
            e0 = ctx.casci(1)[0]
            assert ctx.casci_info()[2] > 0, "the dense solver ran, not Davidson"
            e = ctx.caspt2(1, 0.0)
            out.append((e0, e.total, dict(e.e2)))
        stats = ctx.graph_stats()
    return out, stats


def test_graph_replays_equal_direct_launches():
    from gpu_helpers import say
    ref, st0 = _energies(False, 1)
    got, st = _energies(True, 4)
    say("graphs:", st, "direct:", st0)
    assert st0["replays"] == 0
    assert st["enabled"] == 1 and st["captures"] > 0 and st["replays"] > st["captures"]
    e0, tot, e2 = ref[0]
    for k, (g0, gt, ge2) in enumerate(got):
        assert abs(g0 - e0) < 1e-12, k
        assert abs(gt - tot) < 1e-12, k
        for X in "ABCDEFGH":
            assert abs(ge2[X] - e2[X]) < 1e-12, (k, X)
