"""The batching workload: 64 independent seeded instances (nsec ~ U[40,300], nact in
{8,10,12} with nact/2 electrons, ninact ~ U[4,24]; synth.batch_instances) solved together
through the request server with cross-request eigensolve batching and 6-slice split-precision
transforms, against each instance's plain FP64 solve and, for four instances, the oracle.

Odd-electron instances (nact 10, 5 electrons) have Kramers-degenerate CASCI roots; both paths
and the oracle represent the selected root by the same convention (the normalised projection
of one determinant onto the pair's eigenspace: oracle.caspt2_ref.root_vector), so they are
compared like the even ones.

Tolerances: batched vs plain 1e-9 Eh per subspace e2/sumc2 and in total (split precision
contributes ~1e-11); plain vs oracle 1e-9 Eh.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_instances_equal_plain_and_oracle():
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context, Server
    from dirac_caspt2_amd.problem import load_factors
    from gpu_helpers import assert_energies, oracle_problem, say
    from oracle import caspt2_ref as R
    specs = synth.batch_instances(64)
    assert len(specs) == 64 and {s.nact for s in specs} == {8, 10, 12}
    small = sorted(range(64), key=lambda k: (specs[k].nact, specs[k].nmo))
    check_oracle = [k for k in small if specs[k].nelec % 2 == 0][:2] + [k for k in small if specs[k].nelec % 2][:2]
    plain, oracle = {}, {}
    with Server(0, nslots=32, lanes_per_slot=1, eig_batch=32, eig_window=0.05) as srv:
        ids, hs = [], []
        for k, s in enumerate(specs):
            sp = synth.generate(s)
            with Caspt2Context(0) as ctx:
                load_factors(ctx, (s.ninact, s.nact, s.nsec, s.nelec), sp.h, sp.B, sp.ecore, on_device=True)
                ctx.casci(s.selectroot)
                plain[k] = ctx.caspt2(s.selectroot, s.eshift)
            ids.append(srv.add_molecule((s.ninact, s.nact, s.nsec, s.nelec), sp.B))
            hs.append((sp.h, sp.ecore))
            if k in check_oracle:
                oracle[k] = R.run_caspt2(oracle_problem(sp))
            del sp
            if k % 16 == 15:
                say(f"batch: {k + 1}/64 instances solved plain and registered")
        sub = {srv.submit(ids[k], hs[k][0], hs[k][1], s.selectroot, s.eshift, split_slices=6): k
               for k, s in enumerate(specs)}
        got = {}
        while len(got) < len(specs):
            r = srv.wait(300.0)
            assert r is not None, "timed out waiting for a batched result"
            assert r.status == 0, r.error
            got[sub[r.id]] = r
        st = srv.stats()
    say("batch: eigensolve batches", st["eig_batches"], "problems", st["eig_jobs"])
    assert st["eig_batches"] > 0 and st["eig_jobs"] > st["eig_batches"]
    worst = {"even": 0.0, "odd": 0.0}
    for k, s in enumerate(specs):
        a, b = plain[k], got[k]
        d = max(max(abs(a.e2[X] - b.e2[X]), abs(a.sumc2[X] - b.sumc2[X])) for X in "ABCDEFGH")
        d = max(d, abs(a.total - b.total))
        kind = "odd" if s.nelec % 2 else "even"
        worst[kind] = max(worst[kind], d)
        assert d < 1e-9, (k, s, d)
    say("batch: max |batched - plain|", worst)
    for k in check_oracle:
        say("batch: oracle check of instance", k, specs[k])
        assert_energies(plain[k], oracle[k], 1e-9, label=f"instance {k}")
        assert_energies(got[k], oracle[k], 1e-9, label=f"instance {k} batched")
