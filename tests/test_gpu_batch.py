"""The batching workload: 64 independent seeded instances (nsec ~ U[40,300], nact in
{8,10,12} with nact/2 electrons, ninact ~ U[4,24]; synth.batch_instances) solved together
through the request server with cross-request eigensolve batching, in FP64 and with 6-slice
split-precision transforms, against each instance's plain FP64 solve and, for four instances,
the oracle.

Odd-electron instances (nact 10, 5 electrons) have Kramers-degenerate CASCI roots; both paths
and the oracle represent the selected root by the same convention (the normalised projection
of one determinant onto the pair's eigenspace: oracle.caspt2_ref.root_vector), so they are
compared like the even ones.

Tolerances: batched FP64 vs plain FP64 1e-9 Eh per subspace e2 and in total; split precision vs
plain FP64 1e-8 Eh (the bound the split-precision path is held to).  sumc2 (the norm of the
first-order function, not an energy) is compared relative to max(1, sumc2) at 1e-8: it carries
one more power of the denominators than e2, and instance 9 has a near-intruder C subspace
(sumc2_C = 6.7, an S eigenvalue 3.4e-10 from the 1e-8 cutoff), where the rounding of the two
eigensolvers (this library's for the plain solve, batched rocSOLVER in the server) moves sumc2_C
by 1.2e-9 relative while e2_C agrees to 6e-11 Eh.  Plain vs oracle 1e-9 Eh.
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
        got = {}
        for slices in (0, 6):
            sub = {srv.submit(ids[k], hs[k][0], hs[k][1], s.selectroot, s.eshift, split_slices=slices): k
                   for k, s in enumerate(specs)}
            res = {}
            while len(res) < len(specs):
                r = srv.wait(300.0)
                assert r is not None, "timed out waiting for a batched result"
                assert r.status == 0, r.error
                res[sub[r.id]] = r
            got[slices] = res
        st = srv.stats()
    say("batch: eigensolve batches", st["eig_batches"], "problems", st["eig_jobs"])
    assert st["eig_batches"] > 0 and st["eig_jobs"] > st["eig_batches"]
    bad = []
    for slices, tol in ((0, 1e-9), (6, 1e-8)):
        worst = {"even": (0.0, -1), "odd": (0.0, -1)}
        for k, s in enumerate(specs):
            a, b = plain[k], got[slices][k]
            de = max(abs(a.e2[X] - b.e2[X]) for X in "ABCDEFGH")
            de = max(de, abs(a.total - b.total))
            ds = max(abs(a.sumc2[X] - b.sumc2[X]) / max(1.0, abs(a.sumc2[X])) for X in "ABCDEFGH")
            kind = "odd" if s.nelec % 2 else "even"
            if max(de, ds) > worst[kind][0]:
                worst[kind] = (max(de, ds), k)
            if de >= tol or ds >= 1e-8:
                bad.append((slices, k, s, de, ds))
        say(f"batch: slices {slices}: max |batched - plain| (e2/total, scaled sumc2) and instance", worst)
    assert not bad, bad
    for k in check_oracle:
        say("batch: oracle check of instance", k, specs[k])
        assert_energies(plain[k], oracle[k], 1e-9, label=f"instance {k}")
        assert_energies(got[0][k], oracle[k], 1e-9, label=f"instance {k} batched")
