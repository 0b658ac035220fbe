"""The serving workload's request stream under -m gpu: Poisson arrivals of potential-energy-
scan points over the serving molecules (synth.serve_molecules: nsec ~ U[40,280], nact in
{8,12,14} with nact/2 electrons, ninact ~ U[4,24]; the RASPT2 molecules with RAS1/RAS2/RAS3 =
4/4/4 spinors, at most 2 holes/particles), a 20 % RASPT2 share, split-precision transforms (6
slices) and cross-request eigensolve batching, as bench.py's serve stream runs them.

* every request succeeds;
* sampled requests (CAS and RAS) recomputed alone by the plain FP64 path give the served
  energies within 1e-8 Eh per subspace e2 and in total (the split-precision bound): each result
  is the same alone and in the mix;
* the smallest CAS and the smallest RAS molecule's sampled requests match the oracle within
  1e-8 Eh.
"""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_poisson_stream_alone_mix_and_oracle():
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context, Server
    from dirac_caspt2_amd.problem import load_factors
    from gpu_helpers import say
    from oracle import caspt2_ref as R
    mols = synth.serve_molecules(12)
    sps = [synth.generate(s) for s, _ in mols]
    dirs = [synth.scan_displacements(s, s.seed + 1) for s, _ in mols]
    rng = np.random.default_rng(31)
    n, rate = 120, 30.0
    ras_m = [k for k, (_, r) in enumerate(mols) if r is not None]
    cas_m = [k for k, (_, r) in enumerate(mols) if r is None]
    which = np.where(rng.uniform(size=n) < 0.2, rng.choice(ras_m, n), rng.choice(cas_m, n))
    lams = rng.uniform(0.0, 1.0, n)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    say("stream:", n, "requests,", int(np.sum([mols[int(w)][1] is not None for w in which])), "RASPT2")
    with Server(0, nslots=32, lanes_per_slot=1, eig_batch=32, eig_window=0.003) as srv:
        ids = [srv.add_molecule((s.ninact, s.nact, s.nsec, s.nelec), sp.B) for (s, _), sp in zip(mols, sps)]
        sub, lock = {}, threading.Lock()

        def feeder():
            t0 = time.perf_counter()
            for i in range(n):
                dt = arrivals[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                m = int(which[i])
                rid = srv.submit(ids[m], synth.scan_point(sps[m], float(lams[i]), dirs[m]), sps[m].ecore, 1, 0.0,
                                 mols[m][1], split_slices=6)
                with lock:
                    sub[rid] = i
        th = threading.Thread(target=feeder)
        th.start()
        got = {}
        while len(got) < n:
            r = srv.wait(300.0)
            assert r is not None, "timed out waiting for a served result"
            assert r.status == 0, r.error
            with lock:
                got[sub[r.id]] = r
        th.join()
        st = srv.stats()
    lat = np.array([got[i].latency for i in range(n)]) * 1e3
    say("stream: p50 %.0f ms, p99 %.0f ms" % tuple(np.percentile(lat, [50, 99])), st)
    # sampled requests: 4 CAS and 4 RAS ones, recomputed alone in FP64
    sample = [i for i in range(n) if mols[int(which[i])][1] is None][:4] + \
             [i for i in range(n) if mols[int(which[i])][1] is not None][:4]
    nmo = [s.nmo for s, _ in mols]
    small = {min(cas_m, key=lambda k: nmo[k]), min(ras_m, key=lambda k: nmo[k])}
    for i in sample:
        m = int(which[i])
        s, ras = mols[m]
        h = synth.scan_point(sps[m], float(lams[i]), dirs[m])
        with Caspt2Context(0) as ctx:
            load_factors(ctx, (s.ninact, s.nact, s.nsec, s.nelec), h, sps[m].B, sps[m].ecore, on_device=True)
            if ras:
                ctx.set_ras(*ras)
            ctx.casci(1)
            alone = ctx.caspt2(1, 0.0)
        d = max(abs(alone.e2[X] - got[i].e2[X]) for X in "ABCDEFGH")
        say("stream: request", i, "molecule", m, "RAS" if ras else "CAS", "|alone - mix| %.2e" % d)
        assert d < 1e-8 and abs(alone.total - got[i].total) < 1e-8, (i, d)
    for m in small:
        i = next((i for i in sample if int(which[i]) == m), None)
        if i is None:
            continue
        s, ras = mols[m]
        p = R.Problem(s.ninact, s.nact, s.nsec, s.nelec, synth.scan_point(sps[m], float(lams[i]), dirs[m]),
                      R.FactorIntegrals(sps[m].B), sps[m].ecore, 1, ras=R.RasSpec(*ras) if ras else None)
        ref = R.run_caspt2(p)
        d = max(abs(ref.e2[X] - got[i].e2[X]) for X in "ABCDEFGH")
        say("stream: request", i, "vs oracle %.2e" % d)
        assert d < 1e-8 and abs(ref.total - got[i].total) < 1e-8, (i, d)
