"""Continuous batching (dc2_server_*): every request gives the same energies alone and
in a concurrent mix, and matches the oracle.

Tolerances: alone vs in the mix 1e-10 Eh per subspace e2/sumc2 and total (the kernels
and reductions are deterministic; only the overlap differs); against the oracle 1e-9 Eh.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MOLECULES = [
    dict(spec=(4, 8, 40, 4), seed=20240, ras=None),
    dict(spec=(2, 6, 24, 2), seed=7, ras=None),
    dict(spec=(4, 8, 30, 4), seed=11, ras=(2, 4, 2, 1, 1, 0, 0)),   # RASPT2: RAS1/2/3 = 2/4/2, <= 1 hole/particle
]


SHARED = {}


@pytest.fixture(scope="module")
def mix():
    from dirac_caspt2_amd import synth
    mols = []
    for m in MOLECULES:
        spec = synth.SynthSpec(*m["spec"], seed=m["seed"])
        sp = synth.generate(spec)
        mols.append((spec, sp, synth.scan_displacements(spec, m["seed"] + 1), m["ras"]))
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(12):
        mi = int(rng.integers(0, len(mols)))
        reqs.append((mi, float(rng.uniform(0, 1)), 1 + (k % 2) * (mi != 2), 0.0 if k % 3 else 0.05))
    return mols, reqs


def _submit(srv, ids, mols, req):
    mi, lam, root, eshift = req
    spec, sp, direction, ras = mols[mi]
    from dirac_caspt2_amd import synth
    return srv.submit(ids[mi], synth.scan_point(sp, lam, direction), sp.ecore, root, eshift, ras)


def test_served_alone_equals_mixed(mix):
    from dirac_caspt2_amd.api import Server
    from gpu_helpers import say
    mols, reqs = mix
    with Server(0, nslots=4, lanes_per_slot=1) as srv:
        ids = [srv.add_molecule((s.ninact, s.nact, s.nsec, s.nelec), sp.B) for s, sp, _, _ in mols]
        say("serve: alone")
        alone = []
        for r in reqs:
            rid = _submit(srv, ids, mols, r)
            res = srv.wait(60.0)
            assert res is not None and res.id == rid and res.status == 0, res
            alone.append(res)
        say("serve: mixed")
        sub = {_submit(srv, ids, mols, r): k for k, r in enumerate(reqs)}
        got = {}
        while len(got) < len(reqs):
            res = srv.wait(60.0)
            assert res is not None, "timed out waiting for a served result"
            assert res.status == 0, res.error
            got[sub[res.id]] = res
        st = srv.stats()
        say("serve: stats", st, "slots used", sorted({g.slot for g in got.values()}))
        assert st["pending"] == 0 and st["running"] == 0
    for k in range(len(reqs)):
        a, b = alone[k], got[k]
        for X in "ABCDEFGH":
            assert abs(a.e2[X] - b.e2[X]) < 1e-10, (k, X)
            assert abs(a.sumc2[X] - b.sumc2[X]) < 1e-10, (k, X)
        assert abs(a.total - b.total) < 1e-10, k
    # concurrency: some requests of the mix overlapped in time
    spans = sorted((g.t_start, g.t_end) for g in got.values())
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    SHARED["alone"] = alone


def test_served_matches_oracle(mix):
    from dirac_caspt2_amd import synth
    from oracle import caspt2_ref as R
    mols, reqs = mix
    alone = SHARED.get("alone")
    if alone is None:
        pytest.skip("needs test_served_alone_equals_mixed")
    for k in (0, 1, 2, 5):
        mi, lam, root, eshift = reqs[k]
        spec, sp, direction, ras = mols[mi]
        p = R.Problem(spec.ninact, spec.nact, spec.nsec, spec.nelec, synth.scan_point(sp, lam, direction),
                      R.FactorIntegrals(sp.B), sp.ecore, root, eshift=eshift,
                      ras=R.RasSpec(*ras) if ras else None)
        ref = R.run_caspt2(p)
        for X in "ABCDEFGH":
            assert abs(alone[k].e2[X] - ref.e2[X]) < 1e-9, (k, X, alone[k].e2[X], ref.e2[X])
        assert abs(alone[k].total - ref.total) < 1e-9, k


def test_served_with_batched_eigensolves(mix):
    """The same mix with the eigensolves of concurrent requests batched (groups of equal
    size solved by one batched rocSOLVER call): every energy within 1e-9 Eh of the request
    served alone without batching (eigenvectors may differ by rounding-level rotations
    inside degenerate sets; every E2 here is invariant under them or gauge-fixed)."""
    from dirac_caspt2_amd.api import Server
    from gpu_helpers import say
    mols, reqs = mix
    alone = SHARED.get("alone")
    if alone is None:
        pytest.skip("needs test_served_alone_equals_mixed")
    with Server(0, nslots=6, lanes_per_slot=1, eig_batch=8, eig_window=0.002) as srv:
        ids = [srv.add_molecule((s.ninact, s.nact, s.nsec, s.nelec), sp.B) for s, sp, _, _ in mols]
        sub = {_submit(srv, ids, mols, r): k for k, r in enumerate(reqs)}
        got = {}
        while len(got) < len(reqs):
            res = srv.wait(60.0)
            assert res is not None, "timed out waiting for a served result"
            assert res.status == 0, res.error
            got[sub[res.id]] = res
        st = srv.stats()
    say("serve batched: stats", st)
    assert st["eig_batches"] > 0 and st["eig_jobs"] > st["eig_batches"]
    for k in range(len(reqs)):
        for X in "ABCDEFGH":
            assert abs(alone[k].e2[X] - got[k].e2[X]) < 1e-9, (k, X)
        assert abs(alone[k].total - got[k].total) < 1e-9, k


def test_admission_counts_memory_slots_hold(mix):
    """Slots keep their scratch between requests, so admission counts what every slot holds:
    with a budget of one large request plus a little, two slots first run two small
    requests each (each slot then holds a small footprint), then a large request that fits
    only after an idle slot releases its memory.  The held memory never exceeds the budget
    and every energy equals the request served alone (1e-10 Eh)."""
    from dirac_caspt2_amd.api import Server, request_footprint
    from gpu_helpers import say
    mols, reqs = mix
    alone = SHARED.get("alone")
    if alone is None:
        pytest.skip("needs test_served_alone_equals_mixed")
    fp = [request_footprint(s.ninact, s.nact, s.nsec, s.nelec) for s, _, _, _ in mols]
    big = int(np.argmax(fp))
    small = int(np.argmin(fp))
    assert fp[small] < fp[big]
    budget = fp[big] + fp[small] // 2        # one large request, not a large and a small one
    ks = [k for k, r in enumerate(reqs) if r[0] == small][:2]
    kb = [k for k, r in enumerate(reqs) if r[0] == big][:1]
    assert ks and kb
    with Server(0, nslots=2, mem_budget=budget, lanes_per_slot=1) as srv:
        ids = [srv.add_molecule((s.ninact, s.nact, s.nsec, s.nelec), sp.B) for s, sp, _, _ in mols]
        got = {}
        for order in (ks + ks, kb):   # the small requests on both slots, then the large one
            sub = {_submit(srv, ids, mols, reqs[k]): k for k in order}
            while sub:
                res = srv.wait(60.0)
                assert res is not None, "timed out waiting for a served result"
                assert res.status == 0, res.error
                got.setdefault(sub.pop(res.id), res)
        st = srv.stats()
    say("serve admission: stats", st, "budget", budget, "footprints", fp)
    assert st["peak"] <= budget
    for k, res in got.items():
        for X in "ABCDEFGH":
            assert abs(alone[k].e2[X] - res.e2[X]) < 1e-10, (k, X)
        assert abs(alone[k].total - res.total) < 1e-10, k


def test_served_checkpoints_and_resume(mix, tmp_path):
    """With a checkpoint directory each request writes caspt2_restart_1_<root> as its subspaces
    finish (the completed prefix); resubmitting a request from such a file (its first four
    subspaces taken from it) gives the uninterrupted energies within 1e-10 Eh."""
    from dirac_caspt2_amd.api import Server
    from dirac_caspt2_amd.formats import read_restart
    from gpu_helpers import say
    mols, reqs = mix
    spec, sp, direction, ras = mols[0]
    from dirac_caspt2_amd import synth
    h = synth.scan_point(sp, 0.3, direction)
    with Server(0, nslots=2, lanes_per_slot=4) as srv:
        mid = srv.add_molecule((spec.ninact, spec.nact, spec.nsec, spec.nelec), sp.B)
        srv.set_checkpoint_dir(str(tmp_path))
        rid = srv.submit(mid, h, sp.ecore, 1, 0.0)
        full = srv.wait(120.0)
        assert full is not None and full.id == rid and full.status == 0, full
        rs = read_restart(str(tmp_path / str(rid) / "caspt2_restart_1_1"), 1, 1)
        say("checkpoint after the request:", rs.next_subspace, sorted(rs.e2))
        assert rs.next_subspace == "I" and sorted(rs.e2) == list("ABCDEFGH")
        for X in "ABCDEFGH":
            assert rs.e2[X] == full.e2[X] and rs.sumc2[X] == full.sumc2[X]
        part = ({X: rs.e2[X] for X in "ABCD"}, {X: rs.sumc2[X] for X in "ABCD"})
        rid2 = srv.submit(mid, h, sp.ecore, 1, 0.0, restart=(4, part[0], part[1]))
        res = srv.wait(120.0)
        assert res is not None and res.id == rid2 and res.status == 0, res
        rs2 = read_restart(str(tmp_path / str(rid2) / "caspt2_restart_1_1"), 1, 1)
        assert rs2.next_subspace == "I"
    for X in "ABCDEFGH":
        assert abs(res.e2[X] - full.e2[X]) < 1e-10, X
        assert abs(res.sumc2[X] - full.sumc2[X]) < 1e-10, X
    assert abs(res.total - full.total) < 1e-10
