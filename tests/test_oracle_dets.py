"""Bit-exact pins for the determinant machinery of the oracle.

* ras_det_check against the reference's own unit test golden
  (test/unit_test/ras3_bitcheck/expected, copied to tests/golden/);
* Gosper enumeration against itertools.combinations (order and count);
* one_e_exct against the commutator identity of unitary-group generators,
  [E_pq, E_rs] = d_qr E_ps - d_ps E_rq, evaluated on every determinant, and
  against products of independent creation/annihilation operators.
"""
import itertools
import os

import numpy as np

from oracle import caspt2_ref as R

HERE = os.path.dirname(os.path.abspath(__file__))


def test_ras3_bitcheck_expected():
    # test/unit_test/ras3_bitcheck/active.inp: nact 8, .ras1 1,2 (ras2 empty), .ras3 3..8 with 4 electrons max
    ras = R.RasSpec(ras1_size=2, ras2_size=0, ras3_size=6, ras1_max_hole=2, ras3_max_elec=4)
    got = []
    for i in range(1, 2 ** 8):
        if R.satisfy_ras3_condition(i, ras):
            got.append("%4d%20s" % (i, format(i, "b")))
    exp = [ln.rstrip("\n") for ln in open(os.path.join(HERE, "golden", "ras3_bitcheck_expected.txt"))]
    exp = [e for e in exp if e.strip()]
    assert [g.split() for g in got] == [e.split() for e in exp]
    assert len(got) == 227


def test_gosper_order_and_count():
    for nact, nelec in [(4, 2), (8, 4), (12, 6), (14, 6), (10, 5)]:
        dets = R.search_cas_configuration(nact, nelec)
        ref = sorted(sum(1 << b for b in c) for c in itertools.combinations(range(nact), nelec))
        assert dets == ref


def _apply(op, det):
    """Apply a list of (kind, orbital) operators right-to-left; independent Jordan-Wigner."""
    sign = 1
    for kind, q in reversed(op):
        occ = (det >> q) & 1
        above = bin(det >> (q + 1)).count("1")
        if kind == "a":
            if not occ:
                return 0, 0
            det ^= 1 << q
        else:
            if occ:
                return 0, 0
            det ^= 1 << q
        sign *= -1 if above % 2 else 1
    return det, sign


def _E(det, p, q, nact):
    new, ph = R.one_e_exct(det, p + 1, q + 1, nact)
    if new == 0:
        return 0, 0
    return new, (1 if ph % 2 == 0 else -1)


def test_one_e_exct_matches_creation_annihilation():
    nact = 6
    for nelec in (1, 2, 3):
        for det in R.search_cas_configuration(nact, nelec):
            for p in range(nact):
                for q in range(nact):
                    assert _E(det, p, q, nact) == _apply([("c", p), ("a", q)], det)


def test_anticommutators_of_independent_operators():
    nact = 5
    for det in range(1 << nact):
        for p in range(nact):
            for q in range(nact):
                a = _apply([("a", p), ("c", q)], det)
                b = _apply([("c", q), ("a", p)], det)
                tot = {}
                for d, s in (a, b):
                    if s:
                        tot[d] = tot.get(d, 0) + s
                tot = {k: v for k, v in tot.items() if v}
                assert tot == ({det: 1} if p == q else {})


def test_unitary_group_commutator():
    nact, nelec = 6, 3
    dets = R.search_cas_configuration(nact, nelec)

    def Eop(p, q, vec):
        out = {}
        for d, c in vec.items():
            n, s = _E(d, p, q, nact)
            if s:
                out[n] = out.get(n, 0) + s * c
        return out

    def add(a, b, f=1):
        o = dict(a)
        for k, v in b.items():
            o[k] = o.get(k, 0) + f * v
        return {k: v for k, v in o.items() if v}
    for det in dets:
        v = {det: 1}
        for p, q, r, s in itertools.product(range(nact), repeat=4):
            lhs = add(Eop(p, q, Eop(r, s, v)), Eop(r, s, Eop(p, q, v)), -1)
            rhs = {}
            if q == r:
                rhs = add(rhs, Eop(p, s, v))
            if p == s:
                rhs = add(rhs, Eop(r, q, v), -1)
            assert lhs == rhs


def test_excitation_tables_match_one_e_exct():
    tab = R.ExcitationTables(6, 3)
    for k, d in enumerate(tab.full):
        for p in range(6):
            for q in range(6):
                n, s = _E(d, p, q, 6)
                if s == 0:
                    assert tab.tgt[p, q, k] == -1
                else:
                    assert tab.full[tab.tgt[p, q, k]] == n and tab.sgn[p, q, k] == s
