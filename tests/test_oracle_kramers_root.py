"""The degenerate-root convention (oracle.caspt2_ref.root_vector) for odd electron counts.

* the CASCI roots of an odd-electron synthetic problem come in degenerate (Kramers) pairs;
* the convention's vector lies in the pair's eigenspace (H psi = E psi) and does not depend
  on the basis the eigensolver returned for the pair: mixing the two eigenvectors by a
  random 2 x 2 unitary leaves it unchanged to 1e-12, and with it every subspace E2;
* a plain eigenvector of the pair is NOT basis independent in E2 (so the convention pins
  something): two bases of the pair give E2 values that differ by far more than 1e-9.
"""
import numpy as np

from dirac_caspt2_amd import synth
from oracle import caspt2_ref as R


def _problem(seed=3):
    spec = synth.SynthSpec(ninact=2, nact=6, nsec=10, nelec=3, seed=seed)
    sp = synth.generate(spec)
    return R.Problem(spec.ninact, spec.nact, spec.nsec, spec.nelec, sp.h, R.FactorIntegrals(sp.B), sp.ecore, 1)


def _mix(v, rng):
    q, _ = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))
    out = v.copy()
    out[:, :2] = v[:, :2] @ q
    return out


def test_odd_electron_roots_are_kramers_pairs():
    p = _problem()
    dets, w, v, eig = R.casci(p)
    assert abs(w[0] - w[1]) < 1e-10 and abs(w[2] - w[3]) < 1e-10
    assert w[2] - w[1] > 1e-4


def test_root_vector_is_basis_independent_eigenvector():
    p = _problem()
    dets, w, v, eig = R.casci(p)
    H = R.casmat(p, dets)
    psi = R.root_vector(w, v, 1)
    assert abs(np.linalg.norm(psi) - 1.0) < 1e-13
    assert np.max(np.abs(H @ psi - w[0] * psi)) < 1e-11
    rng = np.random.default_rng(0)
    for _ in range(3):
        vm = _mix(v, rng)
        assert np.max(np.abs(R.root_vector(w, vm, 1) - psi)) < 1e-12
        assert np.max(np.abs(R.root_vector(w, vm, 2) - psi)) < 1e-12   # either root of the pair


def test_e2_of_convention_is_basis_independent_and_plain_vector_is_not():
    p = _problem()
    cas = R.casci(p)
    dets, w, v, eig = cas
    ref = R.run_caspt2(p, cas)
    rng = np.random.default_rng(1)
    vm = _mix(v, rng)
    mixed = R.run_caspt2(p, (dets, w, vm, eig))
    for X in "ABCDEFGH":
        assert abs(mixed.e2[X] - ref.e2[X]) < 1e-11, X
    # without the convention: the solver's raw vectors of two bases
    tab = R.ExcitationTables(p.nact, p.nelec)
    e_raw = []
    for vv in (v, vm):
        st = R.rotate(p, vv[:, 0], dets, tab, need=list("ABCDEFGH"))
        e_raw.append(sum(R.SOLVERS[X](st)[0] for X in "ABCDEFGH"))
    assert abs(e_raw[0] - e_raw[1]) > 1e-7
