"""Seeded synthetic Kramers-paired spinor integrals (the workload generator).

Spinors come in Kramers pairs stored as consecutive indices (2k, 2k+1) in CAS
order (inactive, active, secondary; energy-ascending inside each space), the
layout read_mrconee.F90:500-541 produces.  Orbital energies are drawn per
pair: inactive U[-5,-0.5], active U[-0.5,0.3], secondary log-spaced over
[0.3,20] Eh; partners carry bit-identical energies.

One-electron h: Hermitian, diag = orbital energies, off-diagonal N(0,1e-2)
complex, with the time-reversal (quaternion) block structure
h_{p~q~} = conj(h_pq), h_{p~q} = -conj(h_{pq~}).

Two-electron integrals: (pq|rs) = sum_P B^P_pq B^P_rs, naux = 3*nspinor,
B^P Hermitian with the same quaternion structure and magnitudes
amplitude*CN(0,1)*exp(-|e_p-e_q|/4)/sqrt(naux).  This gives (pq|rs)=(rs|pq),
(pq|rs)=conj(qp|sr) and (p~q~|rs)=(qp|rs).

amplitude defaults to 0.25.  With amplitude 1 the 192-spinor workload
(16/12/164, 6 electrons) is unphysical: the two-electron terms push the lowest
secondary Fock eigenvalue to -5.05 Eh (below the inactive ones), the D
subspace returns E2 = +21.2 Eh and the smallest |alpha + w| denominator is
2.3e-4 (an intruder state).  At 0.25 every orbital space keeps its energy
window and the smallest denominator is above 0.1 Eh.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass
class SynthSpec:
    ninact: int
    nact: int
    nsec: int
    nelec: int
    seed: int = 0
    selectroot: int = 1
    eshift: float = 0.0
    shift: float = 0.0            # rigid displacement of all orbital energies (scan points)

    @property
    def nmo(self) -> int:
        return self.ninact + self.nact + self.nsec


# configurations named by the workload they describe
CONFIG_SMALL = SynthSpec(ninact=4, nact=8, nsec=40, nelec=4, seed=20240)        # 52 spinors, 70 dets
CONFIG_FP64 = SynthSpec(ninact=16, nact=12, nsec=164, nelec=6, seed=20241)      # 192 spinors, 924 dets
CONFIG_SPLIT = SynthSpec(ninact=30, nact=14, nsec=280, nelec=6, seed=20242)     # 324 spinors, 3003 dets


def _quaternion_pairs(rng, nkr: int, scale: np.ndarray, nmat: int = 1, diag_real=True):
    """nmat random Hermitian (2nkr x 2nkr) matrices with time-reversal block structure.

    Block (K,L) = [[A, B], [-conj(B), conj(A)]] and block (L,K) = block(K,L)^H;
    diagonal blocks are a*I with a real.
    """
    n = 2 * nkr
    A = (rng.standard_normal((nmat, nkr, nkr)) + 1j * rng.standard_normal((nmat, nkr, nkr))) / np.sqrt(2.0)
    Bq = (rng.standard_normal((nmat, nkr, nkr)) + 1j * rng.standard_normal((nmat, nkr, nkr))) / np.sqrt(2.0)
    A *= scale[None]
    Bq *= scale[None]
    iu = np.triu_indices(nkr, 1)
    # Hermitian in the Kramers-pair labels: upper blocks drawn, lower = H of upper
    M = np.zeros((nmat, n, n), dtype=np.complex128)
    ev, od = slice(0, n, 2), slice(1, n, 2)
    Au = np.zeros_like(A); Bu = np.zeros_like(Bq)
    Au[:, iu[0], iu[1]] = A[:, iu[0], iu[1]]
    Bu[:, iu[0], iu[1]] = Bq[:, iu[0], iu[1]]
    M[:, ev, ev] = Au
    M[:, ev, od] = Bu
    M[:, od, ev] = -np.conj(Bu)
    M[:, od, od] = np.conj(Au)
    M = M + np.conj(M.transpose(0, 2, 1))
    d = rng.standard_normal((nmat, nkr)) * np.diagonal(scale)[None]
    k = np.arange(nkr)
    M[:, 2 * k, 2 * k] = d
    M[:, 2 * k + 1, 2 * k + 1] = d
    return M


def orbital_energies(spec: SynthSpec, rng) -> np.ndarray:
    ki, ka, ks = spec.ninact // 2, spec.nact // 2, spec.nsec // 2
    ei = np.sort(rng.uniform(-5.0, -0.5, ki))
    ea = np.sort(rng.uniform(-0.5, 0.3, ka))
    es = np.geomspace(0.3, 20.0, ks) if ks else np.zeros(0)
    e = np.concatenate([ei, ea, es]) + spec.shift
    return np.repeat(e, 2)


@dataclass
class SynthProblem:
    spec: SynthSpec
    eps0: np.ndarray          # (nmo,) input orbital energies
    h: np.ndarray             # (nmo, nmo) complex
    B: np.ndarray             # (naux, nmo, nmo) complex, (pq|rs) = sum_P B[P,p,q] B[P,r,s]
    ecore: float

    def eri_block(self, s1: slice, s2: slice, s3: slice, s4: slice) -> np.ndarray:
        L = self.B[:, s1, s2]
        R = self.B[:, s3, s4]
        out = L.reshape(L.shape[0], -1).T @ R.reshape(R.shape[0], -1)
        return out.reshape(L.shape[1], L.shape[2], R.shape[1], R.shape[2])


def generate(spec: SynthSpec, naux: Optional[int] = None, amplitude: float = 0.25) -> SynthProblem:
    if spec.ninact % 2 or spec.nact % 2 or spec.nsec % 2:
        raise ValueError("space sizes must be even (Kramers pairs)")
    rng = np.random.default_rng(spec.seed)
    nmo = spec.nmo
    nkr = nmo // 2
    eps = orbital_energies(spec, rng)
    ek = eps[0::2]
    naux = 3 * nmo if naux is None else naux
    # one-electron part: off-diagonal N(0, 1e-2) complex, quaternion structured
    hs = np.full((nkr, nkr), 1.0e-2)
    h = _quaternion_pairs(rng, nkr, hs, 1)[0]
    k = np.arange(nkr)
    h[2 * k, 2 * k] = ek
    h[2 * k + 1, 2 * k + 1] = ek
    # density-fitting-like factors
    sc = amplitude * np.exp(-np.abs(ek[:, None] - ek[None, :]) / 4.0) / np.sqrt(naux)
    B = _quaternion_pairs(rng, nkr, sc, naux)
    ecore = float(rng.uniform(5.0, 10.0))
    return SynthProblem(spec, eps, h, B, ecore)


def scan_displacements(spec: SynthSpec, seed: int, sigma: float = 0.05) -> np.ndarray:
    """Per-spinor orbital-energy displacement direction of a potential-energy scan: one
    N(0, sigma) Eh value per Kramers pair, shared by both partners."""
    rng = np.random.default_rng(seed)
    d = rng.normal(0.0, sigma, spec.nmo // 2)
    return np.repeat(d, 2)


def scan_point(sp: SynthProblem, lam: float, direction: np.ndarray) -> np.ndarray:
    """One-electron matrix of scan point lam in [0, 1]: h + lam * diag(direction).  The
    two-electron integrals of the molecule are unchanged along the scan."""
    h = sp.h.copy()
    k = np.arange(sp.spec.nmo)
    h[k, k] += lam * direction
    return h


def batch_instances(n: int = 64, seed: int = 20243):
    """Independent instances of the batching workload, one seed each: nsec ~ U[40, 300],
    nact in {8, 10, 12} with nact/2 electrons, ninact ~ U[4, 24] (all even: Kramers pairs)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nact = int(rng.choice([8, 10, 12]))
        ninact = 2 * int(rng.integers(2, 13))
        nsec = 2 * int(rng.integers(20, 151))
        out.append(SynthSpec(ninact=ninact, nact=nact, nsec=nsec, nelec=nact // 2, seed=seed * 1000 + k))
    return out


# RAS restriction of the RASPT2 share of the serving workload: RAS1/RAS2/RAS3 of 4 spinors
# each (nact = 12, 6 electrons), at most 2 holes in RAS1 and 2 electrons in RAS3
SERVE_RAS = (4, 4, 4, 2, 2, 0, 0)


def serve_molecules(n: int = 12, seed: int = 20244):
    """Molecules of the serving workload: nsec ~ U[40, 280], nact in {8, 12, 14} with nact/2
    electrons, ninact ~ U[4, 24]; every third molecule is a RASPT2 one (SERVE_RAS)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        ras = k % 3 == 2
        nact = 12 if ras else int(rng.choice([8, 12, 14]))
        ninact = 2 * int(rng.integers(2, 13))
        nsec = 2 * int(rng.integers(20, 141))
        out.append((SynthSpec(ninact=ninact, nact=nact, nsec=nsec, nelec=nact // 2, seed=seed * 1000 + k),
                    SERVE_RAS if ras else None))
    return out
