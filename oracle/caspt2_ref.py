"""numpy restatement of the DIRAC-CASPT2 hot path (TEST INFRASTRUCTURE ONLY).

Every function cites the reference routine it restates (paths are relative to
the reference repository).  Indices are 0-based here; the reference is 1-based.
Spinor order is the reference's "CAS order": inactive, active, secondary,
each in ascending orbital energy (read_mrconee.F90:500-541).

Parity pinning (see tests/test_oracle_*.py):
  * determinant enumeration / RAS filter / one_e_exct: pinned bit-exact against
    test/unit_test/ras3_bitcheck/expected and against brute force;
  * casmat + cdiag: pinned against an independent second-quantisation CI
    Hamiltonian and against the CASCI energies of the runnable fixtures;
  * fockcasci/fockdiag/tracic/calce0/transforms/RDMs/solvers A-H: pinned
    end-to-end against the per-subspace e2/sumc2 goldens of the four runnable
    fixtures (real integrals, point-group symmetry, RAS), and for complex
    integrals by the invariance of every e2 under a random unitary rotation of
    the input spinors within each (space, irrep) block, plus closed forms
    (H subspace = MP2-like double sum; RDM trace identities; nelec=1 tracic).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

SMAT_LIN_DEP_THRESHOLD = 1.0e-8   # module_global_variables.F90:181
BMAT_NO_CUTOFF = 0.0              # module_global_variables.F90:182
EFFH_SKIP_C = 1.0e-10             # solve_C_subspace.F90:629

SUBSPACES = ("A", "B", "C", "D", "E", "F", "G", "H")


# ---------------------------------------------------------------------------
# symmetry tables (read_mrconee.F90:383-453)
# ---------------------------------------------------------------------------
class Symmetry:
    """Double-group multiplication tables as built by create_multiplication_table.

    ``multb`` is MULTB (2n x 2n, values 1..2n, fermion irreps 1..n, boson n+1..2n).
    ``irpamo`` holds the 1-based fermion irrep of every spinor in CAS order.
    """

    def __init__(self, nsymrpa: int, multb: np.ndarray, irpamo: Sequence[int]):
        n = int(nsymrpa)
        self.n = n
        self.multb = np.asarray(multb, dtype=np.int64).reshape(2 * n, 2 * n)
        m2 = np.zeros_like(self.multb)
        for i0 in range(2 * n):                     # read_mrconee.F90:408-413
            for j0 in range(2 * n):
                k0 = self.multb[i0, j0]
                m2[i0, k0 - 1] = j0 + 1
        self.multb2 = m2
        self.S = self.multb[n:, n:] - n             # read_mrconee.F90:428
        self.D = m2[:n, :n] - n                     # read_mrconee.F90:430
        self.DS = self.multb[n:, :n].T.copy()       # read_mrconee.F90:443-453
        self.irpamo = np.asarray(irpamo, dtype=np.int64)

    @classmethod
    def c1(cls, nmo: int) -> "Symmetry":
        return cls(1, np.array([[2, 1], [1, 2]]), np.ones(nmo, dtype=np.int64))

    def mD(self, a, b):
        return int(self.D[a - 1, b - 1])

    def mS(self, a, b):
        return int(self.S[a - 1, b - 1])

    def mDS(self, a, b):
        return int(self.DS[a - 1, b - 1])

    @staticmethod
    def partner(irp: int) -> int:
        """irp - (-1)**mod(irp,2): the Kramers-partner irrep (solve_B_subspace.F90:103)."""
        return irp + 1 if irp % 2 == 1 else irp - 1


@dataclass
class RasSpec:
    """RAS restriction (read_input_module.F90:150-172, ras_det_check.F90)."""
    ras1_size: int = 0
    ras2_size: int = 0
    ras3_size: int = 0
    ras1_max_hole: int = 0
    ras3_max_elec: int = 0
    min_hole_ras1: int = 0
    min_elec_ras3: int = 0

    @property
    def active(self) -> bool:
        return (self.ras1_size + self.ras2_size + self.ras3_size) != 0


def popcount(x: int) -> int:
    return bin(x).count("1")


def satisfy_ras1_condition(det: int, ras: RasSpec) -> bool:
    """ras_det_check.F90:8-19."""
    ras1_bit = (1 << ras.ras1_size) - 1
    n = popcount(det & ras1_bit)
    return (ras.ras1_size - ras.ras1_max_hole) <= n <= (ras.ras1_size - ras.min_hole_ras1)


def satisfy_ras3_condition(det: int, ras: RasSpec) -> bool:
    """ras_det_check.F90:20-39."""
    shift = 0
    x = det
    if ras.ras1_size != 0:
        x >>= ras.ras1_size
        shift += ras.ras1_size
    if ras.ras2_size != 0:
        x >>= ras.ras2_size
        shift += ras.ras2_size
    ras3_bit = x << shift
    n = popcount(det & ras3_bit)
    return n <= ras.ras3_max_elec and ras.min_elec_ras3 <= n


def satisfy_ras_conditions(det: int, ras: Optional[RasSpec]) -> bool:
    """search_cas_configuration.F90:92-111."""
    if ras is None:
        return True
    if ras.ras1_size != 0 and not satisfy_ras1_condition(det, ras):
        return False
    if ras.ras3_size != 0 and not satisfy_ras3_condition(det, ras):
        return False
    return True


def next_configuration(cur: int) -> int:
    """Gosper's next bit permutation as written at search_cas_configuration.F90:82-89."""
    t = (cur | (cur - 1)) + 1
    return t | ((((t & -t) // (cur & -cur)) >> 1) - 1)


def cas_configuration_totsym(det: int, nact: int, ninact: int, sym: Symmetry) -> int:
    """search_cas_configuration.F90:112-149."""
    if sym.n == 1:
        return 1
    ret = 1
    ielec = 0
    for i in range(1, nact + 1):
        if (det >> (i - 1)) & 1:
            ielec += 1
            jsym = int(sym.irpamo[ninact + i - 1])
            if ielec % 2 == 1:
                ret = sym.mDS(jsym, ret)
            else:
                ret = sym.mD(jsym + 1 if jsym % 2 == 1 else jsym - 1, ret)
    if ielec % 2 == 0:
        ret += sym.n
    return ret


def search_cas_configuration(nact: int, nelec: int, ninact: int = 0,
                             ras: Optional[RasSpec] = None,
                             sym: Optional[Symmetry] = None,
                             totsym: int = 1) -> List[int]:
    """Determinant bit strings in the reference's order (search_cas_configuration.F90:34-58)."""
    dets: List[int] = []
    cur = (1 << nelec) - 1
    while cur < (1 << nact):
        if satisfy_ras_conditions(cur, ras):
            if sym is None or sym.n == 1:
                dets.append(cur)
            elif cas_configuration_totsym(cur, nact, ninact, sym) == totsym:
                dets.append(cur)
        cur = next_configuration(cur)
    return dets


def one_e_exct(det: int, cre: int, ann: int, nact: int) -> Tuple[int, int]:
    """E_{cre,ann} on a bit string; 1-based orbital labels (one_e_exct.F90:3-51).

    Returns (new_det, phase) with new_det == 0 when the result vanishes; the
    sign is (-1)**phase.
    """
    if ann == cre and (det >> (ann - 1)) & 1:
        return det, 0
    if (det >> (ann - 1)) & 1 and not (det >> (cre - 1)) & 1:
        new = det - (1 << (ann - 1)) + (1 << (cre - 1))
        ia = (1 << nact) - (1 << ann) if ann < nact else 0
        ib = (1 << nact) - (1 << cre) if cre < nact else 0
        phase = popcount(det & ia) + popcount((det - (1 << (ann - 1))) & ib)
        return new, phase
    return 0, 0


# ---------------------------------------------------------------------------
# integrals
# ---------------------------------------------------------------------------
class Integrals:
    """Two-electron integral provider: block(s1,s2,s3,s4)[i,j,k,l] = (ij|kl)."""
    nmo: int

    def block(self, s1: slice, s2: slice, s3: slice, s4: slice) -> np.ndarray:
        raise NotImplementedError


class DenseIntegrals(Integrals):
    def __init__(self, eri: np.ndarray):
        self.eri = np.asarray(eri)
        self.nmo = self.eri.shape[0]

    def block(self, s1, s2, s3, s4):
        return np.ascontiguousarray(self.eri[s1, s2, s3, s4]).astype(np.complex128)


class FactorIntegrals(Integrals):
    """(pq|rs) = sum_P B[P,p,q] B[P,r,s] (the synthetic generator's form)."""

    def __init__(self, B: np.ndarray):
        self.B = np.asarray(B, dtype=np.complex128)
        self.nmo = self.B.shape[1]

    def block(self, s1, s2, s3, s4):
        L = self.B[:, s1, s2]
        R = self.B[:, s3, s4]
        n1, n2 = L.shape[1:]
        n3, n4 = R.shape[1:]
        out = L.reshape(L.shape[0], -1).T @ R.reshape(R.shape[0], -1)
        return out.reshape(n1, n2, n3, n4)


@dataclass
class Problem:
    ninact: int
    nact: int
    nsec: int
    nelec: int
    h: np.ndarray               # (nmo, nmo) one-electron integrals, CAS order
    eri: Integrals
    ecore: float = 0.0
    selectroot: int = 1         # 1-based
    totsym: int = 1
    eshift: float = 0.0
    sym: Optional[Symmetry] = None
    ras: Optional[RasSpec] = None

    def __post_init__(self):
        if self.sym is None:
            self.sym = Symmetry.c1(self.nmo)

    @property
    def nmo(self) -> int:
        return self.ninact + self.nact + self.nsec

    @property
    def nmoc(self) -> int:
        return self.ninact + self.nact

    @property
    def sl_i(self):
        return slice(0, self.ninact)

    @property
    def sl_a(self):
        return slice(self.ninact, self.ninact + self.nact)

    @property
    def sl_s(self):
        return slice(self.ninact + self.nact, self.nmo)

    @property
    def sl_o(self):
        return slice(0, self.ninact + self.nact)


# ---------------------------------------------------------------------------
# CASCI
# ---------------------------------------------------------------------------
def casmat(p: Problem, dets: Sequence[int]) -> np.ndarray:
    """CASCI matrix <i|H|j> restated from casmat.F90:3-222 (casmat_complex)."""
    ni, na, ne = p.ninact, p.nact, p.nelec
    O = p.eri.block(p.sl_o, p.sl_o, p.sl_o, p.sl_o)
    h = p.h
    index = {d: k for k, d in enumerate(dets)}
    n = len(dets)
    mat = np.zeros((n, n), dtype=np.complex128)
    inact = list(range(ni))
    for i, det in enumerate(dets):
        oc = [imo for imo in range(1, na + 1) if (det >> (imo - 1)) & 1]
        vi = [imo for imo in range(1, na + 1) if not (det >> (imo - 1)) & 1]
        goc = [ni + r - 1 for r in oc]
        # diagonal (casmat.F90:57-104)
        d = 0.0 + 0.0j
        for k in inact:
            d += h[k, k]
        for g in goc:
            d += h[g, g]
        lst = inact + goc
        mat0 = 0.0 + 0.0j
        for a0 in range(len(lst)):
            r = lst[a0]
            for b0 in range(a0 + 1, len(lst)):
                s = lst[b0]
                mat0 += 0.5 * O[r, r, s, s]
                mat0 -= 0.5 * O[r, s, s, r]
        mat[i, i] += d + mat0 + np.conj(mat0)
        # singles (casmat.F90:108-169)
        occ_all = np.array(lst, dtype=np.int64)
        for r in oc:
            ir = ni + r - 1
            for a in vi:
                ia = ni + a - 1
                new, ph = one_e_exct(det, a, r, na)
                j = index.get(new)
                if j is None or j <= i:
                    continue
                val = h[ir, ia]
                val += np.sum(O[ir, ia, occ_all, occ_all]) - np.sum(O[ir, occ_all, occ_all, ia])
                val = val if ph % 2 == 0 else -val
                mat[i, j] = val
                mat[j, i] = np.conj(val)
        # doubles (casmat.F90:171-213)
        for i0 in range(len(oc)):
            for j0 in range(i0 + 1, len(oc)):
                r, s = oc[i0], oc[j0]
                ir, is_ = ni + r - 1, ni + s - 1
                for k0 in range(len(vi)):
                    for l0 in range(k0 + 1, len(vi)):
                        a, b = vi[k0], vi[l0]
                        ia, ib = ni + a - 1, ni + b - 1
                        new1, ph1 = one_e_exct(det, a, r, na)
                        new2, ph2 = one_e_exct(new1, b, s, na)
                        j = index.get(new2)
                        if j is None or j <= i:
                            continue
                        val = O[ir, ia, is_, ib] - O[ir, ib, is_, ia]
                        val = val if (ph1 + ph2) % 2 == 0 else -val
                        mat[i, j] = val
                        mat[j, i] = np.conj(val)
    return mat


def cdiag(c: np.ndarray, threshold: float = 0.0):
    """ZHEEV('V','U') + linear-dependency count (diag.F90:66-128)."""
    w, v = np.linalg.eigh(c, UPLO="U")
    if threshold == 0.0:
        dimm = c.shape[0]
    else:
        dimm = int(np.count_nonzero(w >= threshold))
    return w, v, dimm


# roots closer than this (Eh) form one degenerate group (a Kramers pair for odd electron counts)
DEGENERACY_TOL = 1.0e-8


def root_vector(w: np.ndarray, v: np.ndarray, selectroot: int, tol: float = DEGENERACY_TOL) -> np.ndarray:
    """CI vector of root `selectroot` (1-based), with a convention for degenerate roots.

    A non-degenerate root is its eigenvector.  For a root inside a group of eigenvalues
    within `tol` (odd electron counts: the Kramers pair), the eigensolver's vector is an
    arbitrary unit vector of the group's eigenspace (ZHEEV's choice in the reference,
    diag.F90:109, casci.F90:54), and E2 depends on it.  The convention here takes the
    normalised projection P e_m of one determinant onto the eigenspace, P the projector of
    the group: m has the largest weight P_mm = sum_k |v_mk|^2, the lowest index among
    weights within a relative 1e-6 of that largest one (the two determinants of a Kramers
    image pair carry equal weights).  The vector does not depend on the basis the
    eigensolver returned for the group; its component m is real and positive.
    """
    k = selectroot - 1
    grp = [j for j in range(len(w)) if abs(w[j] - w[k]) < tol]
    if len(grp) == 1:
        return v[:, k]
    V = v[:, grp]
    wt = np.sum(np.abs(V) ** 2, axis=1)
    m = int(np.nonzero(wt >= wt.max() * (1.0 - 1.0e-6))[0][0])
    return V @ (np.conj(V[m, :]) / np.sqrt(wt[m]))


def casci(p: Problem):
    """casci.F90:31-72: determinant space, CI matrix, eigenpairs, total energies."""
    dets = search_cas_configuration(p.nact, p.nelec, p.ninact, p.ras, p.sym, p.totsym)
    if len(dets) == 0:
        raise ValueError("the number of CASCI configurations is 0")   # check_det_cnt
    mat = casmat(p, dets)
    w, v, _ = cdiag(mat, 0.0)
    return dets, w, v, w + p.ecore


# ---------------------------------------------------------------------------
# determinant-space operator engine (density_matrix.F90)
# ---------------------------------------------------------------------------
class ExcitationTables:
    """E_pq on every determinant of the FULL (unrestricted) nelec-in-nact space.

    density_matrix.F90 only tests membership of the FINAL determinant in the
    CAS space (dim*_density: ``exists(dict_cas_idx_reverse, i)``), so
    intermediate strings may leave a symmetry- or RAS-restricted space; the
    engine therefore works on the full string space and the CI vector is
    zero outside the selected determinants.
    """

    def __init__(self, nact: int, nelec: int):
        self.nact = nact
        full = search_cas_configuration(nact, nelec)
        self.full = full
        self.index = {d: k for k, d in enumerate(full)}
        nf = len(full)
        self.tgt = -np.ones((nact, nact, nf), dtype=np.int64)
        self.sgn = np.zeros((nact, nact, nf), dtype=np.float64)
        for k, d in enumerate(full):
            for p in range(nact):
                for q in range(nact):
                    new, ph = one_e_exct(d, p + 1, q + 1, nact)
                    if new == 0:
                        continue
                    self.tgt[p, q, k] = self.index[new]
                    self.sgn[p, q, k] = 1.0 if ph % 2 == 0 else -1.0
        self.occ = np.array([[(d >> w) & 1 for w in range(nact)] for d in full], dtype=np.float64)

    def embed(self, dets: Sequence[int], c: np.ndarray) -> np.ndarray:
        out = np.zeros(len(self.full), dtype=np.complex128)
        for k, d in enumerate(dets):
            out[self.index[d]] = c[k]
        return out

    def apply(self, p: int, q: int, X: np.ndarray) -> np.ndarray:
        """E_pq applied along the last axis of X."""
        Y = np.zeros_like(X)
        src = np.nonzero(self.tgt[p, q] >= 0)[0]
        Y[..., self.tgt[p, q, src]] = self.sgn[p, q, src] * X[..., src]
        return Y


class Densities:
    """<0|E..E|0> tensors for the rotated CI vector (density_matrix.F90:3-247).

    D1[a,b]=<E_ab>, D2[a,b,c,d]=<E_ab E_cd>, D3[a..f]=<E_ab E_cd E_ef>;
    the *F variants carry sum_w eps_w E_ww as the right-most operator
    (the eps-weighted dim(k+1)_density sums of the bXmat routines).
    """

    def __init__(self, tab: ExcitationTables, c_full: np.ndarray, eps_act: np.ndarray,
                 need_d3: bool = True):
        na = tab.nact
        self.na = na
        fvec = tab.occ @ np.asarray(eps_act, dtype=np.float64)
        cF = fvec * c_full
        cc = np.conj(c_full)
        Y1 = np.zeros((na, na, len(c_full)), dtype=np.complex128)
        Y1F = np.zeros_like(Y1)
        for p in range(na):
            for q in range(na):
                Y1[p, q] = tab.apply(p, q, c_full)
                Y1F[p, q] = tab.apply(p, q, cF)
        self.D1 = Y1 @ cc
        self.D1F = Y1F @ cc
        Y2 = np.zeros((na, na, na, na, len(c_full)), dtype=np.complex128)
        Y2F = np.zeros_like(Y2)
        for p in range(na):
            for q in range(na):
                Y2[p, q] = tab.apply(p, q, Y1)
                Y2F[p, q] = tab.apply(p, q, Y1F)
        self.D2 = Y2 @ cc
        self.D2F = Y2F @ cc
        if need_d3:
            # <E_ab X> = (E_ba |0>)^H X|0>
            bra = np.conj(Y1.transpose(1, 0, 2)).reshape(na * na, -1)
            self.D3 = (bra @ Y2.reshape(na ** 4, -1).T).reshape((na,) * 6)
            self.D3F = (bra @ Y2F.reshape(na ** 4, -1).T).reshape((na,) * 6)


# ---------------------------------------------------------------------------
# Fock matrix, diagonalisation, CI rotation, e0
# ---------------------------------------------------------------------------
def fockcasci(p: Problem, D1: np.ndarray) -> np.ndarray:
    """CASCI Fock matrix, fockcasci.F90:4-90 (occupied and secondary diagonal blocks)."""
    ni, na = p.ninact, p.nact
    nmo = p.nmo
    f = np.zeros((nmo, nmo), dtype=np.complex128)
    so, si, sa, ss = p.sl_o, p.sl_i, p.sl_a, p.sl_s
    O = p.eri.block(so, so, so, so)
    kk = np.arange(ni)
    fo = p.h[so, so].copy()
    if ni:
        fo += O[:, :, kk, kk].sum(axis=2) - O[:, kk, kk, :].sum(axis=1)
    Oa_j = O[:, :, ni:, ni:]        # (i,j,k,l) k,l active
    Oa_k = O[:, ni:, ni:, :]        # (i,l,k,j)
    fo += np.einsum("kl,ijkl->ij", D1, Oa_j) - np.einsum("kl,ilkj->ij", D1, Oa_k)
    f[so, so] = fo
    if p.nsec:
        fs = p.h[ss, ss].copy()
        if ni:
            J = p.eri.block(ss, ss, si, si)
            K = p.eri.block(ss, si, si, ss)
            fs += J[:, :, kk, kk].sum(axis=2) - K[:, kk, kk, :].sum(axis=1)
        Ja = p.eri.block(ss, ss, sa, sa)
        Ka = p.eri.block(ss, sa, sa, ss)
        fs += np.einsum("kl,ijkl->ij", D1, Ja) - np.einsum("kl,ilkj->ij", D1, Ka)
        f[ss, ss] = fs
    # only the upper triangle is computed; the lower one is its conjugate
    iu = np.triu_indices(nmo)
    up = np.zeros_like(f)
    up[iu] = f[iu]
    full = up + np.conj(np.triu(up, 1)).T
    idx = np.arange(nmo)
    full[idx, idx] = np.conj(up[idx, idx])
    return full


def canonical_gauge(w: np.ndarray, V: np.ndarray, rel_tol: float = 1.0e-8,
                    tie_tol: float = 1.0e-10) -> np.ndarray:
    """Fix the gauge of Hermitian eigenvectors independently of the eigensolver.

    Not a reference step: ZHEEV returns eigenvectors in an implementation-defined
    gauge (arbitrary phases, arbitrary mixing inside Kramers-degenerate pairs).
    Every quantity of the A-D, F-H subspaces is invariant under that gauge, but
    the E-subspace V of solve_E_subspace.F90:448-460 multiplies (ai|pj) by
    <E_tp> (not <E_pt>), which is not covariant for complex spinors.  Both the
    oracle and the GPU path therefore apply this deterministic gauge:
    eigenvalues are clustered (|w_k - w_l| <= rel_tol*max(1,|w|)); inside a
    cluster with projector P, pivot rows r_1, r_2, ... are chosen greedily as
    argmax diag(P_k) (ties within tie_tol go to the lowest row) and the basis is
    u_k = P_k e_{r_k} / |P_k e_{r_k}|, P_{k+1} = P_k - u_k u_k^H.  For a single
    vector this makes its largest component real and positive.
    """
    V = V.copy()
    n = len(w)
    i0 = 0
    while i0 < n:
        i1 = i0 + 1
        while i1 < n and abs(w[i1] - w[i0]) <= rel_tol * max(1.0, abs(w[i0])):
            i1 += 1
        Vc = V[:, i0:i1]
        P = Vc @ np.conj(Vc).T
        out = []
        for _ in range(i1 - i0):
            d = np.real(np.diag(P))
            r = int(np.nonzero(d >= d.max() - tie_tol)[0][0])
            u = P[:, r] / np.sqrt(d[r])
            out.append(u)
            P = P - np.outer(u, np.conj(u))
        V[:, i0:i1] = np.stack(out, axis=1)
        i0 = i1
    return V


def fockdiag(p: Problem, fock: np.ndarray, gauge: bool = True):
    """Per space and irrep ZHEEV of the Fock matrix (fockdiag.F90:4-78, diag.F90:230-375).

    Returns the block-diagonal transformation U (new spinor j = sum_i old i * U[i,j])
    and the orbital energies eps.  ``gauge`` applies canonical_gauge to every block.
    """
    nmo = p.nmo
    U = np.zeros((nmo, nmo), dtype=np.complex128)
    eps = np.zeros(nmo)
    for sl in (p.sl_i, p.sl_a, p.sl_s):
        lo, hi = sl.start, sl.stop
        if hi <= lo:
            continue
        for isym in range(1, p.sym.n + 1):
            ind = [i for i in range(lo, hi) if p.sym.irpamo[i] == isym]
            if not ind:
                continue
            ind = np.array(ind)
            blk = fock[np.ix_(ind, ind)]
            # facsym(i,j), i>=j from fock_cmplx; facsym(j,i)=conj (diag.F90:276-281)
            low = np.tril(blk)
            herm = low + np.conj(np.tril(low, -1)).T
            di = np.arange(len(ind))
            herm[di, di] = np.conj(low[di, di])
            w, v, _ = cdiag(herm, 0.0)
            if gauge:
                v = canonical_gauge(w, v)
            eps[ind] = w
            U[np.ix_(ind, ind)] = v
    return U, eps


def tracic(p: Problem, dets: Sequence[int], Uact: np.ndarray, c: np.ndarray) -> np.ndarray:
    """Rotate the CI vector to the new active spinors (trac.F90:76-156, sdet.F90:58-110).

    ds(k,k~) = det(Uact[occ_k, occ_k~]) and ds c~ = c.
    """
    ne = p.nelec
    occ = np.array([[w for w in range(p.nact) if (d >> w) & 1] for d in dets], dtype=np.int64)
    n = len(dets)
    ds = np.zeros((n, n), dtype=np.complex128)
    chunk = max(1, 2_000_000 // max(1, n * ne * ne))
    for k0 in range(0, n, chunk):
        k1 = min(n, k0 + chunk)
        rows = occ[k0:k1]                                       # (m, ne)
        sub = Uact[rows[:, None, :, None], occ[None, :, None, :]]  # (m, n, ne, ne)
        ds[k0:k1] = np.linalg.det(sub)
    return np.linalg.solve(ds, c)


def calce0(D1: np.ndarray, eps_act: np.ndarray) -> float:
    """e0 = sum_w eps_w Re D_ww (calce0.F90:35-43)."""
    return float(np.sum(np.real(np.diag(D1)) * eps_act))


def transform_block(T: np.ndarray, U1, U2, U3, U4) -> np.ndarray:
    """(i'j'|k'l') = sum conj(U1_ii') U2_jj' conj(U3_kk') U4_ll' (ij|kl) (module_intra.F90:141-261)."""
    T = np.einsum("ijkl,ld->ijkd", T, U4, optimize=True)
    T = np.einsum("ijkd,kc->ijcd", T, np.conj(U3), optimize=True)
    T = np.einsum("ijcd,jb->ibcd", T, U2, optimize=True)
    T = np.einsum("ibcd,ia->abcd", T, np.conj(U1), optimize=True)
    return T


# ---------------------------------------------------------------------------
# subspace solver core (solve_A_subspace.F90:150-266 and twins)
# ---------------------------------------------------------------------------
def hermitize_upper(M: np.ndarray) -> np.ndarray:
    """Keep M(i,j), j>=i; set M(j,i)=conj(M(i,j)) (this also conjugates the diagonal)."""
    up = np.triu(M)
    out = up + np.conj(np.triu(up, 1)).T
    d = np.arange(M.shape[0])
    out[d, d] = np.conj(up[d, d])
    return out


@dataclass
class BlockResult:
    e2: float
    sumc2: float
    dimn: int
    dimm: int
    smin_margin: float   # min |w_S - 1e-8| over the block (cutoff-flip margin)


def solve_block(Sf: np.ndarray, Bf: np.ndarray, V: np.ndarray, alpha: np.ndarray,
                thr: float = SMAT_LIN_DEP_THRESHOLD) -> BlockResult:
    """Common S -> cutoff -> S^-1/2 -> B~ -> eig -> E2 sequence of every solve_X_subspace.

    Sf, Bf: formula values for all (i,j) (only j>=i is used);
    V: (dimn, nperturber) with V[:, k] = vc of perturber k; alpha: (nperturber,).
    """
    dimn = Sf.shape[0]
    S = hermitize_upper(Sf)
    ws, vs, dimm = cdiag(S, thr)
    margin = float(np.min(np.abs(ws - thr))) if dimn else float("inf")
    if dimm == 0:
        return BlockResult(0.0, 0.0, dimn, 0, margin)
    B = hermitize_upper(Bf)
    keep = ws >= thr                                   # ccutoff, cutoff.F90:64-69
    uc = vs[:, keep] * (ws[keep] ** -0.5)[None, :]     # ulambda_s_half.F90:66-73
    bc1 = np.conj(uc).T @ B @ uc
    wb, W, _ = cdiag(bc1, BMAT_NO_CUTOFF)
    X = np.conj(W).T @ (np.conj(uc).T @ V)             # two gemv per perturber
    den = alpha[None, :] + wb[:, None]
    a2 = np.abs(X) ** 2
    e2 = -float(np.sum(a2 / den))
    sumc2 = float(np.sum(a2 / den ** 2))
    return BlockResult(e2, sumc2, dimn, dimm, margin)


# ---------------------------------------------------------------------------
# the transformed state handed to the solvers
# ---------------------------------------------------------------------------
@dataclass
class RotatedState:
    p: Problem
    U: np.ndarray
    eps: np.ndarray
    e0: float
    dens: Densities
    hp: np.ndarray                      # U^H h U (tramo1)
    blocks: Dict[str, np.ndarray] = field(default_factory=dict)

    def eps_i(self):
        return self.eps[self.p.sl_i]

    def eps_a(self):
        return self.eps[self.p.sl_a]

    def eps_s(self):
        return self.eps[self.p.sl_s]


# spaces: i inactive, a active, s secondary -- the 13 subspace blocks
# (r4dcaspt2_tra.F90:189-420)
SUBSPACE_BLOCKS = {
    "A1": "aiaa", "A2": "aiii", "B": "aiai",
    "C1": "saaa", "C2": "saii", "C3": "siia",
    "D1": "siaa", "D2": "saai", "D3": "siii",
    "E": "siai", "F": "sasa", "G": "sisa", "H": "sisi",
}


def rotate(p: Problem, c_casci: np.ndarray, dets: Sequence[int],
           tab: Optional[ExcitationTables] = None, need: Sequence[str] = SUBSPACES) -> RotatedState:
    """Fock build -> fockdiag -> tracic -> e0 -> transforms (r4dcaspt2_tra.F90:113-172)."""
    if tab is None:
        tab = ExcitationTables(p.nact, p.nelec)
    c0 = tab.embed(dets, c_casci)
    D1_0 = np.zeros((p.nact, p.nact), dtype=np.complex128)
    for a in range(p.nact):
        for b in range(p.nact):
            D1_0[a, b] = np.vdot(c0, tab.apply(a, b, c0))
    fock = fockcasci(p, D1_0)
    U, eps = fockdiag(p, fock)
    Uact = U[p.sl_a, p.sl_a]
    c_new = tracic(p, dets, Uact, c_casci)
    cfull = tab.embed(dets, c_new)
    dens = Densities(tab, cfull, eps[p.sl_a], need_d3=any(x in need for x in "AC"))
    e0 = calce0(dens.D1, eps[p.sl_a])
    hp = np.conj(U).T @ p.h @ U
    st = RotatedState(p, U, eps, e0, dens, hp)
    sl = {"i": p.sl_i, "a": p.sl_a, "s": p.sl_s}
    want = set()
    for X in need:
        want |= {k for k in SUBSPACE_BLOCKS if k[0] == X}
    for name in sorted(want):
        code = SUBSPACE_BLOCKS[name]
        s = [sl[ch] for ch in code]
        if any(x.stop <= x.start for x in s):
            continue
        T = p.eri.block(*s)
        Us = [U[x, x] for x in s]
        st.blocks[name] = transform_block(T, *Us)
    return st


# ---------------------------------------------------------------------------
# subspace definitions: index tuples, S/B formulas, V, alpha, symmetry filters
# ---------------------------------------------------------------------------
def _irp(p: Problem, g: int) -> int:
    return int(p.sym.irpamo[g])


def solve_A(st: RotatedState):
    """E_ti E_uv|0> (solve_A_subspace.F90:26-678)."""
    p, d = st.p, st.dens
    ni, na = p.ninact, p.nact
    ea, ei = st.eps_a(), st.eps_i()
    I = np.eye(na)
    # S(xyz,tuv) = -<E_zy E_tx E_uv> + d_tx <E_zy E_uv>       (:312-330)
    Sf = -d.D3.transpose(3, 1, 0, 2, 4, 5) + np.einsum("tx,zyuv->xyztuv", I, d.D2)
    # B = sum_w eps_w{-<E_zy E_tx E_uv E_ww> + d_tx<E_zy E_uv E_ww>} + S(e_u+e_t-e_v)  (:374-409)
    e = ea[None, None, None, None, :, None] + ea[None, None, None, :, None, None] - ea[None, None, None, None, None, :]
    Bf = -d.D3F.transpose(3, 1, 0, 2, 4, 5) + np.einsum("tx,zyuv->xyztuv", I, d.D2F) + Sf * e
    # V(i,tuv)  (:541-672)
    A1, A2 = st.blocks["A1"], st.blocks["A2"]          # (a i|a a), (a i|i i)
    effh = st.hp[p.sl_a, p.sl_i].copy()
    kk = np.arange(ni)
    effh += A2[:, :, kk, kk].sum(axis=2) - A2[:, kk, kk, :].sum(axis=1)
    V = -np.einsum("vuptqr,piqr->ituv", d.D3, A1, optimize=True)
    V += np.einsum("vupq,tipq->ituv", d.D2, A1, optimize=True)
    V -= np.einsum("vupt,pi->ituv", d.D2, effh, optimize=True)
    V += np.einsum("vu,ti->ituv", d.D1, effh, optimize=True)
    n3 = na ** 3
    Sf, Bf = Sf.reshape(n3, n3), Bf.reshape(n3, n3)
    Vt = V.reshape(ni, n3).T
    alpha = -ei - st.e0 + p.eshift
    tuples = [(x, y, z) for x in range(na) for y in range(na) for z in range(na)]

    def tuple_ok(isym, t):
        x, y, z = t
        g = lambda q: _irp(p, ni + q)
        return p.sym.mS(p.sym.mD(g(x), isym), p.sym.mD(g(y), g(z))) == 1

    def pert_ok(isym, k):
        return _irp(p, k) == isym
    return _run_blocks(p, Sf, Bf, Vt, alpha, tuples, list(range(ni)), tuple_ok, pert_ok)


def solve_B(st: RotatedState):
    """E_ti E_uj|0>, t>u, i>j (solve_B_subspace.F90:27-591)."""
    p, d = st.p, st.dens
    ni, na = p.ninact, p.nact
    ea, ei = st.eps_a(), st.eps_i()
    pairs = [(t, u) for t in range(na) for u in range(t)]
    T = np.array([q[0] for q in pairs], dtype=np.int64)
    Uu = np.array([q[1] for q in pairs], dtype=np.int64)
    X, Y = T[:, None], Uu[:, None]          # row (x,y)
    Tt, Ut = T[None, :], Uu[None, :]        # col (t,u)
    dl = lambda a, b: (a == b).astype(np.float64)
    # S(xy,tu) (:317-345)
    Sf = (d.D2[Tt, X, Ut, Y] - dl(Tt, X) * d.D1[Ut, Y] - dl(Ut, Y) * d.D1[Tt, X]
          - dl(Tt, Y) * d.D1[Ut, X] + dl(Tt, X) * dl(Ut, Y))
    # B(xy,tu) (:404-447)
    Bf = (d.D2F[Tt, X, Ut, Y] - dl(Tt, X) * d.D1F[Ut, Y] - dl(Ut, Y) * d.D1F[Tt, X]
          - dl(Tt, Y) * d.D1F[Ut, X] + dl(Tt, X) * dl(Ut, Y) * st.e0
          + Sf * (ea[Tt] + ea[Ut]))
    # V(ij,tu), i>j (:530-585)
    Bi = st.blocks["B"]                      # (p i|q j)
    Vfull = np.einsum("piqj,ptqu->ijtu", Bi, d.D2, optimize=True)
    Vfull += np.einsum("uipj,pt->ijtu", Bi, d.D1, optimize=True)
    Vfull -= np.einsum("tipj,pu->ijtu", Bi, d.D1, optimize=True)
    Vfull -= np.einsum("piuj,pt->ijtu", Bi, d.D1, optimize=True)
    Vfull += np.einsum("tiuj->ijtu", Bi) - np.einsum("uitj->ijtu", Bi)
    perts = [(i, j) for i in range(ni) for j in range(i)]
    if perts:
        Vt = np.stack([Vfull[i, j][T, Uu] for (i, j) in perts], axis=1)
    else:
        Vt = np.zeros((len(pairs), 0), dtype=np.complex128)
    alpha = np.array([-ei[i] - ei[j] - st.e0 + p.eshift for (i, j) in perts])
    sym = p.sym

    def tuple_ok(isym, q):
        t, u = q
        return sym.mD(_irp(p, ni + t), Symmetry.partner(_irp(p, ni + u))) == isym

    def pert_ok(isym, q):
        i, j = q
        return sym.mS(sym.mD(Symmetry.partner(_irp(p, i)), _irp(p, j)), isym) == 1
    return _run_blocks(p, Sf, Bf, Vt, alpha, pairs, perts, tuple_ok, pert_ok)


def solve_C(st: RotatedState):
    """E_at E_uv|0> (solve_C_subspace.F90:26-658)."""
    p, d = st.p, st.dens
    ni, na, ns = p.ninact, p.nact, p.nsec
    ea, es = st.eps_a(), st.eps_s()
    # S(xyz,tuv) = <E_zy E_xt E_uv>  (:308-330)
    Sf = d.D3.transpose(2, 1, 0, 3, 4, 5)
    e = ea[None, None, None, None, :, None] - ea[None, None, None, None, None, :] - ea[None, None, None, :, None, None]
    Bf = d.D3F.transpose(2, 1, 0, 3, 4, 5) + Sf * e           # (:384-414)
    C1 = st.blocks["C1"]                                      # (a p|q r)
    effh = st.hp[p.sl_s, p.sl_a].copy()
    if ni:
        C2, C3 = st.blocks["C2"], st.blocks["C3"]             # (a p|k k), (a k|k p)
        kk = np.arange(ni)
        effh += C2[:, :, kk, kk].sum(axis=2)
        effh -= C3[:, kk, kk, :].sum(axis=1)
    ww = np.arange(na)
    effh -= C1[:, ww, ww, :].sum(axis=1)
    V = np.einsum("ajkl,vutjkl->atuv", C1, d.D3, optimize=True)
    eff = np.where(np.abs(effh) < EFFH_SKIP_C, 0.0, effh)   # (:629)
    V += np.einsum("ap,vutp->atuv", eff, d.D2, optimize=True)
    n3 = na ** 3
    Sf, Bf = Sf.reshape(n3, n3), Bf.reshape(n3, n3)
    Vt = V.reshape(ns, n3).T
    alpha = es - st.e0 + p.eshift
    tuples = [(x, y, z) for x in range(na) for y in range(na) for z in range(na)]
    sym = p.sym

    def tuple_ok(isym, t):
        x, y, z = t
        g = lambda q: _irp(p, ni + q)
        return sym.mS(sym.mD(isym, g(x)), sym.mD(g(y), g(z))) == 1

    def pert_ok(isym, a):
        return _irp(p, ni + na + a) == isym
    return _run_blocks(p, Sf, Bf, Vt, alpha, tuples, list(range(ns)), tuple_ok, pert_ok)


def solve_D(st: RotatedState):
    """E_ai E_tu|0> (solve_D_subspace.F90:25-609)."""
    p, d = st.p, st.dens
    ni, na, ns = p.ninact, p.nact, p.nsec
    ea, ei, es = st.eps_a(), st.eps_i(), st.eps_s()
    pairs = [(t, u) for t in range(na) for u in range(na)]
    Sf = d.D2.transpose(1, 0, 2, 3).reshape(na * na, na * na)     # <E_yx E_tu> (:326-337)
    e = (ea[:, None] - ea[None, :]).reshape(1, na * na)
    Bf = d.D2F.transpose(1, 0, 2, 3).reshape(na * na, na * na) + Sf * e   # (:387-406)
    D1b, D2b, D3b = st.blocks["D1"], st.blocks["D2"], st.blocks["D3"]
    effh = st.hp[p.sl_s, p.sl_i].copy()
    kk = np.arange(ni)
    effh += D3b[:, :, kk, kk].sum(axis=2) - D3b[:, kk, kk, :].sum(axis=1)
    V = np.einsum("aipq,utpq->aitu", D1b, d.D2, optimize=True)
    V -= np.einsum("aqpi,utpq->aitu", D2b, d.D2, optimize=True)
    V += np.einsum("ai,ut->aitu", effh, d.D1, optimize=True)
    perts = [(a, i) for a in range(ns) for i in range(ni)]
    Vt = V.reshape(ns * ni, na * na).T
    alpha = (es[:, None] - ei[None, :]).reshape(-1) - st.e0 + p.eshift
    sym = p.sym

    def tuple_ok(isym, q):
        t, u = q
        return sym.mD(_irp(p, ni + t), _irp(p, ni + u)) == isym

    def pert_ok(isym, q):
        a, i = q
        return sym.mS(sym.mD(_irp(p, ni + na + a), _irp(p, i)), isym) == 1
    return _run_blocks(p, Sf, Bf, Vt, alpha, pairs, perts, tuple_ok, pert_ok)


def solve_E(st: RotatedState):
    """E_ti E_aj|0>, i>j (solve_E_subspace.F90:26-476)."""
    p, d = st.p, st.dens
    ni, na, ns = p.ninact, p.nact, p.nsec
    ea, ei, es = st.eps_a(), st.eps_i(), st.eps_s()
    I = np.eye(na)
    Sf = I - d.D1.T                                      # S(u,t) = d_ut - <E_tu> (:303-318)
    Bf = -d.D1F.T + I * st.e0 + Sf * ea[None, :]         # (:362-392)
    E = st.blocks["E"]                                   # (a i|p j)
    Vf = np.einsum("tp,aipj->aijt", d.D1, E, optimize=True)
    Vf -= np.einsum("tp,ajpi->aijt", d.D1, E, optimize=True)
    Vf += np.einsum("ajti->aijt", E) - np.einsum("aitj->aijt", E)
    perts = [(a, i, j) for i in range(ni) for j in range(i) for a in range(ns)]
    if perts:
        Vt = np.stack([Vf[a, i, j] for (a, i, j) in perts], axis=1)
    else:
        Vt = np.zeros((na, 0), dtype=np.complex128)
    alpha = np.array([es[a] - ei[i] - ei[j] - st.e0 + p.eshift for (a, i, j) in perts])
    sym = p.sym

    def tuple_ok(isym, t):
        return _irp(p, ni + t) == isym

    def pert_ok(isym, q):
        a, i, j = q
        return sym.mS(sym.mD(isym, _irp(p, i)), sym.mD(_irp(p, ni + na + a), _irp(p, j))) == 1
    return _run_blocks(p, Sf, Bf, Vt, alpha, list(range(na)), perts, tuple_ok, pert_ok)


def solve_F(st: RotatedState):
    """E_at E_bu|0>, a>b, t>u (solve_F_subspace.F90:26-544)."""
    p, d = st.p, st.dens
    ni, na, ns = p.ninact, p.nact, p.nsec
    ea, es = st.eps_a(), st.eps_s()
    pairs = [(t, u) for t in range(na) for u in range(t)]
    T = np.array([q[0] for q in pairs], dtype=np.int64)
    Uu = np.array([q[1] for q in pairs], dtype=np.int64)
    Vv, Xx = T[:, None], Uu[:, None]       # row (v,x)
    Tt, Ut = T[None, :], Uu[None, :]       # col (t,u)
    dl = lambda a, b: (a == b).astype(np.float64)
    Sf = d.D2[Vv, Tt, Xx, Ut] - dl(Xx, Tt) * d.D1[Vv, Ut]                       # (:325-340)
    Bf = d.D2F[Vv, Tt, Xx, Ut] - dl(Xx, Tt) * d.D1F[Vv, Ut] + Sf * (-ea[Ut] - ea[Tt])   # (:395-420)
    Fb = st.blocks["F"]                                  # (a p|b q)
    Vf = np.einsum("tpuq,apbq->abtu", d.D2, Fb, optimize=True)
    Vf -= np.einsum("tp,aubp->abtu", d.D1, Fb, optimize=True)
    perts = [(a, b) for a in range(ns) for b in range(a)]
    if perts:
        Vt = np.stack([Vf[a, b][T, Uu] for (a, b) in perts], axis=1)
    else:
        Vt = np.zeros((len(pairs), 0), dtype=np.complex128)
    alpha = np.array([es[a] + es[b] - st.e0 + p.eshift for (a, b) in perts])
    sym = p.sym

    def tuple_ok(isym, q):
        t, u = q
        return sym.mD(Symmetry.partner(_irp(p, ni + u)), _irp(p, ni + t)) == isym

    def pert_ok(isym, q):
        a, b = q
        return sym.mS(sym.mD(_irp(p, ni + na + a), Symmetry.partner(_irp(p, ni + na + b))), isym) == 1
    return _run_blocks(p, Sf, Bf, Vt, alpha, pairs, perts, tuple_ok, pert_ok)


def solve_G(st: RotatedState):
    """E_ai E_bt|0>, a>b (solve_G_subspace.F90:26-469)."""
    p, d = st.p, st.dens
    ni, na, ns = p.ninact, p.nact, p.nsec
    ea, ei, es = st.eps_a(), st.eps_i(), st.eps_s()
    Sf = d.D1.copy()                                     # S(i,j) = <E_{t_i t_j}> (:313-325)
    Bf = d.D1F - Sf * ea[:, None]                        # (:380-395)
    G = st.blocks["G"]                                   # (a i|b p)
    Vf = np.einsum("tp,aibp->abit", d.D1, G, optimize=True)
    Vf -= np.einsum("tp,biap->abit", d.D1, G, optimize=True)
    perts = [(a, b, i) for a in range(ns) for b in range(a) for i in range(ni)]
    if perts:
        A_ = np.array([q[0] for q in perts]); B_ = np.array([q[1] for q in perts]); I_ = np.array([q[2] for q in perts])
        Vt = Vf[A_, B_, I_].T
    else:
        Vt = np.zeros((na, 0), dtype=np.complex128)
    alpha = np.array([-ei[i] + es[a] + es[b] - st.e0 + p.eshift for (a, b, i) in perts])
    sym = p.sym

    def tuple_ok(isym, t):
        return _irp(p, ni + t) == isym

    def pert_ok(isym, q):
        a, b, i = q
        return sym.mS(sym.mD(_irp(p, ni + na + b), isym), sym.mD(_irp(p, ni + na + a), _irp(p, i))) == 1
    return _run_blocks(p, Sf, Bf, Vt, alpha, list(range(na)), perts, tuple_ok, pert_ok)


def solve_H(st: RotatedState):
    """E_ai E_bj|0>, a>b, i>j: S = 1, elementwise E2 (solve_H_subspace.F90:23-191)."""
    p = st.p
    ni, na, ns = p.ninact, p.nact, p.nsec
    ei, es = st.eps_i(), st.eps_s()
    H = st.blocks["H"]                                   # (a i|b j)
    V = H - H.transpose(0, 3, 2, 1)                      # (ai|bj) - (aj|bi)
    a_, b_ = np.tril_indices(ns, -1)
    i_, j_ = np.tril_indices(ni, -1)
    Vp = V[a_[:, None], i_[None, :], b_[:, None], j_[None, :]]
    den = (es[a_][:, None] + es[b_][:, None] - ei[i_][None, :] - ei[j_][None, :] + p.eshift)
    if p.sym.n != 1:
        g = lambda q: _irp(p, q)
        ok = np.array([[p.sym.mS(p.sym.mD(g(ni + na + a), g(i)), p.sym.mD(g(ni + na + b), g(j))) == 1
                        for (i, j) in zip(i_, j_)] for (a, b) in zip(a_, b_)], dtype=bool).reshape(Vp.shape)
    else:
        ok = np.ones(Vp.shape, dtype=bool)
    a2 = np.abs(Vp) ** 2
    e2 = -float(np.sum(np.where(ok, a2 / den, 0.0)))
    sumc2 = float(np.sum(np.where(ok, a2 / den ** 2, 0.0)))
    return e2, sumc2, []


def _run_blocks(p, Sf, Bf, Vt, alpha, tuples, perts, tuple_ok, pert_ok):
    """Loop isym = 1..nsymrpa as the reference does; C1 is one block."""
    e2 = 0.0
    sumc2 = 0.0
    blocks: List[Tuple[int, BlockResult]] = []
    nsym = p.sym.n
    for isym in range(1, nsym + 1):
        if nsym == 1:
            tsel = np.arange(len(tuples))
            psel = np.arange(len(perts))
        else:
            tsel = np.array([k for k, t in enumerate(tuples) if tuple_ok(isym, t)], dtype=np.int64)
            psel = np.array([k for k, q in enumerate(perts) if pert_ok(isym, q)], dtype=np.int64)
        if len(tsel) == 0:
            continue
        r = solve_block(Sf[np.ix_(tsel, tsel)], Bf[np.ix_(tsel, tsel)],
                        Vt[np.ix_(tsel, psel)] if len(psel) else np.zeros((len(tsel), 0), complex),
                        np.asarray(alpha)[psel] if len(psel) else np.zeros(0))
        if r.dimm == 0:          # 'If (dimm == 0) cycle' (solve_A_subspace.F90:162-168): no e2X(isym) line
            continue
        blocks.append((isym, r))
        e2 += r.e2
        sumc2 += r.sumc2
    return e2, sumc2, blocks


SOLVERS = {"A": solve_A, "B": solve_B, "C": solve_C, "D": solve_D,
           "E": solve_E, "F": solve_F, "G": solve_G, "H": solve_H}


def active_subspaces(p: Problem) -> List[str]:
    """Skipping rules of r4dcaspt2_tra.F90:174,210,240,279,318,347,376,405."""
    out = []
    for X in SUBSPACES:
        if X in "AB" and p.ninact == 0:
            continue
        if X in "CF" and p.nsec == 0:
            continue
        if X in "DEGH" and (p.ninact == 0 or p.nsec == 0):
            continue
        out.append(X)
    return out


@dataclass
class Caspt2Result:
    casci_energies: np.ndarray     # eigen(:) = ecas + ecore
    e0: float
    eps: np.ndarray
    e2: Dict[str, float]
    sumc2: Dict[str, float]
    e2_blocks: Dict[str, list]
    e2all: float
    sumc2_all: float
    weight0: float
    total_second_order: float
    total: float


def run_caspt2(p: Problem, casci_out=None, subspaces: Optional[Sequence[str]] = None) -> Caspt2Result:
    """Whole path: CASCI then CASPT2 (r4dcasci.F90, r4dcaspt2_tra.F90:164-442)."""
    if p.ninact == 0 and p.nsec == 0:
        raise ValueError("The CASPT2 energy cannot be defined when ninact = 0 and nsec = 0.")
    if casci_out is None:
        casci_out = casci(p)
    dets, w, v, eig = casci_out
    subs = active_subspaces(p) if subspaces is None else [s for s in subspaces if s in active_subspaces(p)]
    tab = ExcitationTables(p.nact, p.nelec)
    st = rotate(p, root_vector(w, v, p.selectroot), dets, tab, need=subs)
    e2: Dict[str, float] = {X: 0.0 for X in SUBSPACES}
    s2: Dict[str, float] = {X: 0.0 for X in SUBSPACES}
    blk: Dict[str, list] = {X: [] for X in SUBSPACES}
    for X in subs:
        e, s, b = SOLVERS[X](st)
        e2[X], s2[X], blk[X] = e, s, b
    e2all = sum(e2.values())
    sumc2 = sum(s2.values())
    weight0 = 1.0 / (1.0 + sumc2)
    tot2 = e2all - p.eshift * sumc2
    total = e2all + float(eig[p.selectroot - 1]) - p.eshift * sumc2
    return Caspt2Result(eig, st.e0, st.eps, e2, s2, blk, e2all, sumc2, weight0, tot2, total)
