"""IVO (improved virtual orbitals) restatement, TEST INFRASTRUCTURE ONLY.

fockivo.F90:4-409 driven by r4divo.F90:1-114: the secondary-space Fock matrix with the
HOMO electrons removed (fockivo.F90:40-78), diagonalised per irreducible representation
(fockivo.F90:190-230, cdiag = ZHEEV), and the DIRAC MO coefficient file DFPCMO rewritten
with the virtual MOs rotated by the eigenvectors (fockivo.F90:232-330) as DFPCMONEW.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .dirac_files import cas_order, dense_from_records, mdcintnew_records, read_mrconee_file


@dataclass
class IvoInput:
    ninact: int = 0
    nact: int = 0
    nsec: int = 0
    nelec: int = 0
    diracver: int = 21
    scheme: Optional[int] = None
    nhomo: int = 0
    occ: List[int] = field(default_factory=lambda: [0, 0])     # .nocc / .noccg, .noccu
    vcut: List[int] = field(default_factory=lambda: [0, 0])    # .nvcut / .nvcutg, .nvcutu

    @property
    def mdcint_scheme(self) -> int:
        if self.scheme is not None:
            return self.scheme
        return 6 if self.diracver <= 22 else 4


def read_ivo_input(path: str) -> IvoInput:
    """The keywords of active.ivo.inp the IVO path reads (read_input_module.F90)."""
    lines = [ln.strip() for ln in open(path).read().splitlines()]
    inp = IvoInput()
    keys = {".ninact": "ninact", ".nact": "nact", ".nsec": "nsec", ".nelec": "nelec", ".diracver": "diracver",
            ".scheme": "scheme", ".nhomo": "nhomo"}
    for i, ln in enumerate(lines[:-1]):
        k = ln.lower()
        nxt = lines[i + 1].split()[0] if lines[i + 1].split() else ""
        if k in keys:
            setattr(inp, keys[k], int(nxt))
        elif k in (".nocc", ".noccg"):
            inp.occ[0] = int(nxt)
        elif k == ".noccu":
            inp.occ[1] = int(nxt)
        elif k in (".nvcut", ".nvcutg"):
            inp.vcut[0] = int(nxt)
        elif k == ".nvcutu":
            inp.vcut[1] = int(nxt)
    return inp


@dataclass
class Dfpcmo:
    """DIRAC MO coefficient file (formatted), the layout fockivo.F90:113-178 reads."""
    header: List[str]        # INFO line and title line (DIRAC >= 21), or the title only
    nfsym: int
    b: Optional[int]
    pos: List[int]
    ele: List[int]
    bas: List[int]
    energy_line: str
    coef_label: Optional[str]
    buf: np.ndarray
    eval_label: Optional[str]
    evals: np.ndarray
    sym_label: Optional[str]
    syminfo: np.ndarray
    kappa_label: Optional[str] = None
    kappa: Optional[np.ndarray] = None


def _floats(tokens):
    return np.array([float(t.replace("D", "E")) for t in tokens])


def read_dfpcmo(path: str, diracver: int) -> Dfpcmo:
    lines = open(path).read().splitlines()
    new = diracver >= 21
    p = 0
    header = []
    if new:
        header.append(lines[p]); p += 1
    header.append(lines[p]); p += 1
    dims = [int(x) for x in lines[p].split()]; p += 1
    nfsym = dims[0]
    if new:
        b, rest = dims[1], dims[2:]
    else:
        b, rest = None, dims[1:]
    pos, ele, bas = rest[0::3][:nfsym], rest[1::3][:nfsym], rest[2::3][:nfsym]
    energy_line = lines[p]; p += 1
    total_mo = sum(pos) + sum(ele)
    total_ao = sum((pp + ee) * bb for pp, ee, bb in zip(pos, ele, bas))

    def label():
        nonlocal p
        if not new:
            return None
        s = lines[p]; p += 1
        return s

    def numbers(count, conv):
        nonlocal p
        out = []
        while len(out) < count:
            out.extend(lines[p].split()); p += 1
        return conv(out[:count])

    coef_label = label()
    buf = numbers(total_ao, _floats)
    eval_label = label()
    evals = numbers(total_mo, _floats)
    sym_label = label()
    syminfo = numbers(total_mo, lambda t: np.array([int(x) for x in t]))
    kappa_label, kappa = None, None
    if p < len(lines) and lines[p].strip():
        kappa_label = label()
        kappa = numbers(total_mo, lambda t: np.array([int(x) for x in t]))
    return Dfpcmo(header, nfsym, b, pos, ele, bas, energy_line, coef_label, buf, eval_label, evals, sym_label,
                  syminfo, kappa_label, kappa)


def _fortran_e(x: float, w: int = 22, d: int = 12) -> str:
    """Fortran Ew.d: 0.ddddE+xx."""
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}E}".split("E")
        digits = m.replace(".", "")
        ex = int(e) + 1
        s = ("-" if x < 0 else "") + "0." + digits + ("E%+03d" % ex)
    return s.rjust(w)


def write_dfpcmo(path: str, f: Dfpcmo, diracver: int) -> None:
    """fockivo.F90:292-330."""
    new = diracver >= 21
    out = []
    for h in f.header:
        out.append(h.ljust(150))
    if new:
        out.append("".join(" %d" % x for x in [f.nfsym, f.b] + [v for i in range(f.nfsym)
                                                                  for v in (f.pos[i], f.ele[i], f.bas[i])]))
    else:
        out.append("".join(" %d" % x for x in [f.nfsym] + [v for i in range(f.nfsym)
                                                           for v in (f.pos[i], f.ele[i], f.bas[i])]))
    out.append(f.energy_line.ljust(150))
    if new:
        out.append(f.coef_label.ljust(150))
    for k in range(0, len(f.buf), 6):
        out.append("".join("%22.16f" % v for v in f.buf[k:k + 6]))
    if new:
        out.append(f.eval_label.ljust(150))
    for k in range(0, len(f.evals), 6):
        out.append("".join(_fortran_e(v) for v in f.evals[k:k + 6]))
    if new:
        out.append(f.sym_label.ljust(150))
    for k in range(0, len(f.syminfo), 66):
        out.append("".join(" %d" % v for v in f.syminfo[k:k + 66]))
    if f.kappa is not None:
        if new:
            out.append(f.kappa_label.ljust(150))
        for k in range(0, len(f.kappa), 66):
            out.append("".join(" %d" % v for v in f.kappa[k:k + 66]))
    open(path, "w").write("\n".join(out) + "\n")


@dataclass
class IvoProblem:
    inp: IvoInput
    eps: np.ndarray          # caspt2_mo_energy, CAS order (nmo)
    irpamo: np.ndarray       # CAS order (nmo)
    nsymrpa: int
    multb: np.ndarray        # (2 nsymrpa, 2 nsymrpa)
    kset: np.ndarray         # 0-based spinors k of the HOMO set
    f1h: np.ndarray          # (ns, ns, len(kset)): (ab|kk)
    f2h: np.ndarray          # (ns, len(kset), ns): (ak|kb)


def numh_rule(inp: IvoInput, eps: np.ndarray) -> int:
    """fockivo.F90:40-44: the HOMO degeneracy (spinors within 0.1 Eh of the HOMO energy)."""
    if inp.nhomo:
        return inp.nhomo
    act_end = inp.ninact + inp.nact
    return int(np.count_nonzero(np.abs(eps[:act_end] - eps[inp.nelec + inp.ninact - 1]) < 0.1))


def load_ivo(dirpath: str) -> IvoProblem:
    """An IVO test directory of the reference (active.ivo.inp, MRCONEE, MDCINT*)."""
    inp = read_ivo_input(os.path.join(dirpath, "active.ivo.inp"))
    mr = read_mrconee_file(os.path.join(dirpath, "MRCONEE"))

    class _Inp:   # cas_order only reads the RAS lists
        ras1_list = ras2_list = ras3_list = []
    c2d, d2c, e_sorted = cas_order(mr, _Inp())
    nmom = inp.ninact + inp.nact + inp.nsec
    irpamo = np.array([mr.irpamo_dirac[c2d[i0] - 1] for i0 in range(1, mr.nmo + 1)], dtype=np.int64)
    eri, _ = dense_from_records(mr.nmo, mdcintnew_records(dirpath, d2c, inp.mdcint_scheme), nmom)
    act_end = inp.ninact + inp.nact
    numh = numh_rule(inp, e_sorted)
    kset = np.arange(act_end - numh, act_end)
    sec = np.arange(act_end, nmom)
    f1h = np.stack([eri[np.ix_(sec, sec, [k], [k])][:, :, 0, 0] for k in kset], axis=2)
    f2h = np.stack([eri[np.ix_(sec, [k], [k], sec)][:, 0, 0, :] for k in kset], axis=1)
    return IvoProblem(inp, e_sorted, irpamo, mr.nsymrpa, mr.multb, kset, f1h, f2h)


def numh_of(p: IvoProblem) -> int:
    return len(p.kset)


def fockivo_matrix(p: IvoProblem) -> np.ndarray:
    """fockivo.F90:46-78: F_ij = delta_ij eps_i + sum_k [(ik|kj) - (ij|kk)] / numh over the numh
    highest occupied spinors k (a factor 1/2 on the last two for odd nelec); built on i <= j,
    then F_ji = conj(F_ij) (the diagonal conjugated too)."""
    inp = p.inp
    ni, na, ns = inp.ninact, inp.nact, inp.nsec
    act_end = ni + na
    numh = len(p.kset)
    F = np.zeros((ns, ns), dtype=np.complex128)
    F[np.arange(ns), np.arange(ns)] = p.eps[act_end:act_end + ns]
    for t, k in enumerate(p.kset):
        fac = 0.5 if (k + 1 > act_end - 2 and inp.nelec % 2 == 1) else 1.0
        F += fac * (p.f2h[:, t, :] - p.f1h[:, :, t]) / numh
    iu = np.triu_indices(ns)
    U = np.zeros_like(F)
    U[iu] = F[iu]
    out = U + np.conj(np.triu(U, 1)).T
    out[np.arange(ns), np.arange(ns)] = np.conj(np.diag(U))
    return out


def ivo_irreps(p: IvoProblem, nfsym: int):
    """The irreps fockivo diagonalises: isym = start .. end step 2 per fermion irrep
    (fockivo.F90:160-184), 1 .. nsymrpa/2 for the first and nsymrpa/2+1 .. nsymrpa for the second."""
    half = p.nsymrpa // 2
    out = list(range(1, half + 1, 2))
    if nfsym > 1:
        out += list(range(half + 1, p.nsymrpa + 1, 2))
    return out


def ivo_rotations(p: IvoProblem, nfsym: int = 2):
    """Per irrep isym: the secondary spinors of that irrep (0-based secondary indices, energy
    order) and ZHEEV's eigenpairs of F restricted to them (fockivo.F90:190-230)."""
    from .lapack_zheev import zheev_vu
    F = fockivo_matrix(p)
    inp = p.inp
    act_end = inp.ninact + inp.nact
    irp_sec = p.irpamo[act_end:act_end + inp.nsec]
    out = {}
    for isym in ivo_irreps(p, nfsym):
        mos = np.nonzero(irp_sec == isym)[0]
        if len(mos) == 0:
            continue
        w, V = zheev_vu(F[np.ix_(mos, mos)])
        out[isym] = (mos, w, V)
    return out


def apply_ivo(p: IvoProblem, f: Dfpcmo, rot) -> Dfpcmo:
    """fockivo.F90:128-290: the virtual MO coefficients of every irrep rotated by the eigenvectors."""
    inp = p.inp
    buf = f.buf.copy()
    mo = [f.pos[i] + f.ele[i] for i in range(f.nfsym)]
    basis_all = [mo[i] * f.bas[i] for i in range(f.nfsym)]
    half = p.nsymrpa // 2
    for ir in range(f.nfsym):
        nao = f.bas[ir]
        occ, vcut = inp.occ[ir], inp.vcut[ir]
        num_virt = f.ele[ir] - occ - vcut
        buf0 = (f.pos[ir] + occ) * nao + (basis_all[0] if ir == 1 else 0)
        itrfmo = np.array([[buf[buf0 + imo * nao + iao] for imo in range(num_virt)] for iao in range(nao)],
                          dtype=np.complex128)
        jk = f.pos[ir] + occ + (mo[0] if ir == 1 else 0)      # 0-based index before the first virtual
        lo = jk + 1                                            # 1-based first virtual
        hi = (mo[0] if ir == 1 else 0) + f.pos[ir] + f.ele[ir] - vcut
        isyms = range(1, half + 1, 2) if ir == 0 else range(half + 1, p.nsymrpa + 1, 2)
        for isym in isyms:
            if isym not in rot:
                continue
            mos, w, V = rot[isym]
            want = isym if ir == 0 else isym - half
            seg = f.syminfo[lo - 1:hi]
            if np.all(seg == 0):
                dmo = list(range(lo, hi + 1))
            else:
                dmo = [lo + t for t in range(hi - lo + 1) if abs(seg[t]) == want]
            coeff = np.zeros((nao, len(mos)), dtype=np.complex128)
            for i, d in enumerate(dmo):
                coeff[:, i] = itrfmo[:, d - jk - 1]
            coeff = coeff @ V
            for i, d in enumerate(dmo):
                itrfmo[:, d - jk - 1] = coeff[:, i]
        for iao in range(nao):
            for imo in range(num_virt):
                buf[buf0 + imo * nao + iao] = itrfmo[iao, imo].real
    out = Dfpcmo(**{**f.__dict__})
    out.buf = buf
    return out
