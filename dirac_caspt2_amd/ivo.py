"""IVO (improved virtual orbitals) on the host side of the library: the subprogram
r4divo.F90:1-114 with fockivo.F90:4-409.

The IVO Fock matrix of the secondary spinors and its per-irrep eigenvectors come from the
library (dc2_fockivo: the matrix built on the device from the resident F1/F2 integral blocks,
the small per-irrep eigenproblems solved by ZHEEV's sequence).  This module reads DIRAC's MO
coefficient file DFPCMO, rotates the virtual MOs of every irrep by those eigenvectors
(fockivo.F90:128-290: a basis_ao x nv by nv x nv product per irrep) and writes DFPCMONEW in the
reference's layout (fockivo.F90:292-330).

DFPCMO (formatted; DIRAC >= 21 adds the label lines INFO/COEFS/EVALS/SUPERSYM/KAPPA):
  [INFO]  title  |  nfsym [B] (npos, nele, nbas) x nfsym  |  energy line  |  [COEFS]
  coefficients (6F22.16)  |  [EVALS] eigenvalues (6E22.12)  |  [SUPERSYM] irreps (66(X,I0))
  |  optional [KAPPA] kappa (66(X,I0)).
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import List, Optional, Sequence

import numpy as np


@dataclass
class Dfpcmo:
    header: List[str]
    nfsym: int
    b: Optional[int]
    npos: List[int]
    nele: List[int]
    nbas: List[int]
    energy_line: str
    labels: dict
    coef: np.ndarray
    evals: np.ndarray
    supersym: np.ndarray
    kappa: Optional[np.ndarray] = None


def read_dfpcmo(path: str, diracver: int) -> Dfpcmo:
    """fockivo.F90:113-178."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    labeled = diracver >= 21
    it = iter(lines)
    header = [next(it)] if labeled else []
    header.append(next(it))
    dims = [int(t) for t in next(it).split()]
    nfsym = dims[0]
    b = dims[1] if labeled else None
    trip = dims[2:] if labeled else dims[1:]
    npos, nele, nbas = trip[0::3][:nfsym], trip[1::3][:nfsym], trip[2::3][:nfsym]
    energy = next(it)
    labels = {}

    def block(count, key, kind):
        if labeled:
            labels[key] = next(it)
        vals = []
        while len(vals) < count:
            vals += next(it).split()
        return np.array([kind(v.replace("D", "E")) for v in vals[:count]])

    nmo = sum(npos) + sum(nele)
    nao = sum((p + e) * a for p, e, a in zip(npos, nele, nbas))
    coef = block(nao, "coef", float)
    evals = block(nmo, "evals", float)
    supersym = block(nmo, "supersym", int)
    rest = [ln for ln in it]
    kappa = None
    if rest and rest[0].strip():
        it = iter(rest)
        kappa = block(nmo, "kappa", int)
    return Dfpcmo(header, nfsym, b, npos, nele, nbas, energy, labels, coef, evals, supersym, kappa)


def _e22_12(x: float) -> str:
    """Fortran E22.12 (mantissa 0.d..d)."""
    if x == 0.0:
        return "0.000000000000E+00".rjust(22)
    mant, ex = f"{abs(x):.11E}".split("E")
    return (("-" if x < 0 else "") + "0." + mant.replace(".", "") + "E%+03d" % (int(ex) + 1)).rjust(22)


def write_dfpcmo(path: str, f: Dfpcmo, diracver: int) -> None:
    """fockivo.F90:292-330."""
    labeled = diracver >= 21
    rows = [h.ljust(150) for h in f.header]
    trip = [v for i in range(f.nfsym) for v in (f.npos[i], f.nele[i], f.nbas[i])]
    rows.append("".join(" %d" % v for v in ([f.nfsym, f.b] if labeled else [f.nfsym]) + trip))
    rows.append(f.energy_line.ljust(150))

    def block(key, vals, per, fmt):
        if labeled:
            rows.append(f.labels[key].ljust(150))
        for k in range(0, len(vals), per):
            rows.append("".join(fmt(v) for v in vals[k:k + per]))

    block("coef", f.coef, 6, lambda v: "%22.16f" % v)
    block("evals", f.evals, 6, _e22_12)
    block("supersym", f.supersym, 66, lambda v: " %d" % v)
    if f.kappa is not None:
        block("kappa", f.kappa, 66, lambda v: " %d" % v)
    with open(path, "w") as fh:
        fh.write("\n".join(rows) + "\n")


def ivo_irreps(nsymrpa: int, nfsym: int) -> List[int]:
    """The irreps fockivo diagonalises (fockivo.F90:160-184): odd isym of each fermion irrep's half."""
    half = nsymrpa // 2
    return list(range(1, half + 1, 2)) + (list(range(half + 1, nsymrpa + 1, 2)) if nfsym > 1 else [])


def rotate_virtuals(f: Dfpcmo, U: np.ndarray, irpamo_sec: Sequence[int], nsymrpa: int, occ: Sequence[int],
                    vcut: Sequence[int]) -> Dfpcmo:
    """fockivo.F90:128-290: in every fermion irrep, the coefficients of the virtual MOs of each
    irrep isym (those of DFPCMO's supersym value |s| = isym or isym - nsymrpa/2, all when that
    range is unlabelled) multiplied by the eigenvectors of the secondary spinors of irpamo isym
    (U's block); the real part is kept, as the reference's DBLE does."""
    coef = f.coef.copy()
    half = nsymrpa // 2
    nmo = [f.npos[i] + f.nele[i] for i in range(f.nfsym)]
    irp = np.asarray(irpamo_sec)
    for ir in range(f.nfsym):
        nao = f.nbas[ir]
        nvirt = f.nele[ir] - occ[ir] - vcut[ir]
        mo0 = (nmo[0] if ir else 0) + f.npos[ir] + occ[ir]          # MOs before the first virtual
        c0 = (nmo[0] * f.nbas[0] if ir else 0) + (f.npos[ir] + occ[ir]) * nao
        C = coef[c0:c0 + nao * nvirt].reshape(nvirt, nao).T.astype(np.complex128)   # (AO, virtual)
        seg = f.supersym[mo0:mo0 + nvirt]
        for isym in (range(1, half + 1, 2) if ir == 0 else range(half + 1, nsymrpa + 1, 2)):
            mos = np.nonzero(irp == isym)[0]
            if len(mos) == 0:
                continue
            want = isym if ir == 0 else isym - half
            cols = np.arange(nvirt) if np.all(seg == 0) else np.nonzero(np.abs(seg) == want)[0]
            block = np.zeros((nao, len(mos)), dtype=np.complex128)
            block[:, :len(cols)] = C[:, cols]
            C[:, cols] = (block @ U[np.ix_(mos, mos)])[:, :len(cols)]
        coef[c0:c0 + nao * nvirt] = C.real.T.ravel()
    return replace(f, coef=coef)


def fockivo(ctx, mo_energy: np.ndarray, dfpcmo_in: str, dfpcmo_out: str, diracver: int, occ: Sequence[int],
            vcut: Sequence[int], nsymrpa: int, irpamo: Sequence[int], nhomo: int = 0) -> dict:
    """The IVO step of r4divo.F90:60-80 for a context holding the spaces, symmetry and F1/F2
    integral blocks: DFPCMO in, DFPCMONEW out.  Returns the Fock matrix, eigenvalues and U."""
    f = read_dfpcmo(dfpcmo_in, diracver)
    fock, w, U = ctx.fockivo(nhomo, mo_energy, ivo_irreps(nsymrpa, f.nfsym))
    nmoc = ctx.dims[0] + ctx.dims[1]
    g = rotate_virtuals(f, U, np.asarray(irpamo)[nmoc:nmoc + ctx.dims[2]], nsymrpa, occ, vcut)
    write_dfpcmo(dfpcmo_out, g, diracver)
    return {"fock": fock, "w": w, "U": U}
