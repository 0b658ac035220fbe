"""Readers for the reference's input files (TEST INFRASTRUCTURE ONLY).

* active.inp               -- read_input_module.F90:50-240 (keywords used on the path)
* MRCONEE                   -- read_mrconee.F90:226-717 (64-bit-integer DIRAC builds)
* MDCINT, MDCINXXXX<n>      -- create_newmdcint.F90:3-321 (DIRAC unique set ->
                               "MDCINTNEW" records) then the 16-fold Kramers /
                               hermiticity / particle-exchange expansion of
                               module_2integrals.F90:395-434, applied to every
                               record, giving the dense (pq|rs) tensor in CAS order.

The dense tensor is what the reference's subspace routing
(divide_2_elec_integral_into_subspaces.F90) and intra_* unfolding reconstruct
block by block; (ss|ss) integrals are dropped as the reference drops them.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .caspt2_ref import DenseIntegrals, Problem, RasSpec, Symmetry


# ---------------------------------------------------------------------------
# Fortran sequential unformatted records
# ---------------------------------------------------------------------------
def fortran_records(path: str) -> List[bytes]:
    data = open(path, "rb").read()
    out = []
    pos = 0
    n = len(data)
    while pos + 4 <= n:
        (ln,) = struct.unpack_from("<i", data, pos)
        pos += 4
        out.append(data[pos:pos + ln])
        pos += ln
        (ln2,) = struct.unpack_from("<i", data, pos)
        if ln2 != ln:
            raise ValueError(f"{path}: corrupt record marker at {pos}")
        pos += 4
    return out


def write_fortran_record(f, payload: bytes) -> None:
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


# ---------------------------------------------------------------------------
# active.inp
# ---------------------------------------------------------------------------
def _int_list(s: str) -> List[int]:
    out: List[int] = []
    for tok in s.replace(",", " ").split():
        if ".." in tok:
            a, b = tok.split("..")
            out.extend(range(int(a), int(b) + 1))
        else:
            out.append(int(tok))
    return sorted(out)


@dataclass
class ActiveInput:
    ninact: int = 0
    nact: int = 0
    nsec: int = 0
    nelec: int = 0
    totsym: int = 1
    selectroot: int = 1
    eshift: float = 0.0
    diracver: int = 21
    scheme: Optional[int] = None
    ras1_list: List[int] = field(default_factory=list)
    ras2_list: List[int] = field(default_factory=list)
    ras3_list: List[int] = field(default_factory=list)
    ras1_max_hole: int = 0
    ras3_max_elec: int = 0
    min_hole_ras1: int = 0
    min_elec_ras3: int = 0

    @property
    def mdcint_scheme(self) -> int:
        # module_global_variables.F90:44 default: 6 for DIRAC <= 22, 4 otherwise
        if self.scheme is not None:
            return self.scheme
        return 6 if self.diracver <= 22 else 4

    def ras(self) -> Optional[RasSpec]:
        if not (self.ras1_list or self.ras2_list or self.ras3_list):
            return None
        return RasSpec(len(self.ras1_list), len(self.ras2_list), len(self.ras3_list),
                       self.ras1_max_hole, self.ras3_max_elec, self.min_hole_ras1, self.min_elec_ras3)


def read_active_inp(path: str) -> ActiveInput:
    lines = [ln.strip() for ln in open(path).read().splitlines()]
    inp = ActiveInput()
    i = 0
    while i < len(lines):
        key = lines[i].lower()
        nxt = lines[i + 1] if i + 1 < len(lines) else ""
        if key == ".ninact":
            inp.ninact = int(nxt); i += 2
        elif key == ".nact":
            inp.nact = int(nxt); i += 2
        elif key == ".nsec":
            inp.nsec = int(nxt); i += 2
        elif key == ".nelec":
            inp.nelec = int(nxt); i += 2
        elif key == ".caspt2_ciroots":
            v = [int(x) for x in nxt.split()]
            inp.totsym, inp.selectroot = v[0], v[1]
            i += 2
        elif key == ".eshift":
            inp.eshift = float(nxt); i += 2
        elif key == ".diracver":
            inp.diracver = int(nxt); i += 2
        elif key == ".scheme":
            inp.scheme = int(nxt); i += 2
        elif key == ".ras1":
            inp.ras1_list = _int_list(nxt); inp.ras1_max_hole = int(lines[i + 2].split()[0]); i += 3
        elif key == ".ras2":
            inp.ras2_list = _int_list(nxt); i += 2
        elif key == ".ras3":
            inp.ras3_list = _int_list(nxt); inp.ras3_max_elec = int(lines[i + 2].split()[0]); i += 3
        elif key == ".minholeras1":
            inp.min_hole_ras1 = int(nxt); i += 2
        elif key == ".minelecras3":
            inp.min_elec_ras3 = int(nxt); i += 2
        else:
            i += 1
    return inp


# ---------------------------------------------------------------------------
# MRCONEE
# ---------------------------------------------------------------------------
@dataclass
class Mrconee:
    nmo: int
    ecore: float
    nsymrpa: int
    repna: List[str]
    multb: np.ndarray
    irpamo_dirac: np.ndarray
    energy_dirac: np.ndarray
    h_dirac: np.ndarray              # (nmo, nmo) complex, DIRAC order, Im sign as stored
    hf_energy: float


def read_mrconee_file(path: str) -> Mrconee:
    recs = fortran_records(path)
    r1 = recs[0]
    if len(r1) < 64:
        raise NotImplementedError("only 64-bit-integer DIRAC MRCONEE files are supported")
    nmo, breit, ecore, nfsym, nz, spinfr, norbt, hf = struct.unpack_from("<qqdqqqqd", r1, 0)
    (nsymrp,) = struct.unpack_from("<q", recs[1], 0)
    (nsymrpa,) = struct.unpack_from("<q", recs[2], 0)
    repna = [recs[2][8 + 4 * k: 12 + 4 * k].decode("ascii", "replace") for k in range(2 * nsymrpa)]
    n2 = 2 * nsymrpa
    mb = np.frombuffer(recs[3], dtype="<i8", count=n2 * n2).reshape(n2, n2).T.copy()  # multb(i0,j0), i0 fastest
    rec5 = np.frombuffer(recs[4], dtype=np.uint8, count=24 * nmo)
    irpmo = np.zeros(nmo, np.int64); irpamo = np.zeros(nmo, np.int64); en = np.zeros(nmo)
    for k in range(nmo):
        irpmo[k], irpamo[k], en[k] = struct.unpack_from("<qqd", rec5.tobytes(), 24 * k)
    one = np.frombuffer(recs[5], dtype="<f8", count=2 * nmo * nmo).reshape(nmo, nmo, 2)  # (i0, j0, re/im), j0 fastest
    h = one[:, :, 0] + 1j * one[:, :, 1]
    return Mrconee(int(nmo), float(ecore), int(nsymrpa), repna, mb, irpamo, en, h, float(hf))


def cas_order(mr: Mrconee, inp: ActiveInput) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Energy sort (+ RAS reorder) and the Kramers-pair index maps (read_mrconee.F90:500-541).

    Returns (indmo_cas_to_dirac, indmo_dirac_to_cas, caspt2_mo_energy), 1-based maps.
    """
    nmo = mr.nmo
    e_sorted = np.sort(mr.energy_dirac)
    if inp.ras1_list or inp.ras2_list or inp.ras3_list:
        e_sorted = _ras_order(e_sorted, inp)
    c2d = np.zeros(nmo + 1, np.int64)
    d2c = np.zeros(nmo + 1, np.int64)
    for i0 in range(1, nmo + 1, 2):
        m = 0
        for j0 in range(1, nmo + 1):
            if mr.energy_dirac[j0 - 1] == e_sorted[i0 - 1]:
                if m == 0:
                    c2d[i0] = j0; d2c[j0] = i0; m += 1
                else:
                    c2d[i0 + 1] = j0; d2c[j0] = i0 + 1
    return c2d, d2c, e_sorted


def _ras_order(e: np.ndarray, inp: ActiveInput) -> np.ndarray:
    """sort_list_from_energy_order_to_ras_order (read_mrconee.F90:602-715)."""
    ras_all = set(inp.ras1_list) | set(inp.ras2_list) | set(inp.ras3_list)
    out = []
    order = [k for k in range(1, len(e) + 1) if k not in ras_all]
    ni = inp.ninact
    out.extend(e[k - 1] for k in order[:ni])
    for lst in (inp.ras1_list, inp.ras2_list, inp.ras3_list):
        out.extend(e[k - 1] for k in lst)
    out.extend(e[k - 1] for k in order[ni:])
    return np.array(out[:len(e)])


# ---------------------------------------------------------------------------
# MDCINT family -> dense integrals
# ---------------------------------------------------------------------------
def mdcint_files(dirpath: str) -> List[str]:
    """MDCINT, MDCINXXXX1, MDCINXXXX2, ... (get_filename.F90:1-41, single process)."""
    out = [os.path.join(dirpath, "MDCINT")]
    k = 1
    while True:
        pad = "XXXX" if k < 10 else "XXX" if k < 100 else "XX" if k < 1000 else "X"
        f = os.path.join(dirpath, f"MDCIN{pad}{k}")
        if not os.path.exists(f):
            break
        out.append(f)
        k += 1
    return out


def _mdcint_is_complex(rec: bytes) -> bool:
    """module_realonly.F90:29-77: a record long enough for (re,im) pairs is complex."""
    (nz,) = struct.unpack_from("<q", rec, 16)
    return len(rec) >= 24 + 16 * nz + 16 * nz


def _type2_or_3(i, j, k, l) -> bool:
    return (int(i < 0) + int(j < 0)) == 1 and (int(k < 0) + int(l < 0)) == 1


@dataclass
class MdcintSet:
    """DIRAC's unique integral set as create_newmdcint.F90 reads it: per record (ikr, jkr, nz)
    and its nz entries (indk, indl, rklr[, rkli]) concatenated over the records of all
    MDCINT files in file order; kr[nkr + m] = DIRAC spinor of Kramers label m (the header's
    kr(-nkr:nkr)); ``rkli`` is None for real (realonly) integrals."""
    ikr: np.ndarray
    jkr: np.ndarray
    nz: np.ndarray
    indk: np.ndarray
    indl: np.ndarray
    rklr: np.ndarray
    rkli: Optional[np.ndarray]
    nkr: int
    kr: np.ndarray


def read_mdcint_set(dirpath: str, max_records: Optional[int] = None) -> MdcintSet:
    """The records of MDCINT, MDCINXXXX1, ... (read_mdcint_header and read_one_line_of_mdcint,
    create_newmdcint.F90:209-280; a record with ikr = 0 ends a file), at most max_records."""
    files = mdcint_files(dirpath)
    hdr = fortran_records(files[0])[0]
    (nkr,) = struct.unpack_from("<q", hdr, 18)
    vals = struct.unpack_from("<" + "q" * (2 * nkr), hdr, 26)
    kr = np.zeros(2 * nkr + 1, np.int64)
    for i0 in range(1, nkr + 1):
        kr[nkr + i0] = vals[2 * (i0 - 1)]
        kr[nkr - i0] = vals[2 * (i0 - 1) + 1]
    ikr, jkr, nzs, ks, ls, rr, ri = [], [], [], [], [], [], []
    is_complex = None
    for f in files:
        recs = fortran_records(f)
        for rec in recs[1:]:
            a, b, nz = struct.unpack_from("<qqq", rec, 0)
            if a == 0:
                break
            if max_records is not None and len(ikr) >= max_records:
                break
            if is_complex is None:
                is_complex = _mdcint_is_complex(rec)
            idx = np.frombuffer(rec, dtype="<i8", count=2 * nz, offset=24).reshape(nz, 2)
            ikr.append(a); jkr.append(b); nzs.append(nz)
            ks.append(idx[:, 0].copy()); ls.append(idx[:, 1].copy())
            if is_complex:
                v = np.frombuffer(rec, dtype="<f8", count=2 * nz, offset=24 + 16 * nz).reshape(nz, 2)
                rr.append(v[:, 0].copy()); ri.append(v[:, 1].copy())
            else:
                rr.append(np.frombuffer(rec, dtype="<f8", count=nz, offset=24 + 16 * nz).copy())
    cat = (lambda x, dt: np.concatenate(x).astype(dt) if x else np.zeros(0, dt))
    return MdcintSet(np.array(ikr, np.int64), np.array(jkr, np.int64), np.array(nzs, np.int64),
                     cat(ks, np.int64), cat(ls, np.int64), cat(rr, np.float64),
                     cat(ri, np.float64) if is_complex else None, int(nkr), kr)


def mdcintnew_from_set(ms: MdcintSet, d2c: np.ndarray, scheme: int):
    """Yield MDCINTNEW-equivalent records (i, j, k, l, re, im) (create_newmdcint.F90:86-177) in
    the order the reference writes them: record by record, the record as read (n = 1) then
    its sign-flipped Kramers image (n = 2), each over its nz entries and, when the SCHEME
    amplifies (scheme != 6 and |ikr| != |jkr|), over the nz particle-exchanged entries after
    them.  Indices are 1-based CAS-order spinors (d2c[k] for DIRAC spinor k); ``im`` is the
    value written to MDCINTNEW (the negated DIRAC imaginary part)."""
    is_complex = ms.rkli is not None
    cutoff = 0.25e-12
    off = 0
    for r in range(len(ms.ikr)):
        ikr, jkr, nz = int(ms.ikr[r]), int(ms.jkr[r]), int(ms.nz[r])
        indk = ms.indk[off:off + nz].copy(); indl = ms.indl[off:off + nz].copy()
        rr = ms.rklr[off:off + nz].copy()
        ri = ms.rkli[off:off + nz].copy() if is_complex else np.zeros(nz)
        off += nz
        for n in (1, 2):
            if n == 2:
                ikr, jkr = -ikr, -jkr
                indk, indl = -indk, -indl
                sg = np.sign(ikr * jkr * indk * indl)
                sg[sg == 0] = 1
                rr = -rr * sg
                ri = ri * sg
            amplify = (scheme != 6 and abs(ikr) != abs(jkr))
            nloop = 2 * nz if amplify else nz
            for ii in range(1, nloop + 1):
                inz = nz if ii % nz == 0 else ii % nz
                ci, cj, ck, cl = ikr, jkr, int(indk[inz - 1]), int(indl[inz - 1])
                re, im = float(rr[inz - 1]), float(ri[inz - 1])
                if ii > nz:
                    ci, cj = cj, ci
                    ck, cl = cl, ck
                    if _type2_or_3(ci, cj, ck, cl):
                        ci, cj, ck, cl = -ci, -cj, -ck, -cl
                    else:
                        im = -im
                iii, jjj, kkk, lll = (int(d2c[ms.kr[ms.nkr + x]]) for x in (ci, cj, ck, cl))
                iikr, jjkr, kkkr, llkr = ((1 if x % 2 == 1 else -1) * (x // 2 + x % 2) for x in (iii, jjj, kkk, lll))
                iiit, jjjt, kkkt, lllt = (x + 1 if x % 2 == 1 else x - 1 for x in (iii, jjj, kkk, lll))
                a, b, c, d = abs(iikr), abs(jjkr), abs(kkkr), abs(llkr)
                if is_complex:
                    if abs(re) <= cutoff and abs(im) <= cutoff:
                        continue
                elif abs(re) <= cutoff:
                    continue
                ok = False
                if iikr > 0 and jjkr > 0 and kkkr > 0 and llkr > 0:
                    ok = ((a <= b and c <= d and (a < c or (a == c and b <= d))) or
                          (a <= b and d <= c and (a < d or (a == d and b <= c))))
                elif iikr > 0 and jjkr < 0 and kkkr > 0 and llkr < 0:
                    ok = a <= b and c <= d and (a < c or (a == c and b <= d))
                elif iikr > 0 and jjkr < 0 and kkkr < 0 and llkr > 0:
                    ok = a <= b and c <= d and (a < c or (a == c and b <= d))
                elif iikr > 0 and jjkr < 0 and kkkr < 0 and llkr < 0:
                    ok = a <= b
                if ok:
                    yield iiit, jjjt, kkkt, lllt, re, -im


def mdcintnew_records(dirpath: str, d2c: np.ndarray, scheme: int):
    """Yield MDCINTNEW-equivalent records (i, j, k, l, re, im) of the MDCINT files in dirpath
    (create_newmdcint.F90:86-177): read_mdcint_set then mdcintnew_from_set."""
    yield from mdcintnew_from_set(read_mdcint_set(dirpath), d2c, scheme)


def dense_from_records(nmo: int, records, nmom: int) -> Tuple[np.ndarray, int]:
    """Fill the dense tensor with the 16 symmetry images of every record
    (module_2integrals.F90:395-434).  Indices above nmom (virtuals beyond the
    secondary space) are dropped.  Returns (eri, number_of_conflicting_writes).
    """
    eri = np.zeros((nmom, nmom, nmom, nmom), dtype=np.complex128)
    seen = np.zeros((nmom, nmom, nmom, nmom), dtype=bool)
    conflicts = 0
    tr = lambda x: x + 1 if x % 2 == 1 else x - 1
    for (i, j, k, l, re, im) in records:
        if max(i, j, k, l) > nmom:
            continue
        v = complex(re, im)
        sij = 1 if (i + j) % 2 == 0 else -1
        skl = 1 if (k + l) % 2 == 0 else -1
        it, jt, kt, lt = tr(i), tr(j), tr(k), tr(l)
        imgs = [
            ((i, j, k, l), v), ((jt, it, k, l), v * sij), ((i, j, lt, kt), v * skl), ((jt, it, lt, kt), v * sij * skl),
            ((j, i, l, k), v.conjugate()), ((it, jt, l, k), v.conjugate() * sij),
            ((j, i, kt, lt), v.conjugate() * skl), ((it, jt, kt, lt), v.conjugate() * sij * skl),
            ((k, l, i, j), v), ((lt, kt, i, j), v * skl), ((k, l, jt, it), v * sij), ((lt, kt, jt, it), v * sij * skl),
            ((l, k, j, i), v.conjugate()), ((kt, lt, j, i), v.conjugate() * skl),
            ((l, k, it, jt), v.conjugate() * sij), ((kt, lt, it, jt), v.conjugate() * sij * skl),
        ]
        for (a, b, c, d), x in imgs:
            a, b, c, d = a - 1, b - 1, c - 1, d - 1
            if seen[a, b, c, d]:
                if abs(eri[a, b, c, d] - x) > 1e-10 * max(1.0, abs(x)):
                    conflicts += 1
            eri[a, b, c, d] = x
            seen[a, b, c, d] = True
    return eri, conflicts


@dataclass
class Fixture:
    inp: ActiveInput
    mr: Optional[Mrconee]
    problem: Problem
    conflicts: int
    gold: Optional[dict] = None


def load_fixture_npz(path: str) -> Fixture:
    """Problem + goldens from a tests/golden/fixtures/*.npz file (see make_fixtures.py)."""
    z = np.load(path, allow_pickle=False)
    ninact, nact, nsec, nelec, totsym, selectroot = (int(x) for x in z["dims"])
    nmom = ninact + nact + nsec
    idx = z["idx"].astype(np.int64)
    val = z["val"].astype(np.complex128)
    recs = ((int(a), int(b), int(c), int(d), float(v.real), float(v.imag))
            for (a, b, c, d), v in zip(idx, val))
    eri, conflicts = dense_from_records(nmom, recs, nmom)
    rv = [int(x) for x in z["ras"]]
    ras = RasSpec(*rv) if any(rv[:3]) else None
    sym = Symmetry(int(z["nsymrpa"]), z["multb"], z["irpamo"])
    p = Problem(ninact=ninact, nact=nact, nsec=nsec, nelec=nelec, h=z["h"], eri=DenseIntegrals(eri),
                ecore=float(z["ecore"]), selectroot=selectroot, totsym=totsym, eshift=float(z["eshift"]),
                sym=sym, ras=ras)
    inp = ActiveInput(ninact=ninact, nact=nact, nsec=nsec, nelec=nelec, totsym=totsym,
                      selectroot=selectroot, eshift=float(z["eshift"]))
    gold = {k[5:]: z[k] for k in z.files if k.startswith("gold_")}
    return Fixture(inp, None, p, conflicts, gold)


def load_fixture(dirpath: str) -> Fixture:
    """Build an oracle Problem from a reference test directory (active.inp, MRCONEE, MDCINT*)."""
    inp = read_active_inp(os.path.join(dirpath, "active.inp"))
    mr = read_mrconee_file(os.path.join(dirpath, "MRCONEE"))
    c2d, d2c, e_sorted = cas_order(mr, inp)
    nmo = mr.nmo
    nmom = inp.ninact + inp.nact + inp.nsec
    # one-electron integrals in CAS order, Im sign flip for DIRAC >= 21 (read_mrconee.F90:579-597)
    hd = mr.h_dirac.copy()
    if inp.diracver >= 21:
        hd = hd.real - 1j * hd.imag
    h = np.zeros((nmo, nmo), dtype=np.complex128)
    for i0 in range(1, nmo + 1):
        for j0 in range(1, nmo + 1):
            h[d2c[i0] - 1, d2c[j0] - 1] = hd[i0 - 1, j0 - 1]
    irpamo = np.array([mr.irpamo_dirac[c2d[i0] - 1] for i0 in range(1, nmo + 1)], dtype=np.int64)
    recs = mdcintnew_records(dirpath, d2c, inp.mdcint_scheme)
    eri, conflicts = dense_from_records(nmo, recs, nmom)
    sym = Symmetry(mr.nsymrpa, mr.multb, irpamo[:nmom])
    p = Problem(ninact=inp.ninact, nact=inp.nact, nsec=inp.nsec, nelec=inp.nelec,
                h=h[:nmom, :nmom], eri=DenseIntegrals(eri), ecore=mr.ecore,
                selectroot=inp.selectroot, totsym=inp.totsym, eshift=inp.eshift,
                sym=sym, ras=inp.ras())
    return Fixture(inp, mr, p, conflicts)
