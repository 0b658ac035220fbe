"""The reference's files around the CASCI/CASPT2 path, read and written on the host.

* ``CIDATA_sym<totsym>`` (casci.F90:117-160, read_cidata.F90:1-154): Fortran sequential
  unformatted records; every item is a 30-character key record (``cidata_key_size``,
  module_global_variables.F90:186) followed by its value record(s): ninact, nact, nsec,
  nelec, ndet, nroot, totsym (default INTEGER: 4 bytes, 8 in the -DINT64 build; the record
  length tells which), ecas(1:nroot) (real(8), without ecore), dict_cas_idx_values
  (int64 determinant bit strings), ci_coefficients (real(8) cir(ndet, nroot), then cii in
  the complex build) and the key ``end``.  Record markers are 4-byte lengths before and
  after each record (gfortran's default).
* ``caspt2_restart_<totsym>_<selectroot>`` (module_restart_file.F90:55-172): text lines
  ``totsym: N``, ``selectroot: N``, ``Next subspace: X`` and one line per finished subspace
  ``Subspace: X Sumc2: v Energy: v``.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

CIDATA_KEY_SIZE = 30
SUBSPACES = "ABCDEFGH"


# ---------------------------------------------------------------- Fortran unformatted records
def _write_record(f, payload: bytes) -> None:
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _read_records(path: str) -> List[bytes]:
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {pos}")
        (n,) = struct.unpack_from("<i", data, pos)
        end = pos + 4 + n
        if n < 0 or end + 4 > len(data):
            raise ValueError(f"{path}: record of {n} bytes at byte {pos} runs past the end of the file")
        (n2,) = struct.unpack_from("<i", data, end)
        if n2 != n:
            raise ValueError(f"{path}: record markers {n} and {n2} differ at byte {pos}")
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def _key(name: str) -> bytes:
    if len(name) > CIDATA_KEY_SIZE:
        raise ValueError(f"CIDATA key longer than {CIDATA_KEY_SIZE} characters: {name}")
    return name.encode("ascii").ljust(CIDATA_KEY_SIZE, b" ")


def _int(payload: bytes) -> int:
    if len(payload) == 4:
        return struct.unpack("<i", payload)[0]
    if len(payload) == 8:
        return struct.unpack("<q", payload)[0]
    raise ValueError(f"integer record of {len(payload)} bytes")


@dataclass
class Cidata:
    ninact: int
    nact: int
    nsec: int
    nelec: int
    totsym: int
    ecas: np.ndarray                   # (nroot,) CASCI energies without ecore
    dets: np.ndarray                   # (ndet,) int64 bit strings, the CASCI determinant order
    ci: np.ndarray                     # (ndet, nroot) complex128

    @property
    def ndet(self) -> int:
        return len(self.dets)

    @property
    def nroot(self) -> int:
        return len(self.ecas)


def write_cidata(path: str, cd: Cidata, int_bytes: int = 4, complex_build: bool = True) -> None:
    """Write CIDATA as casci.F90:117-160 does (int_bytes 8: the -DINT64 build; complex_build
    False: the real-only build, which writes cir only)."""
    fmt = {4: "<i", 8: "<q"}[int_bytes]
    with open(path, "wb") as f:
        for name, v in (("ninact", cd.ninact), ("nact", cd.nact), ("nsec", cd.nsec), ("nelec", cd.nelec),
                        ("ndet", cd.ndet), ("nroot", cd.nroot), ("totsym", cd.totsym)):
            _write_record(f, _key(name))
            _write_record(f, struct.pack(fmt, int(v)))
        _write_record(f, _key("ecas"))
        _write_record(f, np.ascontiguousarray(cd.ecas, dtype="<f8").tobytes())
        _write_record(f, _key("dict_cas_idx_values"))
        _write_record(f, np.ascontiguousarray(cd.dets, dtype="<i8").tobytes())
        _write_record(f, _key("ci_coefficients"))
        ci = np.asarray(cd.ci, dtype=np.complex128)
        _write_record(f, np.asfortranarray(ci.real).astype("<f8").tobytes(order="F"))
        if complex_build:
            _write_record(f, np.asfortranarray(ci.imag).astype("<f8").tobytes(order="F"))
        _write_record(f, _key("end"))


def read_cidata(path: str, complex_build: bool = True) -> Cidata:
    """Read CIDATA with the checks of read_cidata.F90 (the file must end with the key
    ``end``; ecas needs nroot, the determinants and coefficients need ndet (and nroot);
    unknown keys are an error)."""
    recs = _read_records(path)
    if not recs or recs[-1].decode("ascii").strip() != "end":
        raise ValueError(f"{path}: cidata file is corrupted (it does not end with the key 'end')")
    vals: Dict[str, object] = {}
    i = 0
    while i < len(recs):
        key = recs[i].decode("ascii").strip()
        i += 1
        if key in ("ninact", "nact", "nsec", "nelec", "ndet", "nroot", "totsym"):
            vals[key] = _int(recs[i])
            i += 1
        elif key == "ecas":
            if "nroot" not in vals:
                raise ValueError(f"{path}: ecas detected before nroot")
            vals[key] = np.frombuffer(recs[i], dtype="<f8").copy()
            i += 1
        elif key == "dict_cas_idx_values":
            if "ndet" not in vals:
                raise ValueError(f"{path}: dict_cas_idx_values detected before ndet")
            vals[key] = np.frombuffer(recs[i], dtype="<i8").copy()
            i += 1
        elif key == "ci_coefficients":
            if "ndet" not in vals or "nroot" not in vals:
                raise ValueError(f"{path}: ci_coefficients detected before ndet or nroot")
            shape = (int(vals["ndet"]), int(vals["nroot"]))
            cir = np.frombuffer(recs[i], dtype="<f8").reshape(shape, order="F")
            i += 1
            cii = np.zeros(shape)
            if complex_build:
                cii = np.frombuffer(recs[i], dtype="<f8").reshape(shape, order="F")
                i += 1
            vals[key] = cir + 1j * cii
        elif key == "end":
            break
        else:
            raise ValueError(f"{path}: unknown keyword in cidata file: {key!r}")
    missing = [k for k in ("ninact", "nact", "nsec", "nelec", "ndet", "nroot", "totsym", "ecas",
                           "dict_cas_idx_values", "ci_coefficients") if k not in vals]
    if missing:
        raise ValueError(f"{path}: missing items {missing}")
    if int(vals["ndet"]) < 0 or int(vals["nroot"]) < 0:
        raise ValueError(f"{path}: invalid ndet or nroot")
    return Cidata(int(vals["ninact"]), int(vals["nact"]), int(vals["nsec"]), int(vals["nelec"]),
                  int(vals["totsym"]), np.asarray(vals["ecas"])[:int(vals["nroot"])],
                  np.asarray(vals["dict_cas_idx_values"]), np.asarray(vals["ci_coefficients"]))


# ---------------------------------------------------------------- caspt2_restart
@dataclass
class Restart:
    totsym: int
    selectroot: int
    next_subspace: str                             # 'A'..'H', or 'I': every subspace is done
    e2: Dict[str, float] = field(default_factory=dict)
    sumc2: Dict[str, float] = field(default_factory=dict)


def restart_filename(totsym: int, selectroot: int) -> str:
    return f"caspt2_restart_{totsym}_{selectroot}"


def _after(line: str, key: str) -> str:
    i = line.find(key)
    if i < 0:
        raise ValueError(f"no {key!r} in restart line {line!r}")
    rest = line[i + len(key):].split()
    if not rest:
        raise ValueError(f"nothing after {key!r} in restart line {line!r}")
    return rest[0]


def read_restart(path: str, totsym: Optional[int] = None, selectroot: Optional[int] = None) -> Optional[Restart]:
    """Parse a caspt2_restart file as read_and_validate_restart_file does; None if the file does
    not exist (the calculation starts from scratch).  totsym/selectroot, when given, must match."""
    if not os.path.exists(path):
        return None
    lines = [ln.rstrip("\n") for ln in open(path)]
    if len(lines) < 3:
        raise ValueError(f"{path}: needs the totsym, selectroot and next subspace lines")

    def header(ln, what):
        if ":" not in ln:
            raise ValueError(f"Error reading {what} info in caspt2_restart")
        return ln.split(":", 1)[1].split()[0]
    rs = Restart(int(header(lines[0], "totsym")), int(header(lines[1], "selectroot")),
                 header(lines[2], "next subspace"))
    if totsym is not None and rs.totsym != totsym:
        raise ValueError(f"totsym in {path} does not match with the input file")
    if selectroot is not None and rs.selectroot != selectroot:
        raise ValueError(f"selectroot in {path} does not match with the input file")
    if rs.next_subspace not in "ABCDEFGHI":
        raise ValueError("Invalid subspace (not A-H), no need to restart CASPT2 calculation")
    for ln in lines[3:]:
        if not ln.strip():
            continue
        x = _after(ln, "Subspace:")
        if x not in SUBSPACES:
            raise ValueError("Invalid subspace (not A-H), no need to restart CASPT2 calculation")
        rs.sumc2[x] = float(_after(ln, "Sumc2:"))
        rs.e2[x] = float(_after(ln, "Energy:"))
    return rs


def write_restart(path: str, rs: Restart) -> None:
    """Write the file read_restart (and the reference) reads; subspaces in A..H order."""
    with open(path, "w") as f:
        f.write(f"totsym: {rs.totsym}\nselectroot: {rs.selectroot}\nNext subspace: {rs.next_subspace}\n")
        for x in SUBSPACES:
            if x in rs.e2:
                f.write(f"Subspace: {x} Sumc2: {rs.sumc2.get(x, 0.0)!r} Energy: {rs.e2[x]!r}\n")
