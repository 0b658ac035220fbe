"""Convert the reference's runnable CASPT2 test fixtures into small npz files.

Reads (this container only) the reference test directories
    test/dev/{c2_h2o_dev, mdcint_scheme_6_dirac23_c32h_n2_dev,
              ninact_0_h2_dev, check_minelecras3_dev}
and writes tests/golden/fixtures/<name>.npz holding

* the integrals as MDCINTNEW-equivalent records (1-based CAS-order spinor
  indices + values) produced by oracle.dirac_files.mdcintnew_records,
* the one-electron integrals in CAS order, ecore, symmetry tables, active.inp
  parameters,
* the expected numbers printed in the reference output
  (reference.<name>.out): CASCI energies, e2X and sumc2,X per subspace,
  e2X(isym) per block, c^2, the total second-order and total energies.

Usage:  python tests/golden/make_fixtures.py [/root/reference]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.dirac_files import (cas_order, mdcintnew_records, read_active_inp,  # noqa: E402
                                read_mrconee_file)

FIXTURES = ["c2_h2o_dev", "mdcint_scheme_6_dirac23_c32h_n2_dev", "ninact_0_h2_dev", "check_minelecras3_dev"]
FLOAT = r"([-+]?\d+\.\d+(?:[EeDd][-+]?\d+)?)"


def parse_reference_output(text: str) -> dict:
    out = {"casci": [float(x) for x in re.findall(r"CASCI Total Energy ROOT\s*\d+\s+" + FLOAT, text)]}
    for X in "abcdefgh":
        m = re.search(r"\be2%s\s+=\s+%s" % (X, FLOAT), text)
        out["e2" + X.upper()] = float(m.group(1)) if m else 0.0
        m = re.search(r"sumc2,%s\s+=\s+%s" % (X, FLOAT), text)
        out["sumc2" + X.upper()] = float(m.group(1)) if m else 0.0
        out["e2blocks" + X.upper()] = [float(v) for v in
                                      re.findall(r"e2%s\(\s*\d+\)\s+=\s+%s" % (X, FLOAT), text)]
    out["c2"] = float(re.search(r"c\^2 is\s+" + FLOAT, text).group(1))
    out["total_second_order"] = float(re.search(r"Total second order energy is\s+" + FLOAT, text).group(1))
    out["total"] = float(re.search(r"Total energy is\s+" + FLOAT, text).group(1))
    return out


def convert(ref_root: str, name: str, outdir: str) -> None:
    d = os.path.join(ref_root, "test", "dev", name)
    inp = read_active_inp(os.path.join(d, "active.inp"))
    mr = read_mrconee_file(os.path.join(d, "MRCONEE"))
    c2d, d2c, _ = cas_order(mr, inp)
    nmom = inp.ninact + inp.nact + inp.nsec
    recs = [r for r in mdcintnew_records(d, d2c, inp.mdcint_scheme) if max(r[:4]) <= nmom]
    idx = np.array([r[:4] for r in recs], dtype=np.uint8)
    val = np.array([complex(r[4], r[5]) for r in recs], dtype=np.complex128)
    hd = mr.h_dirac.copy()
    if inp.diracver >= 21:
        hd = hd.real - 1j * hd.imag
    nmo = mr.nmo
    h = np.zeros((nmo, nmo), dtype=np.complex128)
    for i0 in range(1, nmo + 1):
        for j0 in range(1, nmo + 1):
            h[d2c[i0] - 1, d2c[j0] - 1] = hd[i0 - 1, j0 - 1]
    irpamo = np.array([mr.irpamo_dirac[c2d[i0] - 1] for i0 in range(1, nmo + 1)], dtype=np.int64)
    ras = inp.ras()
    rasv = np.array([0] * 7 if ras is None else
                    [ras.ras1_size, ras.ras2_size, ras.ras3_size, ras.ras1_max_hole,
                     ras.ras3_max_elec, ras.min_hole_ras1, ras.min_elec_ras3], dtype=np.int64)
    gold = parse_reference_output(open(os.path.join(d, "reference.%s.out" % name)).read())
    np.savez_compressed(
        os.path.join(outdir, name + ".npz"),
        idx=idx, val=val if np.any(val.imag) else val.real,
        h=h[:nmom, :nmom], ecore=mr.ecore, nsymrpa=mr.nsymrpa, multb=mr.multb, irpamo=irpamo[:nmom],
        dims=np.array([inp.ninact, inp.nact, inp.nsec, inp.nelec, inp.totsym, inp.selectroot]),
        eshift=inp.eshift, ras=rasv,
        gold_casci=np.array(gold["casci"]),
        gold_e2=np.array([gold["e2" + X] for X in "ABCDEFGH"]),
        gold_sumc2=np.array([gold["sumc2" + X] for X in "ABCDEFGH"]),
        gold_c2=gold["c2"], gold_total2=gold["total_second_order"], gold_total=gold["total"],
        **{"gold_blocks_" + X: np.array(gold["e2blocks" + X]) for X in "ABCDEFGH"},
    )
    print(name, len(recs), "records")


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.join(HERE, "fixtures")
    os.makedirs(out, exist_ok=True)
    for n in FIXTURES:
        convert(ref, n, out)
