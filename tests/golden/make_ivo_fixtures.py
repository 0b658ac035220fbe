"""Write the IVO fixtures of tests/golden/ivo/ from the reference's three IVO test directories
that hold their integrals (run in a container where the reference tree exists):

    python tests/golden/make_ivo_fixtures.py /path/to/reference/test/dev

For each directory: <name>.npz with the spaces and keywords of active.ivo.inp, the orbital
energies and irreps in CAS order, the symmetry tables, the HOMO spinor set and the integral
slices the IVO Fock matrix reads, (ab|kk) and (ak|kb) for secondary a, b and HOMO spinors k
(from MRCONEE and the MDCINT family through the oracle's readers); <name>.DFPCMO and
<name>.DFPCMONEW are the directory's DFPCMO and reference.DFPCMONEW data files, copied as
they are (the input coefficient file and the reference's expected output).
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ivo_ref as I  # noqa: E402

NAMES = ["ivo_c32h_n2_dev_dirac22", "ivo_c32h_n2_dev_dirac19", "ivo_c32h_ar_with_kappa_dev"]


def main(root):
    out = os.path.join(HERE, "ivo")
    os.makedirs(out, exist_ok=True)
    for name in NAMES:
        d = os.path.join(root, name)
        p = I.load_ivo(d)
        inp = p.inp
        np.savez_compressed(os.path.join(out, name + ".npz"),
                            dims=np.array([inp.ninact, inp.nact, inp.nsec, inp.nelec, inp.diracver, inp.nhomo]),
                            occ=np.array(inp.occ), vcut=np.array(inp.vcut), eps=p.eps, irpamo=p.irpamo,
                            nsymrpa=np.array(p.nsymrpa), multb=p.multb, kset=p.kset, f1h=p.f1h, f2h=p.f2h)
        shutil.copyfile(os.path.join(d, "DFPCMO"), os.path.join(out, name + ".DFPCMO"))
        shutil.copyfile(os.path.join(d, "reference.DFPCMONEW"), os.path.join(out, name + ".DFPCMONEW"))
        print(name, "numh", len(p.kset), "nsec", inp.nsec)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/dev")
