"""Store the reference's MDCINT integral sets as small npz inputs for the MDCINT -> MDCINTNEW
conversion tests (this container only; the GPU tests read the npz files).

For each fixture directory test/dev/<name> of the reference this writes
tests/golden/mdcint/<name>.npz holding the records exactly as create_newmdcint.F90 reads
them (oracle.dirac_files.read_mdcint_set): ikr, jkr, nz per record, the concatenated entries
indk, indl, rklr (all four fixtures hold real integrals), nkr and kr(-nkr:nkr) of the MDCINT
header, plus indmo_dirac_to_cas (read_mrconee.F90:500-541, oracle.dirac_files.cas_order),
the SCHEME of active.inp and the size of the spinor set (ninact + nact + nsec).  The two
small fixtures are stored whole; of the two large ones (12 and 18 MB of MDCINT files) the
first MAX_RECORDS records of MDCINT are kept, which reproduce the first records of the full
conversion.

Usage:  python tests/golden/make_mdcint_fixtures.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.dirac_files import cas_order, read_active_inp, read_mdcint_set, read_mrconee_file  # noqa: E402

WHOLE = ["ninact_0_h2_dev", "check_minelecras3_dev"]
PREFIX = ["c2_h2o_dev", "mdcint_scheme_6_dirac23_c32h_n2_dev"]
MAX_RECORDS = 40


def convert(ref_root: str, name: str, outdir: str, max_records=None) -> None:
    d = os.path.join(ref_root, "test", "dev", name)
    inp = read_active_inp(os.path.join(d, "active.inp"))
    mr = read_mrconee_file(os.path.join(d, "MRCONEE"))
    _, d2c, _ = cas_order(mr, inp)
    ms = read_mdcint_set(d, max_records)
    assert ms.rkli is None, "complex MDCINT sets are not stored by this script"
    assert np.abs(ms.indk).max() < 2 ** 15 and np.abs(ms.indl).max() < 2 ** 15
    np.savez_compressed(os.path.join(outdir, name + ".npz"), ikr=ms.ikr, jkr=ms.jkr, nz=ms.nz,
                        indk=ms.indk.astype(np.int16), indl=ms.indl.astype(np.int16), rklr=ms.rklr,
                        nkr=np.int64(ms.nkr), kr=ms.kr, d2c=d2c, scheme=np.int64(inp.mdcint_scheme),
                        nmom=np.int64(inp.ninact + inp.nact + inp.nsec), whole=np.bool_(max_records is None))
    print(name, len(ms.ikr), "records,", len(ms.indk), "entries")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    outdir = os.path.join(HERE, "mdcint")
    os.makedirs(outdir, exist_ok=True)
    for n in WHOLE:
        convert(ref, n, outdir)
    for n in PREFIX:
        convert(ref, n, outdir, MAX_RECORDS)


if __name__ == "__main__":
    main()
