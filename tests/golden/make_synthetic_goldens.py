"""Oracle energies of the seeded synthetic workloads too large to recompute inside a test.

    python tests/golden/make_synthetic_goldens.py [split]

Writes tests/golden/synthetic_<name>.json: the oracle's per-subspace e2/sumc2, e0, CASCI
energy and total energy for the workload regenerated from its seed by
dirac_caspt2_amd.synth (inputs are not stored: the generator is deterministic).
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dirac_caspt2_amd import synth  # noqa: E402
from oracle import caspt2_ref as R  # noqa: E402

WORKLOADS = {"split": synth.CONFIG_SPLIT}


def main(names):
    for name in names:
        spec = WORKLOADS[name]
        sp = synth.generate(spec)
        p = R.Problem(spec.ninact, spec.nact, spec.nsec, spec.nelec, sp.h, R.FactorIntegrals(sp.B), sp.ecore,
                      spec.selectroot, eshift=spec.eshift)
        t0 = time.time()
        res = R.run_caspt2(p)
        out = {"spec": [spec.ninact, spec.nact, spec.nsec, spec.nelec, spec.seed, spec.selectroot, spec.eshift],
               "casci": float(res.casci_energies[spec.selectroot - 1]), "e0": res.e0,
               "e2": res.e2, "sumc2": res.sumc2, "total": res.total,
               "dimm": {X: sum(b.dimm for _, b in res.e2_blocks[X]) for X in "ABCDEFGH"},
               "oracle_seconds": time.time() - t0}
        with open(os.path.join(HERE, f"synthetic_{name}.json"), "w") as f:
            json.dump(out, f, indent=1)
        print(name, out)


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
