"""IVO fixtures (tests/golden/ivo/, written by tests/golden/make_ivo_fixtures.py from the
reference's IVO test directories) and the reference test's file comparison."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
IVO_DIR = os.path.join(HERE, "golden", "ivo")
IVO_NAMES = ["ivo_c32h_n2_dev_dirac22", "ivo_c32h_n2_dev_dirac19", "ivo_c32h_ar_with_kappa_dev"]


def load(name):
    from oracle import ivo_ref as I
    z = np.load(os.path.join(IVO_DIR, name + ".npz"), allow_pickle=False)
    ninact, nact, nsec, nelec, diracver, nhomo = (int(x) for x in z["dims"])
    inp = I.IvoInput(ninact=ninact, nact=nact, nsec=nsec, nelec=nelec, diracver=diracver, nhomo=nhomo,
                     occ=[int(x) for x in z["occ"]], vcut=[int(x) for x in z["vcut"]])
    return I.IvoProblem(inp, z["eps"], z["irpamo"], int(z["nsymrpa"]), z["multb"], z["kset"], z["f1h"], z["f2h"])


def paths(name):
    return os.path.join(IVO_DIR, name + ".DFPCMO"), os.path.join(IVO_DIR, name + ".DFPCMONEW")


def assert_same_file(got_path, ref_path, tol=1e-13):
    """The comparison of the reference's IVO tests (test_ivo_c2_h2o_dev.py:35-46): line by line,
    numbers within tol, other tokens equal."""
    with open(got_path) as a, open(ref_path) as b:
        la, lb = a.read().splitlines(), b.read().splitlines()
    assert len(la) == len(lb), (len(la), len(lb))
    worst = 0.0
    for k, (x, y) in enumerate(zip(la, lb)):
        tx, ty = x.split(), y.split()
        try:
            fx, fy = [float(t) for t in tx], [float(t) for t in ty]
        except ValueError:
            assert tx == ty, (k, x, y)
            continue
        assert len(fx) == len(fy), k
        for u, v in zip(fx, fy):
            worst = max(worst, abs(u - v))
            assert abs(u - v) <= tol, (k, u, v)
    return worst
