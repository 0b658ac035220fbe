import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FIXTURE_DIR = os.path.join(ROOT, "tests", "golden", "fixtures")
FIXTURES = ["c2_h2o_dev", "mdcint_scheme_6_dirac23_c32h_n2_dev", "ninact_0_h2_dev", "check_minelecras3_dev"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def fixture_cache():
    from oracle.dirac_files import load_fixture_npz
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = load_fixture_npz(os.path.join(FIXTURE_DIR, name + ".npz"))
        return cache[name]
    return get
