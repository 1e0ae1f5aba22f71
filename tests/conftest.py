import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as ge
    return ge.load_pkg()


@pytest.fixture(scope="session")
def orc():
    import __graft_entry__ as ge
    o = ge.load_oracle()
    o.lib()
    return o


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def dummy(pkg):
    """The reference's own fixture files (dummy-data.bin / dummy-queries.bin)."""
    dg = pkg.datagen
    vecs, labels, ts = dg.read_base_bin(os.path.join(GOLDEN, "dummy-data.bin"))
    q, metas = dg.read_query_bin(os.path.join(GOLDEN, "dummy-queries.bin"))
    return vecs, labels, ts, q, metas
