import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_models_dump")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref_models():
    import json
    with open(os.path.join(GOLDEN, "ref_models.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ref_offsets():
    import json
    with open(os.path.join(GOLDEN, "ref_offsets.json")) as f:
        return json.load(f)
