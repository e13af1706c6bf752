import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running (full-size workloads)")


@pytest.fixture(scope="session")
def oracle():
    from tests import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def ref():
    from tests import oracle_lib
    r = oracle_lib.Reference.load()
    if r is None:
        pytest.skip("oracle/_ref (compiled reference headers) not built")
    return r


@pytest.fixture(scope="session")
def gpu():
    import mpir_amd
    if not mpir_amd.device_ok():
        pytest.fail("no usable gfx950 device for a test marked gpu")
    return mpir_amd
