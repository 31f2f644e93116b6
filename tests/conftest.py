import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """One libamj context for the GPU tests of the session."""
    from accmultiwayjoins_amd import Context

    c = Context(0)
    name, hbm, cus = c.device_info()
    print(f"\n[gpu] {name}, {hbm / 2**30:.0f} GiB HBM, {cus} CUs", flush=True)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    import oracle_ffi

    return oracle_ffi
