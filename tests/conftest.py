import importlib
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def wr():
    """The package (HIP library binding)."""
    return importlib.import_module("warp-reactor_amd")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def oracle_wr():
    import oracle_lib
    return oracle_lib.OracleWrMath()


@pytest.fixture(scope="session")
def ref():
    import oracle_lib
    r = oracle_lib.Reference.load()
    if r is None:
        pytest.skip("reference CPU build (oracle/_ref) not built here")
    return r


@pytest.fixture(scope="session")
def ctx(wr):
    c = wr.Context(1 << 16)
    yield c
    c.close()
