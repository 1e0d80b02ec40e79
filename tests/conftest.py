import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_collection_modifyitems(config, items):
    # GPU tests print the case they are about to run so a fault names it
    pass


@pytest.fixture(autouse=True)
def _collect_after_gpu_test(request):
    """Device handles caught in reference cycles are destroyed by the garbage collector; collect
    after every GPU test so that this happens right after the test that made them, where a
    failure would name the test, and not later inside pytest's own bookkeeping."""
    yield
    if request.node.get_closest_marker("gpu") is not None:
        import gc
        gc.collect()


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle
