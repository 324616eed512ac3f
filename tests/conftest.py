import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("box2d-rs_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import b2o
    return b2o.Oracle("f32")


@pytest.fixture(scope="session")
def oracle64():
    import b2o
    return b2o.Oracle("f64")


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_makereport(item, call):
    """A device error (status 3: a fault, a lost device) leaves the GPU unusable for the rest of
    the process: end the session at the first one, naming the test, instead of running the
    remaining GPU tests against a dead device."""
    outcome = yield
    rep = outcome.get_result()
    if rep.failed and call.excinfo is not None and getattr(call.excinfo.value, "status", None) == 3:
        pytest.exit("device error in %s: %s" % (item.nodeid, call.excinfo.value), returncode=3)
