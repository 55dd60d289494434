import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def gen_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("gen"))


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible GPU; fails (never skips) without one."""
    import nabc_amd as na
    n = na.device_count()
    assert n >= 1, "no GPU visible to the HIP runtime"
    return na
