import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "drmlt-dpt_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def cfg1():
    from drmlt import scenes
    return scenes.config1()


@pytest.fixture(scope="session")
def cfg2_small():
    from drmlt import scenes
    # the config-2 generator at reduced tessellation for CPU-speed tests
    return scenes.config2(sphere_level=2, grid_res=64)
