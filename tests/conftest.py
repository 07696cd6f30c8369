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


@pytest.fixture(scope="session")
def cfg4_small():
    from drmlt import scenes
    # the config-4 generator at reduced size, every Phong exponent in
    # [600, 1000] (roughness below the 0.03 threshold: relative parameterisation)
    return scenes.config4(sphere_level=1, grid_res=48, instances=48, phong_n=(600.0, 1000.0))


@pytest.fixture(scope="session")
def cfg5_small():
    from drmlt import scenes
    # the config-5 generator at reduced size, every dielectric alpha in
    # [0.02, 0.03) (below the threshold: relative parameterisation)
    return scenes.config5(sphere_level=2, grid_res=32, spheres=48, alpha=(0.02, 0.0299))
