"""Shared test set-up: import paths, the `gpu` marker, seeded workloads and the oracle."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "luvas.io-3d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def small_workload():
    from lr3d import scenes
    return scenes.sphere_small().render()


@pytest.fixture(scope="session")
def ctx():
    import lr3d
    c = lr3d.Context(0)
    yield c
    c.destroy()
