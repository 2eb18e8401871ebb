"""pytest configuration: import paths and the `gpu` marker.

Tests marked `gpu` need a real MI355X (they call the HIP library through the
C ABI); everything else runs on the CPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "lpm-v2_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP library through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    import lpmgpu

    c = lpmgpu.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    return o
