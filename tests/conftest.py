"""Shared test set-up: import paths, the `gpu` marker and the solver context."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpgpu-binary-automaton-synchronizing-algorithms_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libssw.so")


@pytest.fixture(scope="session")
def ctx():
    import ssw
    c = ssw.Context(0)
    yield c
    c.close()
