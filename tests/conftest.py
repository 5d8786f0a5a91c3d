"""Shared test set-up: import paths, the `gpu` marker, library loaders."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "qes-public_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libqesmc.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()  # builds oracle/build/liboracle_qesmc.so if it is missing
    return O


@pytest.fixture(scope="session")
def qesmc():
    import qesmc as Q
    Q.lib()  # raises if libqesmc.so was not built
    return Q


@pytest.fixture(scope="session")
def gpu_ctx(qesmc):
    ctx = qesmc.Context(0)  # raises (no fallback) when no device is present
    yield ctx
    ctx.close()
