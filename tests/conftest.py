import os
import sys

import numpy as np
import pytest

# torch first: its HIP runtime (same SONAME as the system's) is then the one
# the library binds to in every test process, whatever order tests import in.
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole GPU session (the HIP path, never a fallback)."""
    from ambient_image_processor_amd import native
    c = native.Context(0)
    yield c
    c.close()


@pytest.fixture(autouse=True)
def _clear_plan_overrides(request):
    """Launch-plan overrides a test sets on the shared context end with it."""
    yield
    if "ctx" in request.fixturenames:
        request.getfixturevalue("ctx").clear_overrides()


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
