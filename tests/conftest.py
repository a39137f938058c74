"""Shared test setup: import paths, the `gpu` marker, cached corpora."""
import os
import sys

import pytest

# torch (used by the full-size configuration tests for device memory and
# random generation) brings its own copy of the HIP runtime, loaded under a
# file name of its own: imported after libhipibs.so it would be a second HIP
# runtime in the process, and one of the two then finds no device. Imported
# first, it is the process's one runtime and libhipibs.so binds to it.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "gpu-ibs_amd"), os.path.join(ROOT, "tests"),
          os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def small_corpus():
    import corpus
    return corpus.small_config()


@pytest.fixture(scope="session")
def expected():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "expected.json")) as f:
        return json.load(f)
