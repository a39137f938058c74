"""Shared test setup: import paths, the `gpu` marker, cached corpora."""
import os
import sys

import pytest

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
