"""Test configuration: the `gpu` marker, and the build-if-absent rule for the
in-tree libraries (tests load what is built; they build only when nothing
is)."""
from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    need = [os.path.join(ROOT, "oracle", "liboracle.so"),
            os.path.join(ROOT, "cudec_amd", "libcudec.so"),
            os.path.join(ROOT, "cudec_amd", "libcudec_workload.so")]
    if not all(os.path.exists(p) for p in need):
        import __graft_entry__
        __graft_entry__.build()


@pytest.fixture(scope="session")
def lz4_fixtures():
    from tests import corpus
    return corpus.make_lz4_fixtures()


@pytest.fixture(scope="session")
def snappy_fixtures():
    from tests import corpus
    return corpus.make_snappy_fixtures()
