"""Shared fixtures.  Tests marked ``gpu`` need an MI355X; the rest run on CPU.

The CPU oracle (oracle/liboracle.so) and, where it was built, the compiled
reference (oracle/_ref/libdress_ref.so) are the checkers; the code under test
is libdress_hip.so, reached only through its C ABI.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "dress-graph_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (libdress_hip.so compute calls)")


@pytest.fixture(scope="session")
def ora():
    import pyoracle
    return pyoracle.Lib("oracle")


@pytest.fixture(scope="session")
def ref():
    import pyoracle
    if not pyoracle.available("ref"):
        pytest.skip("compiled reference (oracle/_ref) not built")
    return pyoracle.Lib("ref")


@pytest.fixture(scope="session")
def dh():
    import dress_hip
    dress_hip._lib.lib()
    return dress_hip


@pytest.fixture(scope="session")
def gpu(dh):
    if not dh.is_available():
        raise RuntimeError("no GPU visible to libdress_hip.so: " + dh._lib.last_error())
    return dh


@pytest.fixture(scope="session")
def gen():
    from dress_hip import generators
    return generators
