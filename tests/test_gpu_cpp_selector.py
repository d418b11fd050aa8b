"""The C++ selector (include/dress/hip/dress.hpp): a program written
against libdress++'s dress::DRESS, switched to the GPU by include, gives
the oracle's fit and Delta/Nabla histograms on a triangle.  The program is
built by tests/cpp/Makefile (against libdress++'s header) when the
reference tree is present; without the built program the test is skipped."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "cpp", "_bin", "hip_selector")


@pytest.mark.skipif(not os.path.isfile(BIN), reason="C++ selector check not built (needs libdress++'s header)")
def test_cpp_selector_matches_oracle(gpu, ora):
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    it, d0, nd, nsub, nn, nsub1 = out.stdout.split()
    U = np.array([0, 1, 2], np.int32)
    V = np.array([1, 2, 0], np.int32)
    o = ora.fit(3, U, V, max_iterations=100, epsilon=1e-6)
    assert int(it) == o["iterations"]
    assert float(d0) == o["edge_dress"][0] == 2.0
    dv, dc, _, dn = ora.lifted("delta", 3, U, V, k=1)
    nv, nc, _, nnum = ora.lifted("nabla", 3, U, V, k=1)
    assert (int(nd), int(nsub), int(nsub1)) == (len(dv), dn, dn)
    assert int(nn) == len(nv)
