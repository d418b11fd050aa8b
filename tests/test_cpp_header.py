"""The C++ selector header compiles against libdress++'s dress/dress.hpp
(read from the reference tree; skipped where it is absent) and links against
libdress_hip.so's exports."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/libdress++/include"


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_INC, "dress", "dress.hpp")),
                    reason="libdress++ header not available")
def test_hip_selector_header_compiles(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "hip_selector.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), "-I", REF_INC, src],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
