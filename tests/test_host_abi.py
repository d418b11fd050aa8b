"""Host-side checks of libdress_hip.so (no GPU needed).

* the library loads and exports every function include/dress/*.h and
  include/dress/hip/*.h declare;
* the lifted drivers' subgraph enumeration (lexicographic combinations and
  ordered tuples, stride/offset, LCG sampling) equals itertools' order and
  the oracle's sample picks;
* without a GPU the compute entry points fail loudly (an error, no values).
"""
import ctypes
import itertools
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", "dress", "dress.h"),
           os.path.join(ROOT, "include", "dress", "hip", "dress_hip.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith("dress_") and not name.endswith("_t"):
                names.add(name)
    return sorted(names)


def test_headers_declare_the_reference_api():
    names = set(declared_functions())
    for f in ("dress_init_graph", "dress_fit", "dress_free_graph", "dress_get", "dress_delta_fit",
              "dress_delta_fit_flat", "dress_delta_fit_strided", "dress_delta_fit_strided_flat",
              "dress_nabla_fit", "dress_nabla_fit_flat", "dress_nabla_fit_strided_flat",
              "dress_fit_hip", "dress_delta_fit_hip", "dress_delta_fit_hip_strided",
              "dress_delta_fit_hip_strided_flat", "dress_nabla_fit_hip", "dress_nabla_fit_hip_strided",
              "dress_nabla_fit_hip_flat", "dress_nabla_fit_hip_strided_flat"):
        assert f in names, f


def test_library_exports_every_declared_symbol(dh):
    lib = ctypes.CDLL(dh._lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_only_the_c_interface(dh):
    """Internal C++ symbols (namespace dh, std template instances) stay
    local, so bindings that load several DRESS backends cannot clash."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", dh._lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    names = [ln.split()[-1] for ln in out.splitlines() if ln.strip()]
    declared = set(declared_functions())
    extra = [n for n in names if n not in declared]
    assert not extra, extra[:20]
    assert not [n for n in names if n.startswith("_Z")]


def test_redirect_header_maps_to_hip_symbols():
    text = open(os.path.join(ROOT, "include", "dress", "hip", "dress.h")).read()
    pairs = dict(re.findall(r"#define\s+(\w+)\s+(\w+)", text))
    assert pairs["dress_fit"] == "dress_fit_hip"
    assert pairs["dress_delta_fit"] == "dress_delta_fit_hip"
    assert pairs["dress_nabla_fit"] == "dress_nabla_fit_hip"


@pytest.mark.parametrize("kind", ["delta", "nabla"])
@pytest.mark.parametrize("n,k", [(6, 1), (6, 2), (7, 3), (5, 4)])
def test_enumeration_order(dh, kind, n, k):
    gen = itertools.combinations if kind == "delta" else itertools.permutations
    allt = list(gen(range(n), k))
    for off, st in ((0, 1), (1, 3), (2, 4)):
        sel, verts = dh.enumerate_subgraphs(kind, n, k, 0, 0, off, st)
        want = [(i, t) for i, t in enumerate(allt) if i % st == off]
        assert sel.tolist() == [i for i, _ in want]
        assert [tuple(v) for v in verts.tolist()] == [t for _, t in want]


def test_sampled_enumeration_matches_oracle_picks(dh, ora):
    n, k = 12, 2
    total = math.comb(n, k)
    for seed in (0, 3, 99):
        for ns in (1, 5, 40):
            sel, _ = dh.enumerate_subgraphs("delta", n, k, ns, seed, 0, 1)
            if ns >= total:
                assert sel.tolist() == list(range(total))
            else:
                assert sel.tolist() == ora.sample_picks(total, ns, seed).tolist()


def test_counts_match_reference_helpers(dh, ref):
    L = dh._lib.lib()
    for n in range(0, 9):
        for k in range(-1, 10):
            assert L.dress_delta_binom(n, k) == ref.lib.dress_delta_binom(n, k)
            assert L.dress_nabla_perm_count(n, k) == ref.lib.dress_nabla_perm_count(n, k)


@pytest.mark.parametrize("bad", [(3, 0), (0, -1), (2147483647, 1)])
def test_vertex_ids_outside_the_graph_are_rejected(dh, bad):
    """An edge list with an id outside [0, N) is refused by a host scan
    before anything is written with it (on a GPU box too: no device work is
    started, so this test is safe there)."""
    n = 3
    U = np.array([0, 1, bad[0]], np.int32)
    V = np.array([1, 2, bad[1]], np.int32)
    with pytest.raises(dh.HipError, match=r"outside \[0, 3\)"):
        dh.Graph(n, U, V)
    ok = (4, np.array([0, 1, 2], np.int32), np.array([1, 2, 3], np.int32))
    with pytest.raises(dh.HipError, match=r"graph 1.*outside \[0, 3\)"):
        dh.lifted_batch([ok, (n, U, V)], k=1)
    with pytest.raises(dh.HipError, match=r"graph 1.*outside \[0, 3\)"):
        dh.fit_batch([ok, (n, U, V)])


def test_no_gpu_fails_loudly(dh):
    if dh.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(dh.HipError):
        dh.Graph(3, [0, 1, 0], [1, 2, 2])
    assert dh._lib.last_error() != ""
