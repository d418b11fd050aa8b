"""dress_get (virtual-edge queries) on the device against the golden answers
of the compiled reference and the oracle."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_get_golden(gpu):
    z = np.load(os.path.join(GOLDEN, "get_queries.npz"))
    with gpu.Graph(80, z["U"], z["V"]) as g:
        g.fit(100, 1e-6)
        batch = g.get_batch(z["qu"], z["qv"], 100, 1e-6, z["qw"])
        single = np.array([g.get(int(a), int(b), 100, 1e-6, float(w))
                           for a, b, w in zip(z["qu"][:20], z["qv"][:20], z["qw"][:20])])
    assert np.array_equal(batch, z["ans"])
    assert np.array_equal(single, z["ans"][:20])


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_get_variants_vs_oracle(gpu, ora, variant):
    rng = np.random.default_rng(70 + variant)
    n, m = 50, 200
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    W = rng.random(m) + 0.2
    NW = rng.random(n) + 0.5
    qu = rng.integers(0, n, 100).astype(np.int32)
    qv = rng.integers(0, n, 100).astype(np.int32)
    with gpu.Graph(n, U, V, W, NW, variant) as g:
        g.fit(100, 1e-7)
        got = g.get_batch(qu, qv, 50, 1e-7, 0.7)
    og = ora.graph(n, U, V, W, NW, variant, 0)
    ora.fit_graph(og, 100, 1e-7)
    want = np.array([ora.get(og, int(a), int(b), 50, 1e-7, 0.7) for a, b in zip(qu, qv)])
    ora.free_graph(og)
    assert np.array_equal(got, want)


def test_get_out_of_range_is_an_error(gpu):
    with gpu.Graph(3, [0, 1], [1, 2]) as g:
        g.fit()
        with pytest.raises(gpu.HipError):
            g.get(0, 7)
