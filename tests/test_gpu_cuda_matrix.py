"""The reference's CUDA-vs-CPU matrix (tests/c/cuda/test_dress_cuda.c:196-237):
triangle, house and weighted star, in every variant, with and without the
precomputed intercepts, fitted for up to 100 iterations at eps 1e-15.  The
reference asks for edge and vertex values within 1e-12 of the CPU library;
the HIP backend is held to bit equality with the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# the fixtures of test_dress_cuda.c:89-133 (graph data only)
GRAPHS = {
    "triangle": (3, [0, 1, 0], [1, 2, 2], None),
    "house": (5, [0, 0, 1, 1, 2, 3], [1, 3, 2, 4, 3, 4], None),
    "weighted-star": (4, [0, 0, 0], [1, 2, 3], [1.0, 2.0, 3.0]),
}
CASES = [("triangle", v, p) for v in range(4) for p in (0, 1)] + \
        [("house", 0, 0), ("house", 0, 1), ("house", 1, 1), ("house", 2, 1)] + \
        [("weighted-star", 0, 0), ("weighted-star", 0, 1), ("weighted-star", 1, 1)]


@pytest.mark.parametrize("name,variant,pre", CASES)
def test_cuda_matrix_bitwise(gpu, ora, name, variant, pre):
    n, U, V, W = GRAPHS[name]
    U = np.array(U, np.int32)
    V = np.array(V, np.int32)
    W = None if W is None else np.array(W)
    with gpu.Graph(n, U, V, W, None, variant, bool(pre)) as g:
        it, dl = g.fit(100, 1e-15)
        o = ora.fit(n, U, V, W, None, variant, pre, 100, 1e-15)
        assert it == o["iterations"]
        assert np.array_equal(g.edge_dress, o["edge_dress"])
        assert np.array_equal(g.vertex_dress, o["vertex_dress"])
        assert np.max(np.abs(g.edge_dress - o["edge_dress"]), initial=0.0) < 1e-12   # the reference's bound
