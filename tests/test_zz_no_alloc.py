"""dress_hip_get_batch_resident_noalloc (no device allocation: queries are
staged in a fixed device array of the library) through the C ABI: its
answers equal dress_hip_get_batch_resident's bit for bit, on two calls in a
row, the second with fewer queries than the first, and a call beyond its
maximum is refused with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_QUERIES = 65536


def _call(dh, f, g, u, v, w):
    from dress_hip._lib import c_int_p, c_dbl_p
    out = np.full(u.size, -1.0)
    r = f(g.g, int(u.size), u.ctypes.data_as(c_int_p), v.ctypes.data_as(c_int_p), w.ctypes.data_as(c_dbl_p),
          100, 1e-6, out.ctypes.data_as(c_dbl_p))
    return r, out


def test_noalloc_get_batch_equals_allocating_entry_point(gpu, gen):
    L = gpu._lib.lib()
    n, U, V, W = gen.gnp(3000, 12.0, 61, weighted=True)
    rng = np.random.default_rng(5)
    with gpu.Graph(n, U, V, W) as g:
        g.fit(100, 1e-6)
        u0 = np.ascontiguousarray(rng.integers(0, n, 50_000).astype(np.int32))
        v0 = np.ascontiguousarray(np.concatenate([rng.integers(0, n, 40_000), V[:10_000]]).astype(np.int32))
        u0[40_000:] = U[:10_000]                      # existing edges too
        w0 = np.ascontiguousarray(rng.random(u0.size) + 0.5)
        # the allocating entry point first: it places the query rows
        r, ref0 = _call(gpu, L.dress_hip_get_batch_resident, g, u0, v0, w0)
        assert r == 0, gpu._lib.last_error()
        r, got0 = _call(gpu, L.dress_hip_get_batch_resident_noalloc, g, u0, v0, w0)
        assert r == 0, gpu._lib.last_error()
        assert np.array_equal(got0, ref0)
        # a second, smaller call: nothing of the first one's staging shows
        u1, v1, w1 = u0[7:1207].copy(), v0[7:1207].copy(), (w0[7:1207] * 2).copy()
        r, ref1 = _call(gpu, L.dress_hip_get_batch_resident, g, u1, v1, w1)
        assert r == 0
        r, got1 = _call(gpu, L.dress_hip_get_batch_resident_noalloc, g, u1, v1, w1)
        assert r == 0, gpu._lib.last_error()
        assert np.array_equal(got1, ref1)
        assert not np.array_equal(got1, got0[7:1207])
        # beyond the maximum: refused before any device work
        big = np.zeros(MAX_QUERIES + 1, np.int32)
        r, _ = _call(gpu, L.dress_hip_get_batch_resident_noalloc, g, big, big, np.ones(big.size))
        assert r == -1 and "maximum" in gpu._lib.last_error()
