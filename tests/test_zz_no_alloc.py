"""pg_subset_to_sparse_noalloc: the allocation-free toSparse (its device memory is a
static array per device) against pg_subset_to_sparse, bit-exact, on two calls in a
row (the second with a smaller graph), through the C ABI.  Device buffers for the
outputs come from the HIP runtime the library itself loaded (hipMalloc through
ctypes), so pointers and streams belong to one runtime."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _noalloc(pa, hip, sub, n):
    """Runs the entry point into fresh device buffers (filled with a marker first)
    and returns the ids it wrote."""
    L = pa.load()
    ids = ctypes.c_void_p()
    cnt = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ids), 4 * (n + 16)) == 0
    assert hip.hipMalloc(ctypes.byref(cnt), 8) == 0
    try:
        marker = np.full(n + 16, -7, np.int32)
        assert hip.hipMemcpy(ids, marker.ctypes.data, 4 * (n + 16), 1) == 0  # host to device
        rc = L.pg_subset_to_sparse_noalloc(sub._h, ids, cnt, None)
        assert rc == 0, pa.last_error()
        assert hip.hipDeviceSynchronize() == 0
        k = np.zeros(1, np.int64)
        assert hip.hipMemcpy(k.ctypes.data, cnt, 8, 2) == 0  # device to host
        out = np.zeros(n + 16, np.int32)
        assert hip.hipMemcpy(out.ctypes.data, ids, 4 * (n + 16), 2) == 0
        # nothing written past the count
        assert np.all(out[int(k[0]):] == -7)
        return out[: int(k[0])]
    finally:
        hip.hipFree(ids)
        hip.hipFree(cnt)


def test_noalloc_matches_to_sparse(pa):
    hip = _hip()
    rng = np.random.default_rng(17)
    cases = ((16, 1 << 20, 0.3), (10, 1 << 13, 0.05))  # the second call has a smaller graph
    for scale, m, frac in cases:
        say("case: noalloc toSparse, scale %d, %.0f%% of the vertices" % (scale, 100 * frac))
        g = pa.Graph.rmat(scale, m, seed=scale)
        s = g.subset()
        pick = np.nonzero(rng.random(g.n) < frac)[0].astype(np.int32)
        s.from_sparse(pick)
        want = s.to_sparse()
        got = _noalloc(pa, hip, s, g.n)
        assert np.array_equal(np.sort(pick), want)
        assert np.array_equal(got, want), (scale, got.size, want.size)
        # the static array keeps the last call's contents: an empty subset after a
        # full one must come out empty
        s.clear()
        assert _noalloc(pa, hip, s, g.n).size == 0
        s.fill()
        assert np.array_equal(_noalloc(pa, hip, s, g.n), np.arange(g.n, dtype=np.int32))
        s.close()
        g.close()


def test_noalloc_refuses_large_graphs(pa):
    say("case: noalloc toSparse above its maximum size")
    g = pa.Graph.rmat(26, 1 << 10, seed=1)  # 2^26 vertices > 2^25
    s = g.subset()
    L = pa.load()
    rc = L.pg_subset_to_sparse_noalloc(s._h, ctypes.c_void_p(16), ctypes.c_void_p(16), None)
    assert rc == -5  # PG_ERR_UNSUPPORTED, refused before any launch
    assert "exceeds" in pa.last_error()
    s.close()
    g.close()
