"""GPU parity of the asynchronous traversal (numa-BFS-async-pipe.C:140-175,
edgeMapSparseAsyncPipe polymer.h:1512-1751) through pg_bfs_async.  The parents
are whichever claimer won each CAS, so they are checked as the reference's
result is defined: the reached set is the oracle's BFS-reachable set, the source
is its own parent, every other reached vertex's parent is an in-neighbour, and
following parents from any reached vertex ends at the source (a tree)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check_tree(g, src, parent, visited, levels):
    n = g.n
    reached = parent >= 0
    assert np.array_equal(reached, levels >= 0)
    assert visited == int(reached.sum())
    assert parent[src] == src
    v = np.nonzero(reached)[0]
    v = v[v != src]
    if v.size:
        src_of = np.repeat(np.arange(n, dtype=np.int64), np.diff(g.off))
        keys = np.sort(src_of * n + g.adj.astype(np.int64))
        q = parent[v].astype(np.int64) * n + v
        pos = np.searchsorted(keys, q)
        assert np.all(pos < keys.size) and np.all(keys[np.minimum(pos, keys.size - 1)] == q)
        assert np.all(reached[parent[v]])
    anc = np.where(reached, parent, src).astype(np.int64)
    for _ in range(int(np.ceil(np.log2(max(n, 2)))) + 1):
        anc = anc[anc]
    assert np.all(anc == src)


@pytest.mark.parametrize("sym", [False, True])
def test_async_rmat(pa, orc, sym):
    g = orc.rmat(14, 1 << 18, seed=61)
    if sym:
        g = g.symmetrized()
    gg = pa.Graph.from_csr(g.n, g.off, g.adj)
    deg = np.diff(g.off)
    srcs = [0, 77, int(np.nonzero(deg == 0)[0][0]) if np.any(deg == 0) else 5, 4000]
    for src in srcs:
        say("async: sym %d source %d" % (sym, src))
        parent, vis, st = gg.bfs_async(src)
        check_tree(g, src, parent, vis, orc.bfs_levels(g, src))
        say("  reached %d in %.3f ms" % (vis, st.ms))
    gg.close()


def test_async_long_chain_and_star(pa, orc):
    """A 3000-vertex path (every vertex is queued by the one before it) and a
    star whose hub the whole wave expands."""
    n = 3000
    g = orc.from_edges(n, list(range(n - 1)), list(range(1, n)))
    gg = pa.Graph.from_csr(g.n, g.off, g.adj)
    parent, vis, _ = gg.bfs_async(0)
    assert vis == n and np.array_equal(parent[1:], np.arange(n - 1))
    check_tree(g, 0, parent, vis, orc.bfs_levels(g, 0))
    parent, vis, _ = gg.bfs_async(n - 1)  # a sink: only itself
    assert vis == 1 and parent[n - 1] == n - 1 and np.all(parent[:-1] == -1)
    gg.close()
    m = 20000
    g = orc.from_edges(m + 1, [0] * m + list(range(1, m + 1)), list(range(1, m + 1)) + [0] * m)
    gg = pa.Graph.from_csr(g.n, g.off, g.adj)
    for src in (0, 7):
        parent, vis, _ = gg.bfs_async(src)
        check_tree(g, src, parent, vis, orc.bfs_levels(g, src))
    gg.close()
