"""The benchmark's full sizes: PageRank and SpMV on the scale-24 rMat (2^28
edges) against the oracle's fp64 iterations for a few steps; Bellman-Ford on the
symmetrised scale-22 graph against the oracle's Dijkstra, bit-exact; and, through
what does not depend on size, BFS on the symmetrised scale-24 graph by the
properties that define breadth-first levels and parents, CC labels on scale 22 by
the properties that define minimum-id labels, and the 64-source BFS on scale 22
against 64 single-source searches (two independent GPU paths); BP on scale 22
against the oracle's fp64 products where they stay in range."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def test_pagerank_scale24_vs_fp64_pull(pa, orc):
    say("case: pagerank scale 24, 3 iterations against the fp64 oracle pull")
    g = pa.Graph.rmat(24, 1 << 28, seed=1)
    r, _ = g.pagerank(3, 0.85)
    off, adj, _ = g.download()
    hg = orc.Graph(g.n, off, adj)
    hg.transpose()
    ref = orc.pagerank_pull_omp(hg, 3, 0.85, threads=16)
    e = rel_l1(r, ref)
    say("  rel L1 %.3e" % e)
    assert e < 1e-6
    assert np.array_equal(r == 0, ref == 0)
    g.close()


def _check_bfs(off, adj, src, level, parent):
    n = off.size - 1
    deg = np.diff(off)
    assert level[src] == 0 and parent[src] == src
    src_of = np.repeat(np.arange(n, dtype=np.int32), deg)
    lu, lv = level[src_of], level[adj]
    del src_of
    # closure: an edge out of a reached vertex ends at a reached vertex one level
    # deeper at most (and, the graph being symmetric, at most one level shallower)
    reached_u = lu >= 0
    assert np.all(lv[reached_u] >= 0)
    assert np.all(lv[reached_u] <= lu[reached_u] + 1)
    del lu, lv, reached_u
    # parents: one level up and a neighbour (adjacency lists are sorted)
    v = np.nonzero(level > 0)[0]
    p = parent[v]
    assert np.all(level[p] == level[v] - 1)
    rng = np.random.default_rng(5)
    for x in rng.choice(v, min(v.size, 20000), replace=False):
        px = parent[x]
        row = adj[off[px]:off[px + 1]]
        k = np.searchsorted(row, x)
        assert k < row.size and row[k] == x
    assert np.all(parent[level < 0] == -1)


def test_bfs_scale24_properties(pa):
    say("case: bfs symmetrised scale 24")
    gs = pa.Graph.rmat(24, 1 << 28, seed=1, symmetrize=True)
    off, adj, _ = gs.download()
    for src in (0, 16123652):
        level, parent, st = gs.bfs(src)
        say("  source %d: %d reached, %d rounds" % (src, int((level >= 0).sum()), st.rounds))
        _check_bfs(off, adj, src, level, parent)
    gs.close()


def test_cc_scale22_properties(pa):
    say("case: cc symmetrised scale 22")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True)
    off, adj, _ = gs.download()
    lab, st = gs.cc()
    n = gs.n
    src_of = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    assert np.all(lab[src_of] == lab[adj])  # constant on every edge
    assert np.all(lab <= np.arange(n))      # the minimum id of its component...
    assert np.all(lab[lab] == lab)          # ...which is labelled by itself
    say("  %d components, %d rounds" % (int((lab == np.arange(n)).sum()), st.rounds))
    gs.close()


def test_msbfs_scale22_matches_single_searches(pa):
    say("case: 64-source bfs scale 22 against 64 single-source searches")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True)
    rng = np.random.default_rng(13)
    src = rng.choice(gs.n, 64, replace=False)
    src[0] = 0
    L, st = gs.msbfs(src)
    for c in range(64):
        lv, _, _ = gs.bfs(int(src[c]))
        assert np.array_equal(L[c], lv), c
    say("  %d rounds, %d pull rounds" % (st.rounds, st.dense_rounds))
    gs.close()


def test_spmv_scale24_vs_fp64(pa, orc):
    say("case: spmv scale 24, fp32 values, 2 iterations against the fp64 oracle")
    g = pa.Graph.rmat(24, 1 << 28, seed=1, weights="f32", wseed=3)
    x0 = np.random.default_rng(3).random(g.n).astype(np.float32)
    y, _ = g.spmv(x0, 2)
    off, adj, w = g.download()
    ref = orc.spmv(orc.Graph(g.n, off, adj), w.astype(np.float64), x0.astype(np.float64), 2)
    e = rel_l1(y, ref)
    say("  rel L1 %.3e" % e)
    assert e < 1e-6
    g.close()


def test_sssp_scale22_vs_dijkstra(pa, orc):
    say("case: bellman-ford symmetrised scale 22, int32 weights, against Dijkstra")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True, weights="i32", wseed=7)
    off, adj, w = gs.download()
    hg = orc.Graph(gs.n, off, adj)
    for src in (0, 1234567):
        d, st = gs.sssp(src)
        ref = orc.dijkstra(hg, w, src)
        say("  source %d: %d rounds, %d reached" % (src, st.rounds, int((d < pa.INF_DIST).sum())))
        assert np.array_equal(d, ref), src
    gs.close()


def test_bp_scale22_vs_fp64(pa, orc):
    """BP with fp64 state, 10 iterations on the scale-22 rMat, against the oracle's
    fp64 products. Unnormalised products under- and overflow even in fp64 (a
    vertex's product is raised to about its in-degree per iteration), so, as in the
    small-graph test, the entries the oracle keeps inside [1e-30, 1e30] are compared
    (about 65% of them at this size) and must agree to 1e-6 relative L1."""
    say("case: bp scale 22, 10 iterations, fp64 state, against the fp64 oracle")
    g = pa.Graph.rmat(22, 1 << 26, seed=1)
    g.bp_potentials(9)
    vpot = orc.vertex_f32(2 * g.n, 8, 0.1, 1.0)
    prod0 = np.ones(2 * g.n, np.float32)
    got, _ = g.bp(vpot, prod0, 10)
    off, adj, _ = g.download()
    hg = orc.Graph(g.n, off, adj)
    ref = orc.bp(hg, vpot, orc.bp_potentials(hg, 9), prod0.astype(np.float64), 10)
    ok = np.isfinite(ref) & (np.abs(ref) >= 1e-30) & (np.abs(ref) <= 1e30)
    e = rel_l1(got[ok], ref[ok])
    say("  in-range entries %d of %d, rel L1 %.3e" % (ok.sum(), ok.size, e))
    assert ok.sum() > 0.5 * ok.size
    assert e < 1e-6
    g.close()
