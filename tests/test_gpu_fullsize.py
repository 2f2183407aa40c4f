"""The benchmark's workloads at their full sizes against the oracle.

* Value propagation on the scale-24 rMat (2^28 edges): PageRank for all 20
  iterations against the oracle's fp64 pull, SpMV for 20 iterations (fp32 values)
  and BP for 10 iterations (fp64 state) against the oracle's OpenMP pulls, which
  are bit-identical to its push restatements (tests/test_oracle.py).
* Traversal on the symmetrised scale-22 rMat: BFS levels from the nine sources of
  the traversal workload bit-exact against the oracle's BFS (parents checked for
  validity), CC labels bit-exact against union-find, Bellman-Ford distances
  bit-exact against Dijkstra; BFS on the symmetrised scale-24 graph by the
  properties that define breadth-first levels and parents.
* Batching: the 64-source BFS and Bellman-Ford on scale 22 against the oracle on
  sixteen of the columns (and the BFS against 64 single-source GPU searches);
  the 512-graph PageRank batch, every graph against the oracle.
* Differentiation: the weighted-PageRank adjoint on the n = 2^20, m = 2^24 rMat
  against fp64 central differences (h = 1e-6) of the oracle's forward solve.

The oracle runs on the host's cores: ctypes releases the GIL, so independent
oracle calls run on a thread pool."""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (workload definitions shared with the benchmark)

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def say(*a):
    print(*a, flush=True)


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def pool_map(fn, items):
    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        return list(ex.map(fn, items))


# ---- value propagation, scale 24 ----------------------------------------------------------
def test_pagerank_scale24_20_iterations(pa, orc):
    say("case: pagerank scale 24, 20 iterations against the fp64 oracle pull")
    g = pa.Graph.rmat(24, 1 << 28, seed=1)
    r, st = g.pagerank(20, 0.85)
    off, adj, _ = g.download()
    hg = orc.Graph(g.n, off, adj)
    hg.transpose()
    ref = orc.pagerank_pull_omp(hg, 20, 0.85, threads=THREADS)
    e = rel_l1(r, ref)
    say("  rel L1 %.3e over %d edges processed" % (e, st.edges))
    assert e < 1e-6
    # P1: vertices without in-edges hold 0 after an even number of iterations
    assert np.array_equal(r == 0, ref == 0)
    g.close()


def test_spmv_scale24_20_iterations(pa, orc):
    say("case: spmv scale 24, fp32 values, 20 iterations against the fp64 oracle")
    g = pa.Graph.rmat(24, 1 << 28, seed=1, weights="f32", wseed=3)
    x0 = np.random.default_rng(3).random(g.n).astype(np.float32)  # as the benchmark
    y, _ = g.spmv(x0, 20)
    off, adj, w = g.download()
    g.close()
    hg = orc.Graph(g.n, off, adj)
    del off, adj
    ref = orc.spmv_pull_omp(hg, w.astype(np.float64), x0.astype(np.float64), 20,
                            threads=THREADS)
    e = rel_l1(y, ref)
    say("  rel L1 %.3e (max |y| %.3e)" % (e, float(np.abs(ref).max())))
    assert e < 1e-6
    # exact zeros (no in-edges, or only zero-valued sources) stay exact; the shared
    # per-iteration exponent truncates terms 2^-60 below the largest value's scale,
    # so an entry may come out 0 only where the oracle's is negligible
    assert np.all(y[ref == 0] == 0)
    lost = (y == 0) & (ref != 0)
    say("  entries zero on the GPU only: %d, largest oracle value among them %.3e" %
        (int(lost.sum()), float(np.abs(ref[lost]).max()) if lost.any() else 0.0))
    assert np.all(np.abs(ref[lost]) <= 2.0 ** -40 * np.abs(ref).max())


def bp_classes(x):
    """0 zero, 1 finite nonzero, 2 infinite, 3 NaN."""
    c = np.ones(x.shape, np.int8)
    c[x == 0] = 0
    c[np.isinf(x)] = 2
    c[np.isnan(x)] = 3
    return c


def test_bp_scale24_10_iterations(pa, orc):
    """BP with fp64 state, 10 iterations on the scale-24 rMat with the benchmark's
    potentials.  Unnormalised products raise a vertex's value to about its
    in-degree per iteration, so many leave fp64's range in the oracle too.  Checked:
    the entries the oracle keeps in [1e-30, 1e30] to 1e-6 relative L1; every entry
    in [1e-290, 1e290] to 1e-9 relative; and both sides agree on which entries are
    zero, finite, infinite or NaN."""
    say("case: bp scale 24, 10 iterations, fp64 state, against the fp64 oracle")
    g = pa.Graph.rmat(24, 1 << 28, seed=1)
    g.bp_potentials(9)
    vpot = np.random.default_rng(8).uniform(0.1, 1.0, 2 * g.n).astype(np.float32)
    prod0 = np.ones(2 * g.n, np.float32)
    got, _ = g.bp(vpot, prod0, 10)
    off, adj, _ = g.download()
    g.close()
    hg = orc.Graph(g.n, off, adj)
    epot = orc.bp_potentials(hg, 9)
    ref = orc.bp_pull_omp(hg, vpot, epot, prod0.astype(np.float64), 10, threads=THREADS)
    del epot
    ok = np.isfinite(ref) & (np.abs(ref) >= 1e-30) & (np.abs(ref) <= 1e30)
    e = rel_l1(got[ok], ref[ok])
    wide = np.isfinite(ref) & (np.abs(ref) >= 1e-290) & (np.abs(ref) <= 1e290)
    worst = float(np.max(np.abs(got[wide] - ref[wide]) / np.abs(ref[wide])))
    cg, cr = bp_classes(got), bp_classes(ref)
    counts = np.bincount(cr.ravel(), minlength=4)
    say("  in-range entries %d of %d, rel L1 %.3e; worst relative error in [1e-290, 1e290] "
        "%.3e; oracle zero/finite/inf/nan %s; class mismatches %d" %
        (ok.sum(), ok.size, e, worst, counts.tolist(), int((cg != cr).sum())))
    assert ok.sum() > 0.3 * ok.size
    assert e < 1e-6
    assert worst < 1e-9
    assert np.array_equal(cg, cr)



def test_load_adj_scale24_timed(pa, orc, tmp_path):
    """PBBS AdjacencyGraph text of the scale-24 rMat (2^28 edges, about 2.6 GB)
    written by the oracle, loaded by pg_graph_load_adj (read into pinned memory
    in parallel, parsed on the GPU) and compared with the generator's graph; the
    load is timed."""
    import time
    say("case: PBBS text load, scale 24")
    g = pa.Graph.rmat(24, 1 << 28, seed=1)
    off, adj, _ = g.download()
    g.close()
    hg = orc.Graph(len(off) - 1, off, adj)
    path = str(tmp_path / "s24.adj")
    t0 = time.perf_counter()
    orc.write_adj(path, hg)
    size = os.path.getsize(path)
    say("  oracle wrote %d bytes in %.1f s" % (size, time.perf_counter() - t0))
    t0 = time.perf_counter()
    d = pa.Graph.load_adj(path)
    dt = time.perf_counter() - t0
    say("  pg_graph_load_adj: %.3f s (%.2f GB/s of text), n %d m %d" % (dt, size / dt / 1e9, d.n, d.m))
    o2, a2, _ = d.download()
    d.close()
    assert np.array_equal(o2, off) and np.array_equal(a2, adj)

# ---- traversal, scale 22 ---------------------------------------------------------------------
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
    # (parent, vertex) is an edge: CSR rows are sorted, so the keys src * n + dst
    # of all edges are sorted and a binary search finds each pair
    keys = np.repeat(np.arange(n, dtype=np.int64), deg) * n + adj
    want = p.astype(np.int64) * n + v
    k = np.searchsorted(keys, want)
    assert np.all(k < keys.size) and np.array_equal(keys[np.minimum(k, keys.size - 1)], want)
    del keys
    assert np.all(parent[level < 0] == -1)


def test_bfs_scale22_nine_sources(pa, orc):
    say("case: bfs symmetrised scale 22, the traversal workload's nine sources")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True)
    off, adj, _ = gs.download()
    hg = orc.Graph(gs.n, off, adj)
    srcs = bench.traversal_sources(np.diff(off))
    refs = pool_map(lambda s: orc.bfs_levels(hg, s), srcs)
    for s, ref in zip(srcs, refs):
        level, parent, st = gs.bfs(s)
        say("  source %d: %d reached, %d rounds (%d bottom-up)" %
            (s, int((level >= 0).sum()), st.rounds, st.dense_rounds))
        assert np.array_equal(level, ref), s
        _check_bfs(off, adj, s, level, parent)
    # the reference's own setting (threshold 0: every round dense) gives the same levels
    level, parent, st = gs.bfs(srcs[1], threshold=0)
    assert st.dense_rounds == st.rounds
    assert np.array_equal(level, refs[1])
    _check_bfs(off, adj, srcs[1], level, parent)
    gs.close()


def test_cc_scale22_vs_unionfind(pa, orc):
    say("case: cc symmetrised scale 22 against union-find")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True)
    off, adj, _ = gs.download()
    lab, st = gs.cc()
    ref = orc.cc_unionfind(orc.Graph(gs.n, off, adj))
    say("  %d components, %d rounds" % (int((ref == np.arange(gs.n)).sum()), st.rounds))
    assert np.array_equal(lab, ref)
    gs.close()


def test_sssp_scale22_nine_sources_vs_dijkstra(pa, orc):
    say("case: bellman-ford symmetrised scale 22, int32 weights, nine sources, Dijkstra")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True, weights="i32", wseed=7)
    off, adj, w = gs.download()
    hg = orc.Graph(gs.n, off, adj)
    srcs = bench.traversal_sources(np.diff(off))
    refs = pool_map(lambda s: orc.dijkstra(hg, w, s), srcs)
    for s, ref in zip(srcs, refs):
        d, st = gs.sssp(s)
        say("  source %d: %d rounds, %d reached" % (s, st.rounds, int((d < pa.INF_DIST).sum())))
        assert np.array_equal(d, ref), s
    gs.close()


def test_bfs_scale24_properties(pa):
    say("case: bfs symmetrised scale 24")
    gs = pa.Graph.rmat(24, 1 << 28, seed=1, symmetrize=True)
    off, adj, _ = gs.download()
    for src in (0, 16123652):
        level, parent, st = gs.bfs(src)
        say("  source %d: %d reached, %d rounds" % (src, int((level >= 0).sum()), st.rounds))
        _check_bfs(off, adj, src, level, parent)
    gs.close()


# ---- batching ------------------------------------------------------------------------------------
def multisource_sources(n):
    """The benchmark's 64 sources (multisource_extra): seeded, vertex 0 first."""
    rng = np.random.default_rng(13)
    src = rng.choice(n, 64, replace=False)
    src[0] = 0
    return src


def test_msbfs_scale22_vs_oracle(pa, orc):
    say("case: 64-source bfs scale 22 against the oracle and 64 single-source searches")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True)
    src = multisource_sources(gs.n)
    L, st = gs.msbfs(src)
    say("  %d rounds, %d pull rounds" % (st.rounds, st.dense_rounds))
    off, adj, _ = gs.download()
    hg = orc.Graph(gs.n, off, adj)
    cols = list(range(16))
    refs = pool_map(lambda c: orc.bfs_levels(hg, int(src[c])), cols)
    for c, ref in zip(cols, refs):
        assert np.array_equal(L[c], ref), c
    for c in range(64):
        lv, _, _ = gs.bfs(int(src[c]))
        assert np.array_equal(L[c], lv), c
    gs.close()


def test_mssssp_scale22_vs_dijkstra(pa, orc):
    say("case: 64-source bellman-ford scale 22 against Dijkstra on 16 columns")
    gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True, weights="i32", wseed=7)
    src = multisource_sources(gs.n)
    D, st = gs.mssssp(src)
    say("  %d rounds" % st.rounds)
    off, adj, w = gs.download()
    hg = orc.Graph(gs.n, off, adj)
    cols = list(range(16))
    refs = pool_map(lambda c: orc.dijkstra(hg, w, int(src[c])), cols)
    for c, ref in zip(cols, refs):
        assert np.array_equal(D[c], ref), c
    gs.close()


def test_multigraph_512(pa, orc):
    say("case: 512 graphs (n = 2^8..2^16, degree 16) in one captured launch")
    vb, off, adj = bench.multigraph_inputs(pa, 512)
    d = pa.Graph.multi(vb, off, adj)
    d.autotune()
    r, st = d.pagerank_multigraph(20, 0.85)
    say("  n %d, m %d, %.3f ms" % (d.n, d.m, st.ms))

    def one(i):
        a, b = int(vb[i]), int(vb[i + 1])
        h = orc.Graph(b - a, off[a:b + 1] - off[a], adj[off[a]:off[b]] - a)
        ref, _ = orc.pagerank(h, 20, 0.85)
        return rel_l1(r[a:b], ref)

    errs = pool_map(one, range(512))
    say("  worst rel L1 %.3e" % max(errs))
    assert max(errs) < 1e-6
    d.close()


# ---- differentiation ---------------------------------------------------------------------------------
def test_wpr_adjoint_fd_scale20(pa, orc):
    """Gradients of L = g . x(W, v) for weighted PageRank on the n = 2^20, m = 2^24
    rMat (fp32 weights U(0,1], column-normalised, damping 0.85, g ~ N(0,1) seeded)
    against fp64 central differences (h = 1e-6) of the oracle's forward solve, on
    the six edges with the largest gradients, six random edges and six teleport
    entries."""
    say("case: weighted pagerank adjoint, n = 2^20, m = 2^24, against central differences")
    g = pa.Graph.rmat(20, 1 << 24, seed=2, weights="f32", wseed=4)
    off, adj, w = g.download()
    src = np.repeat(np.arange(g.n), np.diff(off))
    tot = np.bincount(src, weights=w.astype(np.float64), minlength=g.n)
    wn = w.astype(np.float64) / tot[src]
    rng = np.random.default_rng(21)
    v = rng.random(g.n)
    v /= v.sum()
    gv = rng.standard_normal(g.n)
    dmp = 0.85
    x, stf = g.wpr_forward(wn.astype(np.float32), v.astype(np.float32), dmp, 1000, 1e-8)
    lam, gw, gt, sta = g.wpr_adjoint(gv.astype(np.float32), 1000, 1e-8 * float(np.abs(gv).sum()))
    say("  forward %d iterations, adjoint %d" % (stf.rounds, sta.rounds))
    hg = orc.Graph(g.n, off, adj)
    g.close()
    # a fixed 200 iterations (tolerance 0; 0.85^200 < 1e-14): every solve below
    # runs the same steps, so its truncation error cancels in the differences
    x_ref, _ = orc.wpr_forward(hg, wn, v, dmp, 200, 0.0)
    assert rel_l1(x, x_ref) < 1e-6
    edges = list(np.argsort(-np.abs(gw))[:6]) + list(rng.choice(g.m, 6, replace=False))
    verts = list(rng.choice(g.n, 6, replace=False))
    h = 1e-6

    def fd(item):
        kind, i = item
        a, b = (wn.copy(), wn.copy()) if kind == "w" else (v.copy(), v.copy())
        a[i] += h
        b[i] -= h
        xp, _ = orc.wpr_forward(hg, a if kind == "w" else wn, a if kind == "v" else v, dmp,
                                200, 0.0)
        xm, _ = orc.wpr_forward(hg, b if kind == "w" else wn, b if kind == "v" else v, dmp,
                                200, 0.0)
        return (gv @ xp - gv @ xm) / (2 * h)

    items = [("w", int(e)) for e in edges] + [("v", int(i)) for i in verts]
    fds = pool_map(fd, items)
    scale_w = float(np.abs(gw).max())
    scale_v = float(np.abs(gt).max())
    for (kind, i), f in zip(items, fds):
        got = float(gw[i] if kind == "w" else gt[i])
        tol = 1e-4 * abs(f) + 1e-6 * (scale_w if kind == "w" else scale_v)
        say("  %s[%d]: adjoint %.9e, central difference %.9e" % (kind, i, got, f))
        assert abs(got - f) <= tol, (kind, i, got, f)


def test_spmv_adjoint_scale22_20_iterations(pa, orc):
    """The SpMV adjoint over 20 iterations on the scale-22 rMat (fp32 values U(0,1],
    x0 U[0,1), g ~ N(0,1)) against the oracle's fp64 adjoint."""
    say("case: spmv adjoint scale 22, 20 iterations, against the fp64 oracle")
    g = pa.Graph.rmat(22, 1 << 26, seed=1, weights="f32", wseed=3)
    rng = np.random.default_rng(3)
    x0 = rng.random(g.n).astype(np.float32)
    gv = rng.standard_normal(g.n).astype(np.float32)
    _, stf = g.spmv(x0, 20, save=True)
    gx, gw, sta = g.spmv_adjoint(gv)
    say("  forward %.3f ms, adjoint %.3f ms" % (stf.ms, sta.ms))
    off, adj, w = g.download()
    g.close()
    rx, rw = orc.spmv_adjoint(orc.Graph(g.n, off, adj), w.astype(np.float64),
                              x0.astype(np.float64), gv.astype(np.float64), 20)
    ex, ew = rel_l1(gx, rx), rel_l1(gw, rw)
    say("  dL/dx0 rel L1 %.3e, dL/dw rel L1 %.3e (max |dL/dw| %.3e)" %
        (ex, ew, float(np.abs(rw).max())))
    assert ex < 1e-6 and ew < 1e-6
