"""GPU parity of the driver programs written against include/polymer.h (the
edgeMap / vertexMap interface: graph<vertex>, vertices, LocalFrontier, functors
with update / updateAtomic / cond run as device code).

Each driver runs as its own process on the GPU and writes its result array;
the checks are the ones the C-ABI paths meet:
  * PageRank, SpMV (fp64 values, fp64 atomics): the reference binary's PR1 and
    integer-weight SpMV goldens and the fp64 oracle; PageRankDelta: the fp64
    oracle;
  * BFS levels, CC labels, Bellman-Ford distances: bit-exact against the
    textbook BFS, union-find and Dijkstra; BFS parents checked for validity;
  * BP (fp64, atomic products): the fp64 oracle on the entries it keeps in range;
  * the header's count reporting: header_checks.cpp against host recounts.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "polymer_amd", "apps", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def say(*a):
    print(*a, flush=True)


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


LAST_STDOUT = [""]


def run(app, args, dtype, count, timeout=120):
    exe = os.path.join(BIN, app)
    if not os.path.exists(exe):
        pytest.fail("driver %s is not built (run __graft_entry__.build())" % exe)
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "out.bin")
        cmd = [exe] + [str(x) for x in args] + ["-o", out]
        say("run:", " ".join(cmd[1:]))
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        say(p.stdout.strip())
        LAST_STDOUT[0] = p.stdout
        assert p.returncode == 0, p.stderr
        r = np.fromfile(out, dtype=dtype)
    assert r.size == count
    return r


def check_parents(g, src, parents, levels):
    """Every reached v != src has an in-neighbour parent one level closer."""
    src_of = np.repeat(np.arange(g.n), np.diff(g.off))
    edges = set(zip(src_of.tolist(), g.adj.tolist()))
    assert parents[src] == src and levels[src] == 0
    for v in np.nonzero(levels > 0)[0][:5000]:
        p = int(parents[v])
        assert (p, int(v)) in edges and levels[p] == levels[v] - 1, v
    assert np.all(parents[levels < 0] == -1)


def test_dropin_pagerank(orc):
    z = np.load(os.path.join(GOLD, "pr1_pagerank.npz"))
    g = orc.rmat(16, 1 << 20, seed=1)
    r = run("pagerank", ["-rmat", 16, 1 << 20, 1, "-rounds", 20], np.float64, g.n)
    ref, _ = orc.pagerank(g, 20, 0.85)
    say("pagerank PR1: vs reference golden %.3e, vs fp64 oracle %.3e" %
        (rel_l1(r, z["rank"]), rel_l1(r, ref)))
    assert rel_l1(r, z["rank"]) < 1e-8  # golden printed with 10 significant digits
    assert rel_l1(r, ref) < 1e-12
    assert np.array_equal(r == 0, ref == 0)


@pytest.mark.parametrize("thr", [-1, 1 << 14])
def test_dropin_pagerank_delta(orc, thr):
    """PageRankDelta on the header: edgeMap DENSE_FORWARD + vertexFilter with the
    reference's threshold 0 (every round dense), and with a threshold that sends
    the late, small frontiers through the sparse push."""
    g = orc.rmat(14, 1 << 18, seed=41)
    args = ["-rmat", 14, 1 << 18, 41, "-rounds", 15] + (["-thr", thr] if thr > 0 else [])
    r = run("pagerank_delta", args, np.float64, g.n)
    sparse = int(LAST_STDOUT[0].split("sparse")[1].split()[0])
    ref, act = orc.pagerank_delta(g, 15, 0.85, 0.01)
    front = int(LAST_STDOUT[0].split("frontier")[1].split()[0])
    say("pagerank_delta thr %d: rel L1 %.3e, frontier %d (oracle %d), sparse rounds %d" %
        (thr, rel_l1(r, ref), front, int(act.sum()), sparse))
    assert rel_l1(r, ref) < 1e-9
    assert front == int(act.sum())
    assert (sparse > 0) == (thr > 0)


@pytest.mark.parametrize("thr", [-1, 0])
def test_dropin_bfs(orc, thr):
    g = orc.rmat(14, 1 << 18, seed=41)
    gs = g.symmetrized()
    for src, sym in ((0, True), (77, True), (0, False)):
        h = gs if sym else g
        args = ["-rmat", 14, 1 << 18, 41, "-r", src, "-thr", thr] + (["-s"] if sym else [])
        out = run("bfs", args, np.int32, 2 * g.n)
        parents, levels = out[:g.n], out[g.n:]
        assert np.array_equal(levels, orc.bfs_levels(h, src)), (src, sym)
        check_parents(h, src, parents, levels)


def odd_graph(orc, n=6001, scale=13, m=1 << 17, seed=61):
    """An rMat graph cut to its first n vertices: n is no multiple of 16, 32 or 64, so
    the header's frontier kernels run their partial-word tails."""
    g = orc.rmat(scale, m, seed=seed)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    keep = (src < n) & (g.adj < n)
    s, d = src[keep], g.adj[keep]
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, s + 1, 1)
    return orc.Graph(n, np.cumsum(off), d)


def test_dropin_odd_sizes_from_file(orc):
    """The drivers on PBBS files of a 6001-vertex graph: PageRank against the fp64
    oracle, BFS levels (every round dense, and direction-switching) and CC labels
    bit-exact on its symmetrised copy."""
    g = odd_graph(orc)
    gs = g.symmetrized()
    assert g.n % 16 and gs.m > 0
    with tempfile.TemporaryDirectory() as d:
        fa, fs = os.path.join(d, "g.adj"), os.path.join(d, "gs.adj")
        orc.write_adj(fa, g)
        orc.write_adj(fs, gs)
        r = run("pagerank", ["-file", fa, "-rounds", 20], np.float64, g.n)
        ref, _ = orc.pagerank(g, 20, 0.85)
        say("pagerank n %d: rel L1 %.3e" % (g.n, rel_l1(r, ref)))
        assert rel_l1(r, ref) < 1e-9
        for src, thr in ((0, 0), (5999, -1), (1234, -1)):
            out = run("bfs", ["-file", fs, "-s", "-r", src, "-thr", thr], np.int32, 2 * g.n)
            parents, levels = out[:g.n], out[g.n:]
            assert np.array_equal(levels, orc.bfs_levels(gs, src)), (src, thr)
            check_parents(gs, src, parents, levels)
        lab = run("components", ["-file", fs, "-s"], np.int32, g.n)
        assert np.array_equal(lab, orc.cc_unionfind(gs))


def test_dropin_components(orc):
    g = orc.rmat(14, 1 << 18, seed=41)
    gs = g.symmetrized()
    for thr in (-1, 0, 1 << 40):  # default switch, always dense, always sparse
        lab = run("components", ["-rmat", 14, 1 << 18, 41, "-thr", thr], np.int32, g.n)
        assert np.array_equal(lab, orc.cc_unionfind(gs)), thr


def test_dropin_bellmanford(orc):
    g = orc.rmat(13, 1 << 17, seed=43)
    gs = g.symmetrized()
    w = orc.weights_i32(gs, 7)
    for src, thr in ((0, -1), (5, 0), (5, 1 << 40)):
        d = run("bellmanford", ["-rmat", 13, 1 << 17, 43, "-s", "-wseed", 7, "-r", src, "-thr", thr],
                np.int32, g.n)
        assert np.array_equal(d, orc.dijkstra(gs, w, src)), (src, thr)


def test_dropin_spmv(orc):
    z = np.load(os.path.join(GOLD, "spmv_int.npz"))
    g = orc.rmat(16, 1 << 20, seed=1)
    iters, ws = int(z["iters"]), int(z["weight_seed"])
    y = run("spmv", ["-rmat", 16, 1 << 20, 1, "-wseed", ws, "-rounds", iters], np.float64, g.n)
    w = orc.weights_i32(g, ws)
    ref = orc.spmv(g, w.astype(np.float64), np.full(g.n, 1.0 / g.n), iters)
    say("spmv: vs reference golden %.3e, vs fp64 oracle %.3e" % (rel_l1(y, z["y"]), rel_l1(y, ref)))
    assert rel_l1(y, z["y"]) < 1e-8
    assert rel_l1(y, ref) < 1e-12


def test_dropin_bp(orc):
    g = orc.rmat(12, 1 << 16, seed=8)
    vpot = orc.vertex_f32(2 * g.n, 8, 0.1, 1.0)
    epot = orc.bp_potentials(g, 9)
    prod0 = np.ones(2 * g.n)
    with tempfile.TemporaryDirectory() as d:
        paths = [os.path.join(d, x) for x in ("v.bin", "e.bin", "p.bin")]
        vpot.astype(np.float32).tofile(paths[0])
        epot.astype(np.float32).tofile(paths[1])
        prod0.astype(np.float64).tofile(paths[2])
        for iters in (1, 3):
            got = run("bp", ["-rmat", 12, 1 << 16, 8, "-rounds", iters, "-in1", paths[0],
                             "-in2", paths[1], "-in3", paths[2]], np.float64, 2 * g.n)
            ref = orc.bp(g, vpot, epot, prod0, iters).ravel()
            ok = np.isfinite(ref) & (np.abs(ref) >= 1e-30) & (np.abs(ref) <= 1e30)
            err = rel_l1(got[ok], ref[ok])
            say("bp iters %d: in-range %d of %d, relL1 %.3e" % (iters, ok.sum(), ok.size, err))
            assert ok.sum() > 0.3 * ok.size
            assert err < 1e-12


def test_dropin_header_count_reporting():
    """The header's count reporting (polymer_amd/apps/header_checks.cpp): a dense
    edgeMap's own count of its output against a host recount and after the flags
    change, pushes on both sides of the small-list limit (successes, their
    out-degree sum, the output list), vertexFilter's count."""
    exe = os.path.join(BIN, "header_checks")
    if not os.path.exists(exe):
        pytest.fail("driver %s is not built (run __graft_entry__.build())" % exe)
    say("run: header_checks")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    say(p.stdout.strip())
    assert p.returncode == 0 and "header_checks: ok" in p.stdout, p.stdout + p.stderr
