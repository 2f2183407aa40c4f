"""GPU parity of graph ingest: the binary graph triple <base>.config / .adj / .idx
(readGraphFromBinary / readWghGraphFromBinary, IO-numa.h:543-720) through
pg_graph_load_bin, against the same graph as PBBS text through pg_graph_load_adj,
the oracle's binary reader and the textbook traversals."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "polymer_amd",
                   "apps", "bin")


def say(*a):
    print(*a, flush=True)


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
        assert p.returncode == 0, p.stderr
        r = np.fromfile(out, dtype=dtype)
    assert r.size == count
    return r


def test_load_bin_matches_text(pa, orc, tmp_path):
    g = orc.rmat(12, 1 << 16, seed=21)
    txt = str(tmp_path / "g.txt")
    orc.write_adj(txt, g)
    gt = pa.Graph.load_adj(txt)
    ot, at, _ = gt.download()
    for ib in (4, 8):
        base = str(tmp_path / ("g%d" % ib))
        orc.write_bin(base, g, ib)
        say("load_bin: n %d m %d, %d-byte offsets" % (g.n, g.m, ib))
        gb = pa.Graph.load_bin(base)
        ob, ab, wb = gb.download()
        assert wb is None
        assert np.array_equal(ob, ot) and np.array_equal(ab, at)
        h = orc.read_bin(base)
        assert np.array_equal(ob, h.off) and np.array_equal(ab, h.adj)
        for src in (0, 5):
            lev, _, _ = gb.bfs(src)
            assert np.array_equal(lev, orc.bfs_levels(h, src)), (ib, src)
        gb.close()
    gt.close()


def test_load_bin_weighted_unit_weights(pa, orc, tmp_path):
    """The weighted binary reader gives every edge weight 1 (IO-numa.h:683), so
    Bellman-Ford distances equal BFS levels (unreached: INF_DIST and -1)."""
    g = orc.rmat(11, 1 << 15, seed=23)
    base = str(tmp_path / "w")
    orc.write_bin(base, g, 8)
    gw = pa.Graph.load_bin(base, weighted=True)
    _, _, w = gw.download()
    assert w is not None and w.dtype == np.int32 and np.all(w == 1)
    lev = orc.bfs_levels(g, 0)
    dist, _ = gw.sssp(0)
    assert np.array_equal(np.where(dist == pa.INF_DIST, -1, dist), lev)
    gw.close()


def test_load_bin_symmetric_flag(pa, orc, tmp_path):
    gs = orc.rmat(11, 1 << 15, seed=25).symmetrized()
    base = str(tmp_path / "s")
    orc.write_bin(base, gs, 4)
    txt = str(tmp_path / "s.txt")
    orc.write_adj(txt, gs)
    a = pa.Graph.load_bin(base, symmetric=True)
    b = pa.Graph.load_adj(txt, symmetric=True)
    la, _ = a.cc()
    lb, _ = b.cc()
    assert np.array_equal(la, lb) and np.array_equal(la, orc.cc_unionfind(gs))
    a.close()
    b.close()


def test_load_bin_errors(pa, orc, tmp_path):
    g = orc.rmat(8, 1 << 10, seed=3)
    base = str(tmp_path / "g")
    orc.write_bin(base, g, 4)
    with open(base + ".idx", "ab") as f:
        f.write(b"\0\0")
    with pytest.raises(RuntimeError, match="File size wrong"):
        pa.Graph.load_bin(base)
    with pytest.raises(RuntimeError, match="Unable to open"):
        pa.Graph.load_bin(str(tmp_path / "missing"))
    # an id beyond n is rejected by the graph build, not read out of bounds
    base2 = str(tmp_path / "bad_id")
    orc.write_bin(base2, g, 4)
    adj = np.fromfile(base2 + ".adj", np.uint32)
    adj[3] = g.n + 7
    adj.tofile(base2 + ".adj")
    with pytest.raises(RuntimeError, match="out of range"):
        pa.Graph.load_bin(base2)


def test_dropin_binary_flag(orc, tmp_path):
    """The drivers' -b (the reference's flag for a binary file): Components on a
    symmetric binary graph gives the union-find labels."""
    gs = orc.rmat(12, 1 << 16, seed=27).symmetrized()
    base = str(tmp_path / "cc")
    orc.write_bin(base, gs, 4)
    lab = run("components", ["-file", base, "-b"], np.int32, gs.n)
    assert np.array_equal(lab, orc.cc_unionfind(gs))


def _dump_bytes(g):
    """The dumpGraphToBin layout (IO-numa.h:837-900) written field by field:
    int64 total, int32 n, int64 m, then per vertex -1, id, out-degree, in-degree,
    out-neighbours, in-neighbours (in-lists in ascending source order)."""
    n, m = g.n, g.m
    deg = np.diff(g.off)
    src_of = np.repeat(np.arange(n, dtype=np.int32), deg)
    order = np.lexsort((src_of, g.adj))
    in_src = src_of[order]
    ioff = np.zeros(n + 1, np.int64)
    np.add.at(ioff, g.adj.astype(np.int64) + 1, 1)
    ioff = np.cumsum(ioff)
    total = 20 + 16 * n + 8 * m
    parts = [np.array([total], "<i8").tobytes(), np.array([n], "<i4").tobytes(),
             np.array([m], "<i8").tobytes()]
    for v in range(n):
        parts.append(np.array([-1, v, deg[v], ioff[v + 1] - ioff[v]], "<i4").tobytes())
        parts.append(g.adj[g.off[v]:g.off[v + 1]].astype("<i4").tobytes())
        parts.append(in_src[ioff[v]:ioff[v + 1]].astype("<i4").tobytes())
    b = b"".join(parts)
    assert len(b) == total
    return b


def test_load_dump_and_dump(pa, orc, tmp_path):
    g = orc.rmat(11, 1 << 14, seed=29)
    p = str(tmp_path / "g.bin")
    want = _dump_bytes(g)
    with open(p, "wb") as f:
        f.write(want)
    gd = pa.Graph.load_dump(p)
    gc = pa.Graph.from_csr(g.n, g.off, g.adj)
    od, ad, _ = gd.download()
    oc, ac, _ = gc.download()
    assert np.array_equal(od, oc) and np.array_equal(ad, ac)
    lev, _, _ = gd.bfs(3)
    assert np.array_equal(lev, orc.bfs_levels(g, 3))
    p2 = str(tmp_path / "g2.bin")
    gc.dump(p2)
    with open(p2, "rb") as f:
        got = f.read()
    assert got == want  # the same bytes as the reference's layout
    gd.close()
    gc.close()


def test_load_dump_rejects_corruption(pa, orc, tmp_path):
    g = orc.rmat(8, 1 << 10, seed=31)
    good = bytearray(_dump_bytes(g))
    cases = {}
    b = bytearray(good)
    b[20:24] = np.array([-2], "<i4").tobytes()  # vertex 0's marker
    cases["marker"] = b
    b = bytearray(good)
    b[32:36] = np.array([int(np.frombuffer(good[32:36], "<i4")[0]) + 1], "<i4").tobytes()
    cases["in-degree"] = b  # vertex 0's recorded in-degree off by one (sizes then disagree)
    cases["truncated"] = good[:-4]
    for what, data in cases.items():
        p = str(tmp_path / ("bad_" + what.replace("-", "")))
        with open(p, "wb") as f:
            f.write(bytes(data))
        with pytest.raises(RuntimeError, match="Bad dump file"):
            pa.Graph.load_dump(p)
        say("dump corruption rejected:", what)


# ---- PBBS text, parsed on the GPU ---------------------------------------------------
def test_load_adj_text_forms(pa, orc, tmp_path):
    """pg_graph_load_adj (IO-numa.h:112-190, IO.h:186-272) against the oracle's
    reader: an unweighted and a weighted rMat as the oracle writes them; the same
    words with tabs, CRLF line ends, runs of blanks, signs and no final newline;
    self-loops kept or dropped (PG_GRAPH_SKIP_SELF_LOOPS, IO-numa.h:225-268)."""
    g = orc.rmat(13, 1 << 17, seed=31)
    w = orc.weights_i32(g, 5)
    p = str(tmp_path / "g.adj")
    orc.write_adj(p, g)
    d = pa.Graph.load_adj(p)
    o, a, _ = d.download()
    assert np.array_equal(o, g.off) and np.array_equal(a, g.adj)
    d.close()
    pw = str(tmp_path / "g.wadj")
    orc.write_adj(pw, g, w)
    d = pa.Graph.load_adj(pw)
    o, a, ww = d.download()
    h, hw = orc.read_adj(pw)
    assert np.array_equal(o, g.off) and np.array_equal(a, g.adj)
    assert ww is not None and np.array_equal(ww, w) and np.array_equal(ww, hw)
    d.close()
    # the same small weighted graph with unusual separators, and a self-loop
    off, adj, wt = [0, 2, 3, 5], [1, 2, 2, 0, 2], [7, +3, 9, 1, 4]
    text = ("WeightedAdjacencyGraph\r\n3\t\t5\r\n" + " ".join(map(str, off[:3])) + "\n\n" +
            "\t".join(map(str, adj)) + "   \r\n" + " +7 3\n9\t1  4")  # no final newline
    q = str(tmp_path / "odd.wadj")
    open(q, "w").write(text)
    for skip in (False, True):
        d = pa.Graph.load_adj(q, skip_self_loops=skip)
        o, a, ww = d.download()
        if skip:  # 2 -> 2 is dropped with its weight
            assert list(o) == [0, 2, 3, 4] and list(a) == [1, 2, 2, 0] and list(ww) == [7, 3, 9, 1]
        else:
            assert list(o) == off and list(a) == adj and list(ww) == [7, 3, 9, 1, 4]
        h, hw = orc.read_adj(q, skip_self_loops=skip)
        assert np.array_equal(o, h.off) and np.array_equal(a, h.adj) and np.array_equal(ww, hw)
        d.close()


def test_load_adj_text_errors(pa, tmp_path):
    """The reader's refusals (IO-numa.h:75-78, 116-122) as error codes: a missing
    file, a wrong header, too few offsets, edges or weights, a word too many, and a
    neighbour id out of range."""
    cases = [
        ("Graph\n1\n0\n0\n", "Bad input file"),
        ("AdjacencyGraph\n3\n2\n0\n1\n", "Bad input file \\(offsets\\)"),
        ("AdjacencyGraph\n3\n2\n0\n1\n1\n2\n", "Bad input file \\(edges\\)"),
        ("WeightedAdjacencyGraph\n3\n2\n0\n1\n1\n2\n0\n5\n", "Bad input file \\(weights\\)"),
        ("AdjacencyGraph\n3\n2\n0\n1\n1\n2\n0\n0\n", "Bad input file \\(length\\)"),
        ("AdjacencyGraph\n3\n2\n0\n1\n1\n2\n7\n", "out of range"),
    ]
    for i, (text, msg) in enumerate(cases):
        p = str(tmp_path / ("e%d.adj" % i))
        open(p, "w").write(text)
        say("case: %r" % msg)
        with pytest.raises(RuntimeError, match=msg):
            pa.Graph.load_adj(p)
    with pytest.raises(RuntimeError, match="Unable to open"):
        pa.Graph.load_adj(str(tmp_path / "missing.adj"))
