"""The reference's own programs built with include/polymer.h in place of its header
(oracle/header_swap.py, into oracle/_ref/swap/), run on the GPU on the PR1-size rMat
graph and checked against the reference binaries' golden output and the oracle:

  numa-PageRank  20 iterations, -result      ranks vs the reference binary's (golden)
  numa-SPMV      10 iterations, -result      values vs the reference binary's (golden)
  numa-BFS       -result -s, four sources    vertices visited and rounds (golden)
  numa-Components -s -result                 labels vs union-find
  numa-BFS-async-pipe -result -s             vertices with a parent (golden reach)
  numa-BellmanFord                           finishes; frontier totals vs Dijkstra's reach
  numa-BP        10 iterations               runs to its timing line

The binaries exist only where the reference sources were present at build time;
without them the tests skip."""
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(ROOT, "oracle", "_ref", "swap")
GOLD = os.path.join(HERE, "golden")


def say(*a):
    print(*a, flush=True)


def exe(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.skip("%s not built (the reference sources were absent at build time)" % name)
    return p


def run(args, timeout=300):
    say("  run:", " ".join(os.path.basename(a) if i == 0 else a for i, a in enumerate(args)))
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stderr[-3000:], r.stdout[-2000:])
    return r.stdout


def tab_values(out, n, dtype):
    vals = np.full(n, -1, dtype=dtype)
    seen = 0
    for line in out.splitlines():
        parts = line.split("\t")
        if len(parts) == 2 and parts[0].isdigit():
            vals[int(parts[0])] = dtype(float(parts[1]))
            seen += 1
    assert seen == n, (seen, n)
    return vals


def rel_l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).sum() / np.abs(b).sum())


@pytest.fixture(scope="module")
def files(orc):
    z = np.load(os.path.join(GOLD, "pr1_pagerank.npz"))
    g = orc.rmat(int(z["scale"]), int(z["m"]), float(z["a"]), float(z["b"]), float(z["c"]),
                 int(z["seed"]))
    gs = g.symmetrized()
    zs = np.load(os.path.join(GOLD, "spmv_int.npz"))
    w = orc.weights_i32(g, int(zs["weight_seed"]))
    with tempfile.TemporaryDirectory() as td:
        paths = dict(adj=os.path.join(td, "pr1.adj"), sym=os.path.join(td, "pr1.sym"),
                     wadj=os.path.join(td, "pr1.wadj"))
        orc.write_adj(paths["adj"], g)
        orc.write_adj(paths["sym"], gs)
        orc.write_adj(paths["wadj"], g, w)
        yield dict(paths, g=g, gs=gs, w=w)


def test_swap_pagerank_vs_reference_golden(files):
    say("case: numa-PageRank built on polymer.h, PR1, 20 iterations")
    z = np.load(os.path.join(GOLD, "pr1_pagerank.npz"))
    out = run([exe("numa-PageRank"), files["adj"], "20", "-1", "-result"])
    r = tab_values(out, files["g"].n, np.float64)
    e = rel_l1(r, z["rank"])
    say("  rel L1 vs the reference binary %.3e" % e)
    assert e < 1e-9  # fp64 both; only the order of the atomic additions differs
    assert re.search(r"PageRank : [0-9.e+-]+", out)


def test_swap_spmv_vs_reference_golden(files):
    say("case: numa-SPMV built on polymer.h, PR1 integer weights, 10 iterations")
    z = np.load(os.path.join(GOLD, "spmv_int.npz"))
    out = run([exe("numa-SPMV"), files["wadj"], str(int(z["iters"])), "-result"])
    y = tab_values(out, files["g"].n, np.float64)
    e = rel_l1(y, z["y"])
    say("  rel L1 vs the reference binary %.3e" % e)
    assert e < 1e-9


def test_swap_bfs_vs_reference_golden(files):
    say("case: numa-BFS built on polymer.h, symmetric PR1")
    gold = json.load(open(os.path.join(GOLD, "sym_bfs.json")))
    for r in gold["runs"]:
        out = run([exe("numa-BFS"), files["sym"], str(r["src"]), "-result", "-s"])
        visited = int(re.search(r"Vertices visited = (\d+)", out).group(1))
        rounds = int(re.search(r"Finished in (\d+) iterations", out).group(1))
        counted = int(re.search(r"Vert visited: (\d+)", out).group(1))
        say("  source %d: visited %d rounds %d (reference %d, %d)" %
            (r["src"], visited, rounds, r["vertices_visited"], r["finished_in"]))
        assert visited == r["vertices_visited"] and rounds == r["finished_in"]
        assert counted == r["vertices_visited"]


def test_swap_bfs_async_pipe_vs_reference_golden(files):
    """numa-BFS-async-pipe.C (edgeMapSparseAsyncPipe, polymer.h:1512-1751): one
    asynchronous traversal; with -result it counts the vertices that got a parent,
    which must be the reachable set the reference's numa-BFS reports."""
    say("case: numa-BFS-async-pipe built on polymer.h, symmetric PR1")
    gold = json.load(open(os.path.join(GOLD, "sym_bfs.json")))
    for r in gold["runs"]:
        out = run([exe("numa-BFS-async-pipe"), files["sym"], str(r["src"]), "-result", "-s"])
        counted = int(re.search(r"Vert visited: (\d+)", out).group(1))
        say("  source %d: %d vertices with a parent (reference %d)" %
            (r["src"], counted, r["vertices_visited"]))
        assert counted == r["vertices_visited"]


def test_swap_components_vs_unionfind(files, orc):
    say("case: numa-Components built on polymer.h, symmetric PR1")
    out = run([exe("numa-Components"), files["sym"], "-s", "-result"])
    n = files["gs"].n
    lab = np.full(n, -1, np.int64)
    for m in re.finditer(r"Result of (\d+) : (-?\d+)", out):
        lab[int(m.group(1))] = int(m.group(2))
    ref = orc.cc_unionfind(files["gs"])
    assert np.array_equal(lab, ref)


def test_swap_bellmanford_runs(files, orc):
    """numa-BellmanFord prints no distances (its -result loop reads a pointer the
    program never sets, numa-BellmanFord.C:44,444, and crashes the reference too);
    checked: it finishes, and every vertex Dijkstra reaches entered some frontier
    (the frontier sizes it sums count re-improvements as well)."""
    say("case: numa-BellmanFord built on polymer.h, PR1 integer weights")
    g = files["g"]
    for src in (1, int(np.argmax(np.diff(g.off)))):
        out = run([exe("numa-BellmanFord"), files["wadj"], str(src)])
        visited = int(re.search(r"Vertices visited = (\d+)", out).group(1))
        reach = int((orc.dijkstra(g, files["w"], src) < 1073741823).sum())
        say("  source %d: frontier total %d, reachable %d" % (src, visited, reach))
        assert visited >= reach
        assert re.search(r"BellmanFord : [0-9.e+-]+", out)


def test_swap_bp_runs(files):
    say("case: numa-BP built on polymer.h, PR1, 10 iterations")
    out = run([exe("numa-BP"), files["adj"], "10"])
    assert re.search(r"BeliefPropagation : [0-9.e+-]+", out)
