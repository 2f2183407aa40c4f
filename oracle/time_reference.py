"""Times the reference's own programs (oracle/_ref/, built from the reference sources
by `make -C oracle ref`, the reference Makefile's default rule) on the PR1 and
scale-22 configurations, in the build container where those sources exist:

  * each program's own timing line (PageRank :, SPMV :, BeliefPropagation :,
    BFS :, Components :, BellmanFord :), which covers its iterations and not the
    graph load, median of --runs runs, without -result;
  * the programs start numa_num_configured_cpus() threads: all the host's cores,
    printed with the results;
  * graphs: the oracle's rMat (a/b/c/d 0.57/0.19/0.19/0.05, seed 1) written as PBBS
    text, the same graphs the GPU runs use; SpMV and Bellman-Ford read integer
    weights U[1,255] (the reference's weighted reader is integer-only); BP has no
    inputs of its own (its potentials are zero-filled, numa-BP.C:557-590).

    python oracle/time_reference.py [--runs 3] [--out results.json] [--scale24]

--scale24 adds numa-PageRank on the scale-24 rMat (n = 2^24, m = 2^28, 20
iterations; the headline configuration), run once (its graph load alone takes
minutes).

numa-BFS and numa-Components can abort while their workers exit (a barrier that
outlives the worker owning it), sometimes before the timing line is printed: such a
run is repeated, up to four times, and the attempts are reported.
"""
import argparse
import json
import os
import re
import statistics
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import pyoracle as O  # noqa: E402

BIN = os.path.join(HERE, "_ref")


def run_timed(args, label, runs, timeout=3600):
    """Median of the program's own '<label> : <seconds>' line over runs."""
    vals = []
    for _ in range(runs):
        for attempt in range(4):
            t0 = time.time()
            p = subprocess.run(["stdbuf", "-oL"] + args, capture_output=True, text=True,
                               timeout=timeout)
            m = re.search(r"^%s : ([0-9.eE+-]+)" % re.escape(label), p.stdout, re.M)
            if m or p.returncode >= 0:
                break
            print("  %s: aborted (exit %d) before its timing line, again" %
                  (os.path.basename(args[0]), p.returncode), flush=True)
        if not m:
            raise RuntimeError("%s: no '%s :' line (exit %d): %s" %
                               (args[0], label, p.returncode, p.stderr[-1500:]))
        vals.append(float(m.group(1)))
        print("  %s %s: %.4f s (wall %.1f s, exit %d)" %
              (os.path.basename(args[0]), label, vals[-1], time.time() - t0, p.returncode),
              flush=True)
    return statistics.median(vals), vals


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--only", default="", help="comma-separated program names (default: all)")
    ap.add_argument("--scale24", action="store_true", help="also numa-PageRank at scale 24, once")
    args = ap.parse_args()
    only = set(x for x in args.only.split(",") if x)

    def want(name):
        return not only or name in only
    missing = [a for a in ("numa-PageRank", "numa-SPMV", "numa-BP", "numa-BFS", "numa-Components",
                           "numa-BellmanFord") if not os.path.exists(os.path.join(BIN, a))]
    if missing:
        print("reference programs not built (%s): run make -C oracle ref" % ", ".join(missing))
        return 1
    cores = os.cpu_count()
    res = {"host_cpus": cores, "runs": args.runs, "results": []}

    def add(name, graph, med, vals, work):
        r = dict(program=name, graph=graph, seconds_median=med, seconds=vals)
        r.update(work)
        res["results"].append(r)

    with tempfile.TemporaryDirectory(prefix="polymer_ref_time_") as td:
        # PR1: n = 2^16, m = 2^20, PageRank 20 iterations
        g = O.rmat(16, 1 << 20, seed=1)
        p = os.path.join(td, "pr1.adj")
        O.write_adj(p, g)
        print("PR1: n %d m %d" % (g.n, g.m), flush=True)
        if want("numa-PageRank"):
            med, vals = run_timed([os.path.join(BIN, "numa-PageRank"), p, "20", "-1"], "PageRank",
                                  args.runs)
            add("numa-PageRank", "PR1", med, vals, dict(iterations=20, edges_per_iteration=g.m))
        # scale 22: n = 2^22, m = 2^26
        g = O.rmat(22, 1 << 26, seed=1)
        w = O.weights_i32(g, 7)
        gs = g.symmetrized()
        ws = O.weights_i32(gs, 7)
        paths = dict(adj=os.path.join(td, "s22.adj"), wadj=os.path.join(td, "s22.wadj"),
                     sym=os.path.join(td, "s22.sym"), wsym=os.path.join(td, "s22.wsym"))
        O.write_adj(paths["adj"], g)
        O.write_adj(paths["wadj"], g, w)
        O.write_adj(paths["sym"], gs)
        O.write_adj(paths["wsym"], gs, ws)
        src = int(np.argmax(np.diff(gs.off)))  # the traversal workload's first source
        print("S22: n %d m %d, symmetric %d entries, source %d" % (g.n, g.m, gs.m, src), flush=True)
        if want("numa-PageRank"):
            med, vals = run_timed([os.path.join(BIN, "numa-PageRank"), paths["adj"], "20", "-1"],
                                  "PageRank", args.runs)
            add("numa-PageRank", "S22", med, vals, dict(iterations=20, edges_per_iteration=g.m))
        if want("numa-SPMV"):
            med, vals = run_timed([os.path.join(BIN, "numa-SPMV"), paths["wadj"], "20"], "SPMV", args.runs)
            add("numa-SPMV", "S22 integer weights", med, vals, dict(iterations=20))
        if want("numa-BP"):
            med, vals = run_timed([os.path.join(BIN, "numa-BP"), paths["adj"], "10"], "BeliefPropagation",
                                  args.runs)
            add("numa-BP", "S22", med, vals, dict(iterations=10))
        if want("numa-BFS"):
            med, vals = run_timed([os.path.join(BIN, "numa-BFS"), paths["sym"], str(src), "-", "-s"], "BFS",
                                  args.runs)
            add("numa-BFS", "S22 symmetric", med, vals, dict(source=src))
        if want("numa-Components"):
            med, vals = run_timed([os.path.join(BIN, "numa-Components"), paths["sym"], "-s"], "Components",
                                  args.runs)
            add("numa-Components", "S22 symmetric", med, vals, {})
        if want("numa-BellmanFord"):
            med, vals = run_timed([os.path.join(BIN, "numa-BellmanFord"), paths["wsym"], str(src), "-",
                                   "-s"], "BellmanFord", args.runs)
            add("numa-BellmanFord", "S22 symmetric, integer weights", med, vals, dict(source=src))
        if args.scale24:
            for f in paths.values():
                os.remove(f)
            del g, w, gs, ws
            g = O.rmat(24, 1 << 28, seed=1)
            p24 = os.path.join(td, "s24.adj")
            O.write_adj(p24, g)
            print("S24: n %d m %d" % (g.n, g.m), flush=True)
            med, vals = run_timed([os.path.join(BIN, "numa-PageRank"), p24, "20", "-1"], "PageRank", 1)
            add("numa-PageRank", "S24", med, vals, dict(iterations=20, edges_per_iteration=g.m))
    print(json.dumps(res, indent=1))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
