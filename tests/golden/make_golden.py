"""Writes the golden fixtures under tests/golden/ from the reference's own
programs.

Run in the build container (the reference sources exist only there):
    python tests/golden/make_golden.py
It builds oracle/_ref/ with oracle/Makefile (the reference Makefile's default
rule), writes PBBS text graphs produced by the oracle's deterministic rMat
generator into a temporary directory, runs the reference programs on them
and stores their printed results as small .npz/.json fixtures.

Fixtures (inputs are regenerated by the oracle from the recorded parameters;
the fixture stores a digest of the graph so generator drift is caught):
(numa-Components is not among them: built with the reference's default rule
it finishes its rounds, prints "Components : <secs>" and then aborts with
"The futex facility returned an unexpected error code" before its -result
loop prints anything; connected-component labels are pinned by union-find.)
  pr1_pagerank.npz   numa-PageRank g.adj 20 -1 -result   (rank per vertex)
  sym_bfs.json       numa-BFS g.sym <src> -result -s     (reached count, rounds; 3 sources)
  spmv_int.npz       numa-SPMV g.wadj 10 -result         (integer weights, x0=1/n)
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref")

PR1 = dict(scale=16, m=1 << 20, a=0.57, b=0.19, c=0.19, seed=1)
WSEED = 7


def graph_digest(g, w=None):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(g.off).tobytes())
    h.update(np.ascontiguousarray(g.adj).tobytes())
    if w is not None:
        h.update(np.ascontiguousarray(w).tobytes())
    return h.hexdigest()


def run(args, timeout=600, tolerate_exit_abort=False):
    """Runs a reference program with line-buffered stdout.  numa-BFS (like
    numa-Components) aborts while exiting -- its sub-worker threads still wait
    on a barrier that lives on a returned worker's stack -- after it has
    printed its round counts; tolerate_exit_abort keeps that output."""
    p = subprocess.run(["stdbuf", "-oL"] + args, capture_output=True, text=True,
                       timeout=timeout)
    if p.returncode != 0 and not (tolerate_exit_abort and p.returncode == -6):
        raise RuntimeError("%s failed (%d): %s" % (args, p.returncode, p.stderr[-2000:]))
    return p.stdout


def parse_tab_values(out, n, dtype):
    vals = np.full(n, np.nan if dtype == np.float64 else -1, dtype=dtype)
    seen = 0
    for line in out.splitlines():
        parts = line.split("\t")
        if len(parts) == 2 and parts[0].isdigit():
            vals[int(parts[0])] = dtype(float(parts[1])) if dtype == np.float64 else dtype(parts[1])
            seen += 1
    if seen != n:
        raise RuntimeError("expected %d values, parsed %d" % (n, seen))
    return vals


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True,
                   capture_output=True)
    g = O.rmat(PR1["scale"], PR1["m"], PR1["a"], PR1["b"], PR1["c"], PR1["seed"])
    gs = g.symmetrized()
    deg = gs.out_degree()
    bfs_src = int(np.argmax(deg))
    with tempfile.TemporaryDirectory() as td:
        fadj = os.path.join(td, "pr1.adj")
        fsym = os.path.join(td, "pr1.sym")
        fw = os.path.join(td, "pr1.wadj")
        O.write_adj(fadj, g)
        O.write_adj(fsym, gs)
        wint = O.weights_i32(g, WSEED)
        O.write_adj(fw, g, wint)

        out = run([os.path.join(REF_BIN, "numa-PageRank"), fadj, "20", "-1", "-result"])
        ranks = parse_tab_values(out, g.n, np.float64)
        np.savez_compressed(os.path.join(HERE, "pr1_pagerank.npz"), rank=ranks,
                            digest=graph_digest(g), iters=20, damping=0.85,
                            **{k: v for k, v in PR1.items()})

        runs = []
        for src in sorted({bfs_src, 1, 12345, 65535}):
            out = run([os.path.join(REF_BIN, "numa-BFS"), fsym, str(src), "-result", "-s"],
                      tolerate_exit_abort=True)
            # "Vertices visited" sums |frontier| over rounds = vertices reached;
            # "Finished in k" counts rounds including the final empty one.
            vsum = int(re.search(r"Vertices visited = (\d+)", out).group(1))
            rounds = int(re.search(r"Finished in (\d+) iterations", out).group(1))
            runs.append(dict(src=src, vertices_visited=vsum, finished_in=rounds))
        with open(os.path.join(HERE, "sym_bfs.json"), "w") as f:
            json.dump(dict(PR1, runs=runs, digest=graph_digest(gs)), f, indent=1)

        out = run([os.path.join(REF_BIN, "numa-SPMV"), fw, "10", "-result"])
        y = parse_tab_values(out, g.n, np.float64)
        np.savez_compressed(os.path.join(HERE, "spmv_int.npz"), y=y, iters=10,
                            weight_seed=WSEED, digest=graph_digest(g, wint),
                            **{k: v for k, v in PR1.items()})
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
