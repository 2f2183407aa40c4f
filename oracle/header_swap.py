"""Builds the reference's own programs against include/polymer.h (the header
swap): each numa-*.C is copied from the reference tree into a scratch directory
under /tmp together with the reference helper headers it includes itself
(gettime.h, parallel.h, custom-barrier.h) and this repository's polymer.h,
polymer-wgh.h and polymer_gpu.h, then compiled with hipcc --hipstdpar (so the
programs' functors compile for the GPU unannotated) and linked with
libpolymer_gpu.so.  The copies never enter the repository; binaries go to
oracle/_ref/swap/ (git-ignored).

    python oracle/header_swap.py            # the six programs (and the asynchronous BFS)
    python oracle/header_swap.py --check    # compile only, as tests/test_header_swap.py

Reroutes: numa-PageRank.C, numa-BFS.C and numa-BP.C define their own edge loops
(edgeMapDenseForwardOTHER, edgeMapNoRep/edgeMapDenseNoRep, edgeMapDenseBPNoRep)
that walk the graph's neighbour lists on the host; those programs compile
unchanged, and to run on the GPU their call sites are pointed at the header's
edgeMapDenseForward, edgeMap and edgeMapDenseBP.  numa-BP.C also takes the array
its functor reads (vertex potentials) from plain malloc, which the GPU cannot
address: it is taken from numa_alloc_local instead.  numa-BFS-async-pipe.C
hard-codes eight NUMA nodes: its commented-out libnuma query is restored.
"""
import argparse
import os
import re
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.environ.get("POLYMER_REFERENCE", "/root/reference")
OUT = os.path.join(HERE, "_ref", "swap")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

APPS = ["numa-SPMV", "numa-Components", "numa-BellmanFord", "numa-PageRank", "numa-BFS",
        "numa-BP", "numa-BFS-async-pipe"]
HELPERS = ["gettime.h", "parallel.h", "custom-barrier.h"]
OURS = ["polymer.h", "polymer-wgh.h", "polymer_gpu.h"]

# (pattern, replacement) applied to the program's text to run it on the GPU
REROUTES = {
    "numa-PageRank": [(r"edgeMapDenseForwardOTHER\(GA, Frontier,", "edgeMapDenseForward(GA, Frontier,")],
    "numa-BFS": [(r"\bedgeMapNoRep\(GA, Frontier,", "edgeMap(GA, Frontier,")],
    "numa-BP": [(r"edgeMapDenseBPNoRep\(GA, Frontier,(.*),output,true,subworker\);",
                 r"edgeMapDenseBP(GA, Frontier,\1,output,true);"),
                # the vertex potentials its functor reads on the GPU come from plain
                # malloc (numa-BP.C:557), which the GPU cannot address
                (r"\(VertexInfo \*\)malloc\(sizeof\(VertexInfo\) \* GA\.n\)",
                 "(VertexInfo *)numa_alloc_local(sizeof(VertexInfo) * GA.n)")],
    # it hard-codes eight NUMA nodes (numa-BFS-async-pipe.C:340, with the call that
    # asks libnuma commented out beside it): the call is restored
    "numa-BFS-async-pipe": [(r"numOfNode = 8;//numa_num_configured_nodes\(\);",
                             "numOfNode = numa_num_configured_nodes();")],
}

FLAGS = ["--offload-arch=gfx950", "-O2", "-std=c++17", "--hipstdpar", "-w", "-x", "hip"]


def available():
    return all(os.path.exists(os.path.join(REF, a + ".C")) for a in APPS)


def prepare(tmp, app, reroute):
    """Scratch copy of one program beside the swapped headers; returns its path."""
    for h in HELPERS:
        shutil.copy(os.path.join(REF, h), tmp)
    for h in OURS:
        shutil.copy(os.path.join(ROOT, "include", h), tmp)
    text = open(os.path.join(REF, app + ".C")).read()
    if reroute:
        for pat, rep in REROUTES.get(app, []):
            text, k = re.subn(pat, rep, text)
            if k != 1:
                raise RuntimeError("%s: reroute %r matched %d times" % (app, pat, k))
    path = os.path.join(tmp, app + ".C")
    with open(path, "w") as f:
        f.write(text)
    return path


def compile_one(app, reroute, link):
    """Compiles (and links, into oracle/_ref/swap/) one program; returns (app, rc, log)."""
    with tempfile.TemporaryDirectory(prefix="polymer_swap_") as tmp:
        src = prepare(tmp, app, reroute)
        obj = os.path.join(tmp, app + ".o")
        cmd = [HIPCC] + FLAGS + ["-I", tmp, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0 or not link:
            return app, r.returncode, r.stderr[-4000:]
        os.makedirs(OUT, exist_ok=True)
        exe = os.path.join(OUT, app)
        cmd = [HIPCC, "--offload-arch=gfx950", "--hipstdpar", obj, "-o", exe,
               "-L" + os.path.join(ROOT, "polymer_amd"), "-lpolymer_gpu",
               "-Wl,-rpath,$ORIGIN/../../../polymer_amd", "-lnuma", "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        return app, r.returncode, r.stderr[-4000:]


def build_all(reroute=True, link=True, jobs=6):
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(lambda a: compile_one(a, reroute, link), APPS))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true", help="compile only, no reroutes")
    args = ap.parse_args()
    if not available():
        print("reference sources not found under %s: nothing to build" % REF)
        return 0
    res = build_all(reroute=not args.check, link=not args.check)
    bad = [(a, log) for a, rc, log in res if rc != 0]
    for a, log in bad:
        print("== %s failed\n%s" % (a, log), file=sys.stderr)
    print("built %d of %d%s" % (len(res) - len(bad), len(res),
                                "" if args.check else " into " + OUT))
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
