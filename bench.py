"""Benchmark: PageRank GTEPS on the scale-24 rMat graph (n=2^24, m=2^28) on one
MI355X, 20 iterations per step (numa-PageRank.C semantics, fp32).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line with the metric, the roofline of the dominant kernel
(the pull iteration, measured with HIP events on the library stream) and a
CPU baseline (the oracle's fp64 OpenMP PageRank on a bounded sample of the
same workload: 3 of the 20 iterations on the same graph).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pagerank_bytes(n, edges, iters):
    """Algorithmic bytes of the pull iterations: 8 per active out-edge (4 B
    neighbour id + 4 B gathered contribution) and 20 per vertex (offsets,
    rank write, contribution write, frontier read/write, out-degree)."""
    return 8.0 * edges + 20.0 * n * iters


def cpu_baseline(pa, g, iters_sample=20):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc

    # the host cores this process may use, at most the 16 a one-GPU box shares out
    # (OMP_NUM_THREADS there)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    off, adj, _ = g.download()
    hg = orc.Graph(g.n, off, adj)
    hg.transpose()
    _, st = g.pagerank(iters_sample, 0.85, out=False)  # exact sum of E_t for the sample
    t0 = time.perf_counter()
    ref = orc.pagerank_pull_omp(hg, iters_sample, 0.85, threads=threads)
    dt = time.perf_counter() - t0
    out = {
        "value": round(st.edges / dt / 1e9, 4),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": "%d of the 20 PageRank iterations (fp64 pull, OpenMP) on the same "
                  "scale-24 graph, %.1f s on %d threads (%.0f thread-seconds)" %
                  (iters_sample, dt, threads, dt * threads),
    }
    if iters_sample == 20:
        # the measured solve's fp32 ranks against these fp64 ones (outside any timing)
        rank, _ = g.pagerank(20, 0.85)
        out["gpu_rank_rel_l1"] = float(np.abs(rank.astype(np.float64) - ref).sum() /
                                       np.abs(ref).sum())
    return out


def traversal_sources(deg, k=8, seed=7):
    """Sources of the scale-22 traversal workload: the vertex of maximum degree,
    then k distinct seeded random vertices with at least one edge (an isolated
    draw is skipped, as Graph500 does)."""
    top = int(np.argmax(deg))
    rng = np.random.default_rng(seed)
    out = [top]
    while len(out) < k + 1:
        v = int(rng.integers(0, deg.size))
        if deg[v] > 0 and v not in out:
            out.append(v)
    return out


def bfs_extra(pa, scale=22, m=1 << 26, nsrc=8):
    gs = pa.Graph.rmat(scale, m, seed=1, symmetrize=True)
    off, _, _ = gs.download()
    srcs = traversal_sources(np.diff(off), nsrc)
    gs.bfs(0, out=False)  # warm-up / workspace
    tot_ms, tot_edges, tot_bytes = 0.0, 0, 0
    for s in srcs:
        _, _, st = gs.bfs(s, out=False)
        tot_ms += st.ms
        tot_edges += st.edges  # directed entries of the reached vertices
        tot_bytes += 4 * st.edges + 16 * st.active  # 4 m_reached + 16 n_reached
    gbs = tot_bytes / (tot_ms / 1e3) / 1e9
    return {"bfs_scale22_gteps_graph500": round(tot_edges / 2 / (tot_ms / 1e3) / 1e9, 3),
            "bfs_scale22_gteps_directed": round(tot_edges / (tot_ms / 1e3) / 1e9, 3),
            "bfs_scale22_algorithmic_gbs": round(gbs, 1),
            "bfs_scale22_hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "bfs_sources": len(srcs)}


def bfs24_runs(gs, nsrc=8):
    """BFS from vertex 0 and seeded random sources with at least one edge
    (Graph500's rule: an isolated draw is skipped); returns the stats of each."""
    rng = np.random.default_rng(9)
    gs.bfs(0, out=False)
    runs, draws, s = [], 0, 0
    while len(runs) < nsrc and draws < 20 * nsrc:
        _, _, st = gs.bfs(s, out=False)
        if st.active > 1:
            runs.append((s, st))
        s = int(rng.integers(0, gs.n))
        draws += 1
    return runs


def bfs24_extra(pa, scale=24, m=1 << 28, nsrc=8):
    """BFS on the symmetrised scale-24 rMat (2^29 directed entries), direction-
    optimising, from the sources of bfs24_runs."""
    gs = pa.Graph.rmat(scale, m, seed=1, symmetrize=True)
    tot_ms, tot_edges, tot_bytes, used = 0.0, 0, 0, 0
    for _, st in bfs24_runs(gs, nsrc):
        tot_ms += st.ms
        tot_edges += st.edges
        tot_bytes += 4 * st.edges + 16 * st.active
        used += 1
    gs.close()
    gbs = tot_bytes / (tot_ms / 1e3) / 1e9
    return {"bfs_scale24_gteps_graph500": round(tot_edges / 2 / (tot_ms / 1e3) / 1e9, 2),
            "bfs_scale24_ms_per_source": round(tot_ms / used, 3),
            "bfs_scale24_algorithmic_gbs": round(gbs, 1),
            "bfs_scale24_hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "bfs_scale24_sources": used}


def async_extra(pa, scale=24, m=1 << 28):
    """The asynchronous traversal (numa-BFS-async-pipe.C) on the symmetrised
    scale-24 rMat from the sources of bfs24_runs: time per source and reached
    edges per second (its result is a spanning tree of the reachable set, so the
    comparison with the breadth-first search is per reached edge)."""
    gs = pa.Graph.rmat(scale, m, seed=1, symmetrize=True)
    srcs = [s for s, _ in bfs24_runs(gs, 4)]
    gs.bfs_async(srcs[0], out=False)
    tot_ms, tot_edges = 0.0, 0
    for s in srcs:
        _, vis, st = gs.bfs_async(s, out=False)
        _, _, stb = gs.bfs(s, out=False)
        tot_ms += st.ms
        tot_edges += stb.edges
    gs.close()
    return {"async_scale24_ms_per_source": round(tot_ms / len(srcs), 3),
            "async_scale24_gteps_graph500": round(tot_edges / 2 / (tot_ms / 1e3) / 1e9, 2)}


def traversal_extra(pa, scale=22, m=1 << 26):
    """Connected components on the symmetrised scale-22 graph; Bellman-Ford (int32
    weights U[1,255]) from the sources of traversal_sources."""
    gs = pa.Graph.rmat(scale, m, seed=1, symmetrize=True, weights="i32", wseed=7)
    gs.cc(out=False)
    _, stc = gs.cc(out=False)
    off, _, _ = gs.download()
    srcs = traversal_sources(np.diff(off))
    gs.sssp(srcs[0], out=False)
    tot_ms, tot_edges = 0.0, 0
    for s in srcs:
        _, st = gs.sssp(s, out=False)
        tot_ms += st.ms
        tot_edges += st.edges
    return {"cc_scale22_ms": round(stc.ms, 3), "cc_rounds": stc.rounds,
            "sssp_scale22_ms_per_source": round(tot_ms / len(srcs), 3),
            "sssp_scale22_gteps_relaxed": round(tot_edges / (tot_ms / 1e3) / 1e9, 2)}


def multisource_extra(pa, scale=22, m=1 << 26, k=64):
    """64 BFS and 64 Bellman-Ford sources on the symmetrised scale-22 graph in one
    batched solve each (a 64-bit source mask per vertex), against the same 64
    sources run one by one."""
    gs = pa.Graph.rmat(scale, m, seed=1, symmetrize=True, weights="i32", wseed=7)
    rng = np.random.default_rng(13)
    src = rng.choice(gs.n, k, replace=False)
    src[0] = 0
    out = {}
    gs.msbfs(src, out=False)
    _, st = gs.msbfs(src, out=False)
    one = 0.0
    for s in src:
        _, _, s1 = gs.bfs(int(s), out=False)
        one += s1.ms
    out["msbfs64_scale22_ms"] = round(st.ms, 3)
    out["bfs64_one_by_one_ms"] = round(one, 3)
    out["msbfs64_speedup"] = round(one / st.ms, 2)
    gs.mssssp(src, out=False)
    _, st = gs.mssssp(src, out=False)
    one = 0.0
    for s in src:
        _, s1 = gs.sssp(int(s), out=False)
        one += s1.ms
    out["mssssp64_scale22_ms"] = round(st.ms, 3)
    out["sssp64_one_by_one_ms"] = round(one, 3)
    out["mssssp64_speedup"] = round(one / st.ms, 2)
    gs.close()
    return out


def dropin_extra(pa):
    """The drop-in path: the PageRank and BFS drivers written against
    include/polymer.h (the reference's functors and iteration structure: fp64
    values, atomic pushes), on the same scale-24 graphs; their loop times."""
    import subprocess
    out = {}
    bindir = os.path.join(ROOT, "polymer_amd", "apps", "bin")
    runs = (("dropin_pagerank_scale24_ms", ["pagerank", "-rmat", "24", str(1 << 28), "1",
                                            "-rounds", "20"]),
            ("dropin_bfs_scale24_ms", ["bfs", "-rmat", "24", str(1 << 28), "1", "-s", "-r", "0"]))
    for key, cmd in runs:
        # a search is a fraction of a millisecond: the BFS driver runs three times
        # (three processes) and the median is reported beside all three
        reps = 3 if cmd[0] == "bfs" else 1
        times = []
        for _ in range(reps):
            r = subprocess.run([os.path.join(bindir, cmd[0])] + cmd[1:], capture_output=True,
                               text=True, timeout=300, check=True)
            for line in r.stdout.splitlines():
                if line.startswith("time_ms "):
                    times.append(round(float(line.split()[1]), 3))
        if times:
            out[key] = sorted(times)[len(times) // 2]
            if reps > 1:
                out[key.replace("_ms", "_runs_ms")] = times
    return out


def value_extra(pa, g, scale=24, m=1 << 28):
    """SpMV (20 iterations, fp32 values U(0,1], x0 U[0,1)) and BP (10 iterations,
    potentials U[0.1,1]) on the scale-24 graph, with their algorithmic bandwidth
    (SpMV 12m + 12n, BP 20 E_t + 28n bytes per iteration)."""
    out = {}
    gw = pa.Graph.rmat(scale, m, seed=1, weights="f32", wseed=3)
    x0 = np.random.default_rng(3).random(gw.n).astype(np.float32)
    gw.spmv(x0, 20, out=False)
    _, st = gw.spmv(x0, 20, out=False)
    out["spmv_scale24_ms"] = round(st.ms, 3)
    gbs = (12 * gw.m + 12 * gw.n) * 20 / (st.ms / 1e3) / 1e9
    out["spmv_scale24_algorithmic_gbs"] = round(gbs, 1)
    out["spmv_scale24_hbm_frac"] = round(gbs / HBM_PEAK_GBS, 4)
    # the adjoint of the same 20 iterations (loss g . y_20, g ~ N(0,1)): the forward
    # keeps its iterates; the backward is 20 tiled pulls over the CSR (12m + 12n bytes
    # each, as the forward's) and one gradient pass (per edge its index, source,
    # fp64 gradient and the destination's 20 backward values, 16 + 4*20 bytes; the
    # iterates of both directions transposed into 128-byte rows, 2 * (4*20 + 128) n)
    gv = np.random.default_rng(4).standard_normal(gw.n).astype(np.float32)
    gw.spmv(x0, 20, out=False, save=True)
    _, stf = gw.spmv(x0, 20, out=False, save=True)
    gw.spmv_adjoint(gv, out=False)
    _, _, sta = gw.spmv_adjoint(gv, out=False)
    out["spmv_scale24_forward_saved_ms"] = round(stf.ms, 3)
    out["spmv_scale24_adjoint_ms"] = round(sta.ms, 3)
    out["spmv_adjoint_over_forward"] = round(sta.ms / st.ms, 3)
    adj_bytes = 20 * (12 * gw.m + 12 * gw.n) + gw.m * (16 + 4 * 20) + 2 * (4 * 20 + 128) * gw.n
    out["spmv_adjoint_algorithmic_gbs"] = round(adj_bytes / (sta.ms / 1e3) / 1e9, 1)
    gw.close()
    g.bp_potentials(9)
    vpot = np.random.default_rng(8).uniform(0.1, 1.0, 2 * g.n).astype(np.float32)
    prod0 = np.ones(2 * g.n, np.float32)
    # the first call also prepares the graph once (the iteration at which each source
    # leaves the frontier, as a byte per in-edge, and the light-destination chunks)
    t0 = time.perf_counter()
    g.bp(vpot, prod0, 10, out=False)
    t1 = time.perf_counter()
    _, st = g.bp(vpot, prod0, 10, out=False)
    t2 = time.perf_counter()
    out["bp_scale24_ms"] = round(st.ms, 3)
    out["bp_scale24_prepare_ms"] = round(((t1 - t0) - (t2 - t1)) * 1e3, 1)
    out["bp_scale24_algorithmic_gbs"] = round((20 * st.edges + 28 * g.n * 10) / (st.ms / 1e3) / 1e9, 1)
    return out


def delta_extra(pa, g, iters=20):
    """PageRankDelta (fp64, epsilon2 0.01, 20 rounds) on the scale-24 PageRank graph:
    time, frontier edges per second (the edges the reference pushes), and the
    algorithmic bandwidth of the reference's formulation: 20 bytes per frontier
    edge (the 4-byte id and an 8-byte read-modify-write of the destination's sum)
    plus 40 bytes per vertex and round (the filter reads and writes p and the sum
    and writes the next contribution)."""
    g.pagerank_delta(iters, 0.85, 0.01, out=False)
    _, st = g.pagerank_delta(iters, 0.85, 0.01, out=False)
    gbs = (20 * st.edges + 40 * g.n * iters) / (st.ms / 1e3) / 1e9
    return {"pagerank_delta_scale24_ms": round(st.ms, 3),
            "pagerank_delta_frontier_gteps": round(st.edges / (st.ms / 1e3) / 1e9, 2),
            "pagerank_delta_frontier_mean": round(st.active / max(iters, 1)),
            "pagerank_delta_pull_rounds": st.dense_rounds,
            "pagerank_delta_algorithmic_gbs": round(gbs, 1),
            "pagerank_delta_hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def calibrate():
    """The counter calibration kernel of every profile run: torch's elementwise
    negation of a 1 GiB fp32 tensor into another (exactly 1 GiB read and 1 GiB
    written in one dispatch)."""
    import torch

    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    torch.cuda.synchronize()
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    log("calibration kernel (neg): 1073741824 bytes read, 1073741824 bytes written")
    del x, y


def profile_run(args):
    """Short run for rocprofv3: a plain streaming kernel of known size (torch's
    elementwise negation of a 1 GiB fp32 tensor into another: exactly 1 GiB
    read and 1 GiB written in one dispatch, the counter calibration), then one
    warm-up and one measured PageRank of --iters iterations."""
    import polymer_amd as pa

    calibrate()
    pa.init(0)
    g = pa.Graph.rmat(args.scale, args.edges, seed=1)
    sched = [int(x) for x in args.schedule.split(",")]
    g.set_schedule(*sched[:4])
    g.set_tiles(sched[4], sched[6], sched[7], max_degree=sched[5])
    log("pull schedule (ipl, block, multi, chunk, tiles, tile max degree, tile edges, cold from):",
        g.get_schedule())
    g.pagerank(args.iters, 0.85, out=False)
    _, st = g.pagerank(args.iters, 0.85, out=False, profile=True)
    log("pagerank %d iterations: %.3f ms, per-iteration pull avg %.1f us, edges %d" %
        (args.iters, st.ms, st.kernel_ms / max(st.kernel_launches, 1) * 1e3, st.edges))


def profile_bp(args):
    """Short run for rocprofv3: BP on the scale-24 rMat with the potentials of
    value_extra, 10 iterations with fp64 state twice (warm-up, measured), then the
    same with fp32 state; one k_bp_pull per iteration (the first call also runs the
    once-per-graph preparation: the frontier's exit iteration per edge and the light
    chunks).  The calibration kernel of profile_run runs first."""
    import polymer_amd as pa

    calibrate()
    pa.init(0)
    g = pa.Graph.rmat(args.scale, args.edges, seed=1)
    g.bp_potentials(9)
    vpot = np.random.default_rng(8).uniform(0.1, 1.0, 2 * g.n).astype(np.float32)
    prod0 = np.ones(2 * g.n, np.float32)
    for fp32 in (False, True):
        g.bp(vpot, prod0, 10, out=False, fp32=fp32)
        _, st = g.bp(vpot, prod0, 10, out=False, fp32=fp32)
        log("bp fp32=%d, 10 iterations: %.3f ms, %d edges, %.1f GB/s algorithmic" %
            (fp32, st.ms, st.edges, (20 * st.edges + 28 * g.n * 10) / (st.ms / 1e3) / 1e9))
    g.close()


def profile_bfs(args):
    """Short run for rocprofv3: the calibration kernel of profile_run, then the
    searches of bfs24_extra on the symmetrised scale-24 rMat (the source selection,
    whose isolated draws are searches too), then the eight sources once more: the
    last eight searches of the run are the measured ones."""
    import polymer_amd as pa

    calibrate()
    pa.init(0)
    gs = pa.Graph.rmat(args.scale, args.edges, seed=1, symmetrize=True)
    for s, _ in bfs24_runs(gs):
        _, _, st = gs.bfs(s, out=False)
        log("bfs source %d: %.3f ms, %d rounds (%d bottom-up), %d edges, %d reached" %
            (s, st.ms, st.rounds, st.dense_rounds, st.edges, st.active))


def profile_extra(args):
    """Short runs for rocprofv3 of the workloads in the JSON line's extras: the
    calibration kernel, then each solve once to warm up and once measured.
      spmv    SpMV on the scale-24 rMat, 20 iterations, fp32 values (value_extra)
      delta   PageRankDelta on the scale-24 rMat, 20 rounds (delta_extra)
      bfs22   BFS on the symmetrised scale-22 rMat from the nine traversal sources
              (bfs_extra)
      ppr     batched PPR, 64 sources, scale 22, 20 iterations, then two of the
              plain runs on the tiles (batched_extra)
      async   the asynchronous traversal from the first two scale-24 BFS sources
              (async_extra)
      spmvadj the SpMV adjoint of 20 iterations on the scale-24 rMat, after the
              forward that keeps its iterates (value_extra)
    (The drop-in drivers are profiled directly: rocprofv3 ... --
    polymer_amd/apps/bin/pagerank -rmat 24 268435456 1 -rounds 20.)"""
    import polymer_amd as pa

    calibrate()
    name = args.profile_extra
    pa.init(0)
    if name == "spmv":
        gw = pa.Graph.rmat(24, 1 << 28, seed=1, weights="f32", wseed=3)
        x0 = np.random.default_rng(3).random(gw.n).astype(np.float32)
        for r in range(2):
            _, st = gw.spmv(x0, 20, out=False)
            log("spmv run %d: %.3f ms" % (r, st.ms))
    elif name == "delta":
        g = pa.Graph.rmat(24, 1 << 28, seed=1)
        for r in range(2):
            _, st = g.pagerank_delta(20, 0.85, 0.01, out=False)
            log("pagerank_delta run %d: %.3f ms, %d frontier edges, %d pull rounds" %
                (r, st.ms, st.edges, st.dense_rounds))
    elif name == "bfs22":
        gs = pa.Graph.rmat(22, 1 << 26, seed=1, symmetrize=True)
        off, _, _ = gs.download()
        srcs = traversal_sources(np.diff(off))
        gs.bfs(0, out=False)
        for s in srcs:
            _, _, st = gs.bfs(s, out=False)
            log("bfs source %d: %.3f ms, %d rounds (%d bottom-up), %d edges, %d reached" %
                (s, st.ms, st.rounds, st.dense_rounds, st.edges, st.active))
    elif name == "ppr":
        g = pa.Graph.rmat(22, 1 << 26, seed=1)
        src = np.random.default_rng(11).choice(g.n, 64, replace=False)
        src[0] = 0
        for r in range(2):
            _, st = g.ppr_batched(src, 20, out=False)
            log("ppr batched run %d: %.3f ms" % (r, st.ms))
        for s in src[:2]:
            _, st = g.ppr(int(s), 20, out=False)
            log("ppr plain source %d: %.3f ms" % (s, st.ms))
    elif name == "spmvadj":
        gw = pa.Graph.rmat(24, 1 << 28, seed=1, weights="f32", wseed=3)
        x0 = np.random.default_rng(3).random(gw.n).astype(np.float32)
        gv = np.random.default_rng(4).standard_normal(gw.n).astype(np.float32)
        for r in range(2):
            _, stf = gw.spmv(x0, 20, out=False, save=True)
            _, _, sta = gw.spmv_adjoint(gv, out=False)
            log("spmv forward (kept iterates) run %d: %.3f ms, adjoint %.3f ms" % (r, stf.ms, sta.ms))
    elif name == "async":
        gs = pa.Graph.rmat(24, 1 << 28, seed=1, symmetrize=True)
        srcs = [s for s, _ in bfs24_runs(gs, 2)]
        for s in srcs:
            _, vis, st = gs.bfs_async(s, out=False)
            log("async source %d: %.3f ms, %d reached" % (s, st.ms, vis))
    else:
        raise SystemExit("unknown --profile-extra %r" % name)


def probe_gather(args):
    """Random 8-byte gathers per second against the size of the gathered array:
    torch.index_select of 2^27 uniform int32 indices from an fp64 array of 2^k
    elements (each gather also streams its 4-byte index in and 8 bytes out).  The
    ceiling of a pull whose functor gathers per edge from arrays larger than the
    last-level caches (the drop-in drivers' p_curr[s] and V[s])."""
    import torch

    dev = torch.device("cuda", 0)
    count = 1 << 27
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    for k in range(15, 28):
        a = torch.zeros(1 << k, dtype=torch.float64, device=dev)
        idx = torch.randint(0, 1 << k, (count,), generator=gen, device=dev, dtype=torch.int32)
        out = torch.empty(count, dtype=torch.float64, device=dev)
        torch.index_select(a, 0, idx, out=out)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        reps = 5
        for _ in range(reps):
            torch.index_select(a, 0, idx, out=out)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        print(json.dumps({"array_mb": round((8 << k) / 1e6, 2), "gathers": count,
                          "ms": round(ms, 3), "ggather_per_s": round(count / ms / 1e6, 1)}),
              flush=True)
        del a, idx, out


def measured_traffic(schedule):
    """HBM bytes per PageRank iteration from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary, when it was taken with the same pull schedule."""
    path = os.path.join(ROOT, "profiles", "pagerank_scale24_pmc.txt")
    if not os.path.exists(path):
        return None, "no pmc summary"
    want = ",".join(str(x) for x in schedule)
    cur, found = None, {}
    for line in open(path):
        line = line.strip()
        if line.startswith("[schedule ") and line.endswith("]"):
            cur = line[len("[schedule "):-1]
        elif line.startswith("hbm_bytes_per_iteration:") and cur:
            found[cur] = int(float(line.split(":", 1)[1]))
    if want not in found:
        return None, "pmc summary has schedules %s only" % sorted(found)
    return found[want], "profiles/pagerank_scale24_pmc.txt [schedule %s]" % want


def batched_extra(pa, scale=22, m=1 << 26, k=64, iters=20):
    """Personalized PageRank with k one-hot teleports: one batched pass (k-wide rows)
    against k plain single-vector runs on the same GPU, same tolerance."""
    g = pa.Graph.rmat(scale, m, seed=1)
    rng = np.random.default_rng(11)
    src = rng.choice(g.n, k, replace=False)
    src[0] = 0
    g.ppr_batched(src, iters, out=False)  # workspace / warm-up
    _, stb = g.ppr_batched(src, iters, out=False)
    out = {"ppr64_scale22_batched_ms": round(stb.ms, 3)}
    # the plain runs on the PageRank tiles (the fastest single-vector pull), then on
    # the degree-binned pull alone for comparison with earlier measurements
    for label, tiles in (("", 1), ("_untiled", 0)):
        g.set_tiles(tiles)
        g.ppr(int(src[0]), iters, out=False)
        plain_ms = 0.0
        for s in src:
            _, st = g.ppr(int(s), iters, out=False)
            plain_ms += st.ms
        out["ppr64_scale22_plain%s_total_ms" % label] = round(plain_ms, 3)
        out["ppr64_batched_speedup%s" % label] = round(plain_ms / stb.ms, 2)
    g.set_tiles(1)
    return out


def adjoint_extra(pa, orc_free=True, scale=22, m=1 << 26):
    """Weighted PageRank forward solve vs adjoint (fixed point + gradient pass), both
    iterated to the same relative tolerance."""
    g = pa.Graph.rmat(scale, m, seed=2, weights="f32", wseed=4)
    off, adj, w = g.download()
    src = np.repeat(np.arange(g.n), np.diff(off))
    tot = np.bincount(src, weights=w.astype(np.float64), minlength=g.n)
    wn = (w / tot[src]).astype(np.float32)
    tele = np.full(g.n, 1.0 / g.n, np.float32)
    gv = np.random.default_rng(0).standard_normal(g.n).astype(np.float32)
    # the same relative tolerance for both fixed points: the L1 change of an
    # iteration below 1e-7 of the iterate's scale (|x|_1 = 1; |lambda| scales with |g|_1)
    rel = 1e-7
    tol_f, tol_a = rel * 1.0, rel * float(np.abs(gv).sum())
    g.wpr_forward(wn, tele, 0.85, 500, tol_f, out=False)
    _, stf = g.wpr_forward(wn, tele, 0.85, 500, tol_f, out=False)
    g.wpr_adjoint(gv, 500, tol_a, out=False)
    _, _, _, sta = g.wpr_adjoint(gv, 500, tol_a, out=False)
    return {"wpr_scale22_forward_ms": round(stf.ms, 3), "wpr_forward_iters": stf.rounds,
            "wpr_scale22_adjoint_ms": round(sta.ms, 3), "wpr_adjoint_iters": sta.rounds,
            "wpr_relative_tolerance": rel, "wpr_forward_tol": tol_f, "wpr_adjoint_tol": tol_a,
            "adjoint_over_forward": round(sta.ms / stf.ms, 3)}


def multigraph_inputs(pa, count=512):
    """The many-graph workload: count independent rMat graphs, n = 2^s with s
    uniform in [8, 16] (log-uniform n in [2^8, 2^16]), average degree 16, seeds
    1000 + i; returns vertex bases and the concatenated CSR in input ids."""
    rng = np.random.default_rng(5)
    offs, adjs, vb, base, ebase = [], [], [0], 0, 0
    for i in range(count):
        s = int(rng.integers(8, 17))
        h = pa.Graph.rmat(s, 16 << s, seed=1000 + i)
        off, adj, _ = h.download()
        h.close()
        offs.append(off[:-1] + ebase)
        adjs.append(adj + base)
        base += h.n
        ebase += h.m
        vb.append(base)
    return (np.array(vb, np.int64), np.concatenate(offs + [np.array([ebase])]).astype(np.int64),
            np.concatenate(adjs).astype(np.int32))


def multigraph_extra(pa, count=512, iters=20):
    """512 independent rMat graphs (multigraph_inputs), PageRank of every graph in
    one captured launch."""
    vb, off, adj = multigraph_inputs(pa, count)
    g = pa.Graph.multi(vb, off, adj)
    g.autotune()
    g.pagerank_multigraph(iters, out=False)
    _, st = g.pagerank_multigraph(iters, out=False)
    return {"multigraph_512_n": g.n, "multigraph_512_m": g.m,
            "multigraph_512_pagerank_ms": round(st.ms, 3),
            "multigraph_512_gteps": round(g.m * iters / (st.ms / 1e3) / 1e9, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--profile", action="store_true",
                    help="short run to be wrapped by rocprofv3 (no JSON line)")
    ap.add_argument("--profile-bfs", action="store_true",
                    help="the same for the scale-24 BFS sources of the extras")
    ap.add_argument("--profile-bp", action="store_true",
                    help="the same for BP on the scale-24 rMat of the extras")
    ap.add_argument("--profile-extra", default=None,
                    help="the same for one extra workload: spmv, delta, bfs22, ppr, async, spmvadj")
    ap.add_argument("--probe-gather", action="store_true",
                    help="random 8-byte gather rate against the gathered array's size")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edges", type=int, default=1 << 28)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--schedule", default="16,16384,65536,16384,1,1073741824,524288,2097152",
                    help="--profile only: ipl,block,multi,chunk,tiles,tile_max_degree,"
                         "tile_edges,cold_from")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.profile:
        profile_run(args)
        return
    if args.profile_bfs:
        profile_bfs(args)
        return
    if args.profile_bp:
        profile_bp(args)
        return
    if args.profile_extra:
        profile_extra(args)
        return
    if args.probe_gather:
        probe_gather(args)
        return

    import polymer_amd as pa

    pa.init(0)
    log("device:", pa.device_name())
    t0 = time.perf_counter()
    g = pa.Graph.rmat(args.scale, args.edges, seed=1)
    log("graph scale %d m %d built in %.1f s" % (args.scale, g.m, time.perf_counter() - t0))
    g.autotune()
    schedule = g.get_schedule()
    log("autotuned pull schedule (ipl, block, multi, chunk, tiles, tile edges):", schedule)

    for _ in range(args.warmup):
        g.pagerank(args.iters, 0.85, out=False)
    # timed region: K steps, each a full 20-iteration PageRank with inputs resident in HBM
    edges = 0
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(args.steps):
        _, st = g.pagerank(args.iters, 0.85, out=False)
        edges += st.edges
        dev_ms += st.ms
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    gteps = edges / wall / 1e9
    log("pagerank: %.3f ms/step wall, %.3f ms/step device, %.2f GTEPS" %
        (ms_per_step, dev_ms / args.steps, gteps))

    # dominant kernel: the pull of one iteration (with tiles, the launches
    # k_cold_gather<PrOp> + k_pull<PrOp> hot segments + k_pull_tiles<PrOp> +
    # k_tile_finish<PrOp>), bracketed by HIP events on the library stream,
    # iteration by iteration
    _, stp = g.pagerank(args.iters, 0.85, out=False, profile=True)
    avg_kernel_s = stp.kernel_ms / max(stp.kernel_launches, 1) / 1e3
    bytes_per_launch = pagerank_bytes(g.n, stp.edges, args.iters) / args.iters
    achieved = bytes_per_launch / avg_kernel_s / 1e9
    traffic, traffic_src = measured_traffic(schedule)
    kernels = ("k_pull<PrOp> (hot) + k_pull_tiles<PrOp> + k_tile_finish<PrOp>" if schedule[4]
               else "k_pull<PrOp>")
    if schedule[4] and schedule[7] >= 0:
        kernels = "k_cold_gather<PrOp> + " + kernels
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": traffic_src,
                "kernel": "pull iteration: " + kernels, "avg_launch_us": round(avg_kernel_s * 1e6, 2),
                "bytes_per_launch": int(bytes_per_launch)}
    # the 20 bytes per vertex of the model count the vertices without in-edges, which
    # the pull visits in the first two iterations only (their rank cannot change
    # after that): the same frac over the bytes of the vertices it touches
    _, adj_in, _ = g.download()
    zero_in = int(np.count_nonzero(np.bincount(adj_in, minlength=g.n) == 0))
    del adj_in
    touched = bytes_per_launch - 20.0 * zero_in * max(args.iters - 2, 0) / args.iters
    roofline["zero_in_degree_vertices"] = zero_in
    roofline["frac_touched"] = round(touched / avg_kernel_s / 1e9 / HBM_PEAK_GBS, 4)
    log("pull iteration avg %.1f us, %.1f GB/s algorithmic" % (avg_kernel_s * 1e6, achieved))

    # the same solve with the ranks copied to host memory in input order (the
    # PCIe-inclusive rate; never the reported value)
    g.pagerank(args.iters, 0.85, out=True)
    t1 = time.perf_counter()
    g.pagerank(args.iters, 0.85, out=True)
    host_ms = (time.perf_counter() - t1) * 1e3
    extra = {"pagerank_step_with_host_copy_ms": round(host_ms, 3)}
    if not args.no_extra:
        for name, fn in (("bfs", bfs_extra), ("bfs24", bfs24_extra), ("async", async_extra),
                         ("traversal", traversal_extra),
                         ("multisource", multisource_extra),
                         ("value", lambda p: value_extra(p, g)),
                         ("delta", lambda p: delta_extra(p, g)), ("batched", batched_extra),
                         ("adjoint", adjoint_extra), ("multigraph", multigraph_extra),
                         ("dropin", dropin_extra)):
            t1 = time.perf_counter()
            try:
                extra.update(fn(pa))
            except Exception as e:  # report, do not hide
                extra[name + "_error"] = str(e)
            log("%s extra done in %.1f s" % (name, time.perf_counter() - t1))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pa, g)
        log("cpu baseline:", cpu)

    out = {
        "metric": "PageRank GTEPS (scale-24 rMat), % of HBM peak in roofline.frac",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: device-generated rMat a/b/c=0.57/0.19/0.19 seed 1, no self-loops "
                "or duplicates",
        "config": {"workload": "pagerank rmat scale-%d m=%d, %d iterations per step, "
                               "damping 0.85" % (args.scale, g.m, args.iters),
                   "n": g.n, "m": g.m, "iters": args.iters,
                   "schedule": ",".join(str(x) for x in schedule)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
