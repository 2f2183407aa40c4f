"""Pins the CPU oracle (oracle/polymer_oracle.c) before it is trusted as the
checker of the GPU library.

Pins, per routine:
  * golden vectors printed by the reference's own programs
    (tests/golden/make_golden.py): PageRank ranks, SpMV (integer weights),
    BFS reached counts and round counts;
  * independent formulations: scipy.sparse matrix iterations, scipy.csgraph
    (connected components, Dijkstra), dense numpy solves, brute force;
  * the survey's counter-example for the reference's Bellman-Ford (P6).
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def rel_l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


def csr_matrix(g, w=None):
    rows = np.repeat(np.arange(g.n), np.diff(g.off))
    data = np.ones(g.m) if w is None else np.asarray(w, np.float64)
    return sp.csr_matrix((data, (rows, g.adj)), shape=(g.n, g.n))


@pytest.fixture(scope="module")
def pr1(orc):
    return orc.rmat(16, 1 << 20, seed=1)


def digest(g, w=None):
    import hashlib
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(g.off).tobytes())
    h.update(np.ascontiguousarray(g.adj).tobytes())
    if w is not None:
        h.update(np.ascontiguousarray(w).tobytes())
    return h.hexdigest()


# ---- generator ------------------------------------------------------------------------
def test_rmat_properties(orc, pr1):
    g = pr1
    assert g.m == 1 << 20 and g.n == 1 << 16
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    assert not np.any(src == g.adj), "self loops"
    key = src.astype(np.int64) * g.n + g.adj
    assert np.all(np.diff(key) > 0), "duplicates or unsorted"
    g2 = orc.rmat(16, 1 << 20, seed=1)
    assert np.array_equal(g2.adj, g.adj)
    g3 = orc.rmat(16, 1 << 20, seed=2)
    assert not np.array_equal(g3.adj[:1000], g.adj[:1000])


def test_rmat_prefix_definition(orc):
    # the graph is the first m distinct non-loop pairs of the counter sequence:
    # a smaller m is a subset of a larger one with the same seed
    a = orc.rmat(10, 4000, seed=5)
    b = orc.rmat(10, 6000, seed=5)
    ka = set((np.repeat(np.arange(a.n), np.diff(a.off)) * a.n + a.adj).tolist())
    kb = set((np.repeat(np.arange(b.n), np.diff(b.off)) * b.n + b.adj).tolist())
    assert ka < kb


def test_rmat_skew(pr1):
    # a=0.57 concentrates edges on low ids: vertex 0 is the heaviest
    deg = np.diff(pr1.off)
    assert deg[0] == deg.max()
    assert deg[: pr1.n // 2].sum() > 0.7 * pr1.m


def test_golden_graph_digest(pr1):
    z = np.load(os.path.join(GOLD, "pr1_pagerank.npz"))
    assert str(z["digest"]) == digest(pr1)


# ---- PageRank -----------------------------------------------------------------------
def test_pagerank_matches_reference_binary(orc, pr1):
    z = np.load(os.path.join(GOLD, "pr1_pagerank.npz"))
    r, ept = orc.pagerank(pr1, int(z["iters"]), float(z["damping"]))
    # the fixture holds the reference's %.9e prints: ~5e-10 relative per value
    assert rel_l1(r, z["rank"]) < 1e-9
    assert np.max(np.abs(r - z["rank"]) / np.maximum(np.abs(z["rank"]), 1e-300)) < 1e-8
    assert ept[0] == pr1.m


def pagerank_matrix_form(g, iters, d):
    """P1 semantics in matrix form (independent of the push loop): frontier
    F_{t+1} = N_out(F_t); only F_t is updated / reset; two buffers swap."""
    A = csr_matrix(g)
    deg = np.diff(g.off).astype(np.float64)
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    cur = np.full(g.n, 1.0 / g.n)
    nxt = np.zeros(g.n)
    F = np.ones(g.n, bool)
    for _ in range(iters):
        contrib = np.where(F, cur * inv, 0.0)
        recv = A.T @ contrib
        N = (A.T @ F.astype(np.float64)) > 0
        nxt = np.where(F, d * (nxt + recv) + (1 - d) / g.n, nxt)
        cur = np.where(F, 0.0, cur)
        cur, nxt = nxt, cur
        F = N
    return cur


def test_pagerank_matrix_form(orc):
    g = orc.rmat(12, 30000, seed=3)
    for T in (1, 2, 5, 20):
        r, _ = orc.pagerank(g, T, 0.85)
        assert rel_l1(r, pagerank_matrix_form(g, T, 0.85)) < 1e-12


def test_pagerank_closed_forms(orc):
    # directed cycle: the uniform vector is a fixed point
    n = 7
    g = orc.from_edges(n, np.arange(n), (np.arange(n) + 1) % n)
    r, _ = orc.pagerank(g, 20, 0.85)
    assert np.allclose(r, 1.0 / n, rtol=0, atol=1e-15)
    # zero in-degree vertices print 0 after an even number of iterations and
    # (1-d)/n after an odd number (P1)
    g = orc.rmat(10, 8000, seed=4)
    indeg = np.bincount(g.adj, minlength=g.n)
    z = indeg == 0
    assert z.any()
    r20, _ = orc.pagerank(g, 20, 0.85)
    r19, _ = orc.pagerank(g, 19, 0.85)
    assert np.all(r20[z] == 0.0)
    assert np.allclose(r19[z], 0.15 / g.n, rtol=1e-15)


def test_pagerank_pull_equals_push(orc, pr1):
    a, _ = orc.pagerank(pr1, 20, 0.85)
    b = orc.pagerank_pull_omp(pr1, 20, 0.85, threads=4)
    assert rel_l1(b, a) < 1e-13


# ---- BFS ----------------------------------------------------------------------------
def test_bfs_reference_counts(orc, pr1):
    info = json.load(open(os.path.join(GOLD, "sym_bfs.json")))
    gs = pr1.symmetrized()
    assert info["digest"] == digest(gs)
    for run in info["runs"]:
        parent, level, rounds = orc.bfs_pull(gs, run["src"])
        assert (level >= 0).sum() == run["vertices_visited"]
        assert rounds == run["finished_in"]
        assert rounds == level.max() + 1


def check_parents(g, level, parent, src):
    reached = level >= 0
    assert parent[src] == src
    assert np.all(parent[~reached] == -1)
    v = np.nonzero(reached)[0]
    v = v[v != src]
    p = parent[v]
    assert np.all(level[p] == level[v] - 1)
    # p is an in-neighbour of v
    toff, tadj, _ = g.transpose()
    for vv, pp in zip(v[:2000], p[:2000]):
        assert pp in tadj[toff[vv]:toff[vv + 1]]


def test_bfs_levels_vs_scipy(orc, pr1):
    gs = pr1.symmetrized()
    M = csr_matrix(gs)
    for src in (0, 17, 4242):
        parent, level, _ = orc.bfs_pull(gs, src)
        lv = orc.bfs_levels(gs, src)
        assert np.array_equal(level, lv)
        d = csg.shortest_path(M, unweighted=True, indices=src)
        ref = np.where(np.isinf(d), -1, d).astype(np.int32)
        assert np.array_equal(level, ref)
        check_parents(gs, level, parent, src)


def test_bfs_directed(orc):
    g = orc.rmat(11, 20000, seed=9)
    parent, level, _ = orc.bfs_pull(g, 1)
    d = csg.shortest_path(csr_matrix(g), unweighted=True, indices=1, directed=True)
    assert np.array_equal(level, np.where(np.isinf(d), -1, d).astype(np.int32))
    check_parents(g, level, parent, 1)


# ---- connected components ------------------------------------------------------------
def min_id_labels(g):
    ncomp, lab = csg.connected_components(csr_matrix(g), directed=False)
    first = np.full(ncomp, g.n, np.int64)
    np.minimum.at(first, lab, np.arange(g.n))
    return first[lab].astype(np.int32)


def test_cc(orc, pr1):
    gs = pr1.symmetrized()
    ref = min_id_labels(gs)
    lab, rounds = orc.cc_labelprop(gs)
    assert np.array_equal(lab, ref)
    assert np.array_equal(orc.cc_unionfind(gs), ref)
    assert rounds >= 2


def test_cc_small_graphs(orc):
    rng = np.random.default_rng(0)
    for trial in range(20):
        n = int(rng.integers(1, 60))
        m = int(rng.integers(0, 2 * n + 1))
        s = rng.integers(0, n, m)
        d = rng.integers(0, n, m)
        keep = s != d
        g = orc.from_edges(n, np.r_[s[keep], d[keep]], np.r_[d[keep], s[keep]])
        ref = min_id_labels(g)
        assert np.array_equal(orc.cc_labelprop(g)[0], ref)
        assert np.array_equal(orc.cc_unionfind(g), ref)


# ---- SSSP ----------------------------------------------------------------------------
def test_dijkstra_and_bf_vs_scipy(orc, pr1):
    gs = pr1.symmetrized()
    w = orc.weights_i32(gs, 7)
    M = csr_matrix(gs, w)
    for src in (0, 31):
        d = orc.dijkstra(gs, w, src)
        ref = csg.dijkstra(M, indices=src)
        ref = np.where(np.isinf(ref), 1073741823, ref).astype(np.int64)
        assert np.array_equal(d.astype(np.int64), ref)
        bf, rounds = orc.sssp_bf(gs, w, src)
        assert np.array_equal(bf, d)


def test_weights_symmetric(orc, pr1):
    gs = pr1.symmetrized()
    w = orc.weights_i32(gs, 7)
    assert w.min() >= 1 and w.max() <= 255
    M = csr_matrix(gs, w)
    assert (M != M.T).nnz == 0


def test_reference_bellman_ford_counterexample(orc):
    # 0->B(10), 0->A(1), A->B(1), B->x(1) with id(B) < id(A): the reference
    # resets Visited on the current frontier, so B's improvement is lost
    B, A, x = 1, 2, 3
    src = [0, 0, A, B]
    dst = [B, A, B, x]
    wts = {(0, B): 10, (0, A): 1, (A, B): 1, (B, x): 1}
    g = orc.from_edges(4, src, dst)
    w = np.array([wts[(int(s), int(d))] for s, d in
                  zip(np.repeat(np.arange(4), np.diff(g.off)), g.adj)], np.int32)
    ref_bf, _ = orc.sssp_bf_reference(g, w, 0)
    assert ref_bf[x] == 11
    assert orc.dijkstra(g, w, 0)[x] == 3
    assert orc.sssp_bf(g, w, 0)[0][x] == 3


# ---- SpMV -------------------------------------------------------------------------------
def test_spmv_matches_reference_binary(orc, pr1):
    z = np.load(os.path.join(GOLD, "spmv_int.npz"))
    w = orc.weights_i32(pr1, int(z["weight_seed"]))
    assert str(z["digest"]) == digest(pr1, w)
    y = orc.spmv(pr1, w.astype(np.float64), np.full(pr1.n, 1.0 / pr1.n), int(z["iters"]))
    assert rel_l1(y, z["y"]) < 1e-9


def test_spmv_vs_scipy(orc):
    g = orc.rmat(12, 40000, seed=11)
    w = orc.values_f32(g, 3).astype(np.float64)
    assert w.min() > 0 and w.max() <= 1.0
    x0 = orc.vertex_f32(g.n, 5).astype(np.float64)
    M = csr_matrix(g, w)
    y = x0.copy()
    for _ in range(6):
        y = M.T @ y
    assert rel_l1(orc.spmv(g, w, x0, 6), y) < 1e-13


# ---- BP ---------------------------------------------------------------------------------
def bp_bruteforce(n, edges, vpot, epot, prod0, iters):
    """numa-BP.C:112-125 written edge by edge, sequential, float64."""
    cur = prod0.astype(np.float64).copy()
    nxt = np.zeros((n, 2))
    F = np.ones(n, bool)
    for _ in range(iters):
        N = np.zeros(n, bool)
        nxt[F] = 1.0
        for e, (s, d) in enumerate(edges):
            if not F[s]:
                continue
            for k in range(2):
                b = sum(float(vpot[d, l]) * float(epot[e, k, l]) * cur[d, l] for l in range(2))
                nxt[d, k] *= b
            N[d] = True
        cur, nxt = nxt, cur
        F = N
    return cur


def test_bp_bruteforce(orc):
    g = orc.rmat(6, 300, seed=2)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    edges = list(zip(src.tolist(), g.adj.tolist()))
    vpot = orc.vertex_f32(2 * g.n, 8, 0.1, 1.0).reshape(g.n, 2)
    epot = orc.bp_potentials(g, 9)
    prod0 = np.ones((g.n, 2))
    for iters in (1, 2, 5):
        ref = bp_bruteforce(g.n, edges, vpot, epot.reshape(-1, 2, 2), prod0, iters)
        got = orc.bp(g, vpot, epot, prod0, iters)
        assert np.allclose(got, ref, rtol=1e-14, atol=0)
        f32 = orc.bp_f32(g, vpot, epot, prod0.astype(np.float32), iters)
        assert rel_l1(f32, ref) < 1e-5


def test_bp_reference_initial_state(orc):
    # the reference starts from zero-filled products: every vertex with an
    # in-edge multiplies a zero message, the rest keep the reset value (1,1)
    g = orc.rmat(8, 2000, seed=2)
    vpot = orc.vertex_f32(2 * g.n, 8, 0.1, 1.0)
    epot = orc.bp_potentials(g, 9)
    out = orc.bp(g, vpot, epot, np.zeros((g.n, 2)), 1)
    indeg = np.bincount(g.adj, minlength=g.n)
    assert np.all(out[indeg > 0] == 0.0)
    assert np.all(out[indeg == 0] == 1.0)


from bp_hand import (BP_HAND_ADJ, BP_HAND_ANSWER, BP_HAND_EPOT, BP_HAND_OFF,  # noqa: E402
                     BP_HAND_PROD0, BP_HAND_VPOT)


def test_bp_hand_derived(orc):
    g = orc.Graph(3, BP_HAND_OFF, BP_HAND_ADJ)
    for iters, want in BP_HAND_ANSWER.items():
        got = orc.bp(g, BP_HAND_VPOT, BP_HAND_EPOT, BP_HAND_PROD0, iters)
        assert np.array_equal(got, want), (iters, got)
        f32 = orc.bp_f32(g, BP_HAND_VPOT, BP_HAND_EPOT, BP_HAND_PROD0.astype(np.float32), iters)
        assert np.array_equal(f32, want.astype(np.float32)), (iters, f32)
    # a planted transposition of W fails the known answer
    epot_t = BP_HAND_EPOT.reshape(3, 2, 2).transpose(0, 2, 1).ravel().copy()
    assert not np.array_equal(orc.bp(g, BP_HAND_VPOT, epot_t, BP_HAND_PROD0, 1), BP_HAND_ANSWER[1])


def test_bp_potential_range(orc, pr1):
    p = orc.bp_potentials(pr1, 9)
    assert p.min() >= 0.1 and p.max() < 1.0 and p.size == 4 * pr1.m


# ---- PPR and weighted PageRank -----------------------------------------------------
def test_ppr_dense(orc):
    g = orc.rmat(9, 4000, seed=12)
    A = csr_matrix(g).toarray()
    deg = A.sum(1)
    P = np.where(deg[:, None] > 0, A / np.maximum(deg, 1)[:, None], 0.0)
    v = np.zeros(g.n)
    v[3] = 1.0
    r = v.copy()
    for _ in range(20):
        r = 0.85 * P.T @ r + 0.15 * v
    assert rel_l1(orc.ppr(g, v, 20, 0.85), r) < 1e-13


def column_normalised(orc, g, seed):
    w = orc.values_f32(g, seed).astype(np.float64)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    tot = np.bincount(src, weights=w, minlength=g.n)
    return w / tot[src]


def test_wpr_forward_adjoint_dense(orc):
    g = orc.rmat(8, 2000, seed=13)
    wn = column_normalised(orc, g, 4)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    W = np.zeros((g.n, g.n))
    W[g.adj, src] = wn
    rng = np.random.default_rng(1)
    v = rng.random(g.n)
    v /= v.sum()
    gv = rng.standard_normal(g.n)
    d = 0.85
    x_ref = np.linalg.solve(np.eye(g.n) - d * W, (1 - d) * v)
    lam_ref = np.linalg.solve((np.eye(g.n) - d * W).T, gv)
    x, it = orc.wpr_forward(g, wn, v, d, 2000, 1e-15)
    assert rel_l1(x, x_ref) < 1e-12
    lam, gw, gt, _ = orc.wpr_adjoint(g, wn, x, gv, d, 2000, 1e-15)
    assert rel_l1(lam, lam_ref) < 1e-11
    assert np.allclose(gt, (1 - d) * lam_ref, rtol=1e-10, atol=1e-14)
    assert np.allclose(gw, d * lam_ref[g.adj] * x_ref[src], rtol=1e-9, atol=1e-15)
    # central differences on a few edges
    for e in (0, 7, g.m // 2):
        h = 1e-6
        wp, wm = wn.copy(), wn.copy()
        wp[e] += h
        wm[e] -= h
        xp, _ = orc.wpr_forward(g, wp, v, d, 2000, 1e-15)
        xm, _ = orc.wpr_forward(g, wm, v, d, 2000, 1e-15)
        fd = (gv @ xp - gv @ xm) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-6 * max(1.0, abs(gw[e]))


# ---- IO -------------------------------------------------------------------------------
def test_pbbs_roundtrip(orc, tmp_path):
    g = orc.rmat(8, 1500, seed=3)
    p = str(tmp_path / "g.adj")
    orc.write_adj(p, g)
    h, w = orc.read_adj(p)
    assert w is None and np.array_equal(h.off, g.off) and np.array_equal(h.adj, g.adj)
    wt = orc.weights_i32(g, 7)
    orc.write_adj(p, g, wt)
    h, w = orc.read_adj(p)
    assert np.array_equal(w, wt)
    # self-loop removal (readGraphFromFileSkipRing)
    g2 = orc.from_edges(3, [0, 0, 1, 2], [0, 1, 1, 0])
    orc.write_adj(p, g2)
    h, _ = orc.read_adj(p, skip_self_loops=True)
    assert h.m == 2 and list(h.adj) == [1, 0]


def test_pbbs_rejects_bad_length(orc, tmp_path):
    p = str(tmp_path / "bad.adj")
    with open(p, "w") as f:
        f.write("AdjacencyGraph\n3\n2\n0\n1\n2\n1\n")  # one edge missing
    with pytest.raises(RuntimeError):
        orc.read_adj(p)


def test_binary_graph_layout_and_round_trip(orc, tmp_path):
    """or_write_bin / or_read_bin (IO-numa.h:543-635): the file layout the
    reference reads (n as text; m 32-bit ids; n offsets of 4 or 8 bytes, no
    closing m) and the same graph as the PBBS text reader gives."""
    g = orc.rmat(10, 1 << 13, seed=9)
    p = str(tmp_path / "g.adj")
    orc.write_adj(p, g)
    ht, _ = orc.read_adj(p)
    for ib in (4, 8):
        base = str(tmp_path / ("g%d" % ib))
        orc.write_bin(base, g, ib)
        assert open(base + ".config").read().split() == [str(g.n)]
        assert os.path.getsize(base + ".adj") == 4 * g.m
        assert os.path.getsize(base + ".idx") == ib * g.n
        idx = np.fromfile(base + ".idx", np.int32 if ib == 4 else np.int64)
        assert np.array_equal(idx, g.off[:-1])
        assert np.array_equal(np.fromfile(base + ".adj", np.uint32), g.adj.astype(np.uint32))
        h = orc.read_bin(base)
        assert h.n == g.n and h.m == g.m
        assert np.array_equal(h.off, ht.off) and np.array_equal(h.adj, ht.adj)
    # a graph whose last vertices have no out-edges: off[n] = m closes the last list
    g2 = orc.from_edges(5, [0, 0, 2], [1, 2, 0])
    base = str(tmp_path / "tail")
    orc.write_bin(base, g2, 4)
    h = orc.read_bin(base)
    assert list(h.off) == [0, 2, 2, 3, 3, 3] and list(h.adj) == [1, 2, 0]


def test_binary_graph_rejects_bad_idx_size(orc, tmp_path):
    g = orc.rmat(8, 1 << 10, seed=3)
    base = str(tmp_path / "g")
    orc.write_bin(base, g, 4)
    with open(base + ".idx", "ab") as f:
        f.write(b"\0\0")  # neither 4n nor 8n bytes ("File size wrong")
    with pytest.raises(RuntimeError):
        orc.read_bin(base)


def _power_iteration(g, iters, d):
    """p_0 = 1/n; p_{t+1} = d * sum over in-edges of p_t[s] / outdeg(s) + (1-d)/n."""
    n = g.n
    deg = np.diff(g.off)
    src = np.repeat(np.arange(n), deg)
    p = np.full(n, 1.0 / n)
    for _ in range(iters):
        c = np.where(deg > 0, p / np.maximum(deg, 1), 0.0)
        p = d * np.bincount(g.adj, weights=c[src], minlength=n) + (1.0 - d) / n
    return p


def test_pagerank_delta_without_threshold_is_the_power_iteration(orc):
    """epsilon2 = 0 keeps every vertex with a nonzero delta active, so the delta
    recurrence telescopes to p_t of the power iteration (round 1's -1/n turns p_1
    into the first delta p_1 - p_0)."""
    for seed, sym in ((31, False), (32, True)):
        g = orc.rmat(10, 1 << 13, seed=seed)
        if sym:
            g = g.symmetrized()
        for iters in (1, 2, 7):
            p, act = orc.pagerank_delta(g, iters, 0.85, 0.0)
            ref = _power_iteration(g, iters, 0.85)
            assert np.abs(p - ref).sum() / ref.sum() < 1e-12, (seed, iters)
            assert act[0] == g.n


def test_pagerank_delta_large_threshold_freezes_after_round_one(orc):
    g = orc.rmat(10, 1 << 13, seed=33)
    p1, _ = orc.pagerank_delta(g, 1, 0.85, 1e9)
    p5, act = orc.pagerank_delta(g, 5, 0.85, 1e9)
    assert np.array_equal(p1, p5) and list(act) == [g.n, 0, 0, 0, 0]
    ref = _power_iteration(g, 1, 0.85)
    assert np.abs(p1 - ref).max() < 1e-15


def test_pagerank_delta_known_answer(orc):
    """Edges 0->1, 0->2, 1->2, 2->0, d = 0.85, epsilon2 = 0.01, derived by hand:
    round 1 gives p = (1/3, d/6 + 0.05, d/2 + 0.05) with deltas (0, -d/6, d/6), so
    1 and 2 stay active; round 2 moves d^2/6 from 2 to 0 and -d^2/6 from 1 to 2
    (0 and 2 active, 1 not); round 3 moves -d^3/6 to 0 and d^3/12 to 1 and 2."""
    g = orc.from_edges(3, [0, 0, 1, 2], [1, 2, 2, 0])
    d = 0.85
    p, act = orc.pagerank_delta(g, 3, d, 0.01)
    want = [1 / 3 + d * d / 6 - d ** 3 / 6,
            d / 6 + 0.05 + d ** 3 / 12,
            d / 2 + 0.05 - d * d / 6 + d ** 3 / 12]
    assert np.allclose(p, want, rtol=0, atol=1e-15), (p, want)
    assert list(act) == [3, 2, 2]
    assert abs(p.sum() - 1.0) < 1e-15


def test_pull_omp_variants_bit_identical(orc):
    """The OpenMP pulls that check the full-size GPU runs form every sum and product
    in the push's order: equal bits to or_spmv / or_bp on skewed graphs (duplicate-free
    rMat with hubs, isolated vertices, a few frontier-shrinking BP rounds)."""
    for scale, m, seed in ((8, 3000, 3), (12, 60000, 4)):
        g = orc.rmat(scale, m, seed=seed)
        w = orc.values_f32(g, 5).astype(np.float64)
        x0 = orc.vertex_f32(g.n, 6, 0.0, 1.0).astype(np.float64)
        for iters in (1, 4):
            assert np.array_equal(orc.spmv_pull_omp(g, w, x0, iters, threads=4),
                                  orc.spmv(g, w, x0, iters))
        vpot = orc.vertex_f32(2 * g.n, 8, 0.1, 1.0)
        epot = orc.bp_potentials(g, 9)
        for prod0 in (np.ones(2 * g.n), orc.vertex_f32(2 * g.n, 2, 0.5, 2.0).astype(np.float64)):
            for iters in (1, 2, 5):
                a = orc.bp_pull_omp(g, vpot, epot, prod0, iters, threads=4)
                b = orc.bp(g, vpot, epot, prod0, iters)
                assert np.array_equal(a, b, equal_nan=True), (scale, iters)
    gh = orc.Graph(3, BP_HAND_OFF, BP_HAND_ADJ)
    for iters, want in BP_HAND_ANSWER.items():
        assert np.array_equal(orc.bp_pull_omp(gh, BP_HAND_VPOT, BP_HAND_EPOT, BP_HAND_PROD0, iters),
                              want)


def test_spmv_adjoint_dense_and_fd(orc):
    """or_spmv_adjoint against the dense-matrix gradient (numpy: y_{t+1} = A y_t,
    mu_t = A^T mu_{t+1}, dL/dA[d, s] = sum_t mu_{t+1}[d] y_t[s]) and fp64 central
    differences of or_spmv's loss (a planted sign or index error fails both)."""
    g = orc.rmat(9, 4000, seed=31)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    rng = np.random.default_rng(2)
    w = rng.uniform(0.05, 1.0, g.m) / 4.0
    x0 = rng.random(g.n)
    gv = rng.standard_normal(g.n)
    T = 5
    A = np.zeros((g.n, g.n))
    A[g.adj, src] = w
    ys = [x0]
    for _ in range(T):
        ys.append(A @ ys[-1])
    mus = [None] * (T + 1)
    mus[T] = gv
    for t in range(T - 1, -1, -1):
        mus[t] = A.T @ mus[t + 1]
    gw_dense = sum(mus[t + 1][g.adj] * ys[t][src] for t in range(T))
    gx, gw = orc.spmv_adjoint(g, w, x0, gv, T)
    assert rel_l1(gx, mus[0]) < 1e-13
    assert rel_l1(gw, gw_dense) < 1e-13
    assert abs(gv @ orc.spmv(g, w, x0, T) - gv @ ys[T]) < 1e-10 * abs(gv @ ys[T])
    h = 1e-6
    for e in (0, 17, g.m // 2, g.m - 1):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (gv @ orc.spmv(g, wp, x0, T) - gv @ orc.spmv(g, wm, x0, T)) / (2 * h)
        assert abs(gw[e] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(gw).max()), (e, gw[e], fd)
    for i in (1, 100, g.n - 1):
        xp, xm = x0.copy(), x0.copy()
        xp[i] += h
        xm[i] -= h
        fd = (gv @ orc.spmv(g, w, xp, T) - gv @ orc.spmv(g, w, xm, T)) / (2 * h)
        assert abs(gx[i] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(gx).max()), (i, gx[i], fd)
