"""GPU parity of the batched and differentiated forms.

  * personalized PageRank: k one-hot teleports as k-wide rows vs the fp64 oracle
    and vs the plain one-vector kernel, 1e-6 relative L1 per source;
  * multi-source BFS / Bellman-Ford: per-source levels / distances bit-exact
    against the textbook BFS and Dijkstra;
  * many graphs in one captured launch: per-graph PageRank vs the oracle;
  * weighted PageRank adjoint: gradients w.r.t. edge weights and teleport vs a
    dense fp64 solve and fp64 central differences (h = 1e-6), also through a
    torch.autograd.Function wrapper (test harness only).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-6


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def say(*a):
    print(*a, flush=True)


def onehot(n, s):
    v = np.zeros(n)
    v[s] = 1.0
    return v


# ---- personalized PageRank ---------------------------------------------------------------
def test_ppr_plain(pa, orc):
    """The single-vector form on the PageRank tiles (with and without the cold
    pre-pass, split tiles) and on the degree-binned pull; 1-3 iterations cover the
    settling of the zero-in-degree tail."""
    say("case: ppr plain")
    g = orc.rmat(14, 1 << 18, seed=31)
    d = pa.Graph.rmat(14, 1 << 18, seed=31)
    for tiles, te, cold in ((1, 0, 0), (1, 4096, 2000), (1, 1 << 30, 6000), (0, 0, 0)):
        d.set_tiles(tiles, te, cold if cold else 0)
        for s in (0, 77, 9999):
            for iters in (1, 2, 3, 20):
                r, _ = d.ppr(s, iters, 0.85)
                ref = orc.ppr(g, onehot(g.n, s), iters, 0.85)
                assert rel_l1(r, ref) < TOL, (tiles, te, cold, s, iters)
    d.set_tiles(1)


@pytest.mark.parametrize("k", [64, 5])
def test_ppr_batched(pa, orc, k):
    say("case: ppr batched k", k)
    g = orc.rmat(14, 1 << 18, seed=31)
    d = pa.Graph.rmat(14, 1 << 18, seed=31)
    rng = np.random.default_rng(k)
    src = rng.choice(g.n, k, replace=False)
    src[0] = 0  # the heaviest vertex (multi-chunk rows)
    R, st = d.ppr_batched(src, 20, 0.85)
    worst = 0.0
    for c in range(k):
        ref = orc.ppr(g, onehot(g.n, src[c]), 20, 0.85)
        worst = max(worst, rel_l1(R[c], ref))
    say("ppr batched worst relL1 %.3e" % worst)
    assert worst < TOL
    plain, _ = d.ppr(int(src[1]), 20, 0.85)
    assert rel_l1(R[1], plain) < TOL


def test_ppr_batched_short(pa, orc):
    """1-3 iterations (the first iteration runs as a push from the sources' rows
    unless it is the last), with a repeated source and a source without out-edges."""
    say("case: ppr batched short")
    g = orc.rmat(14, 1 << 18, seed=31)
    d = pa.Graph.rmat(14, 1 << 18, seed=31)
    outdeg = np.diff(g.off)
    sink = int(np.flatnonzero(outdeg == 0)[0])
    src = np.array([0, 77, 77, sink, 9999, int(np.argmax(outdeg))], np.int64)
    for iters in (1, 2, 3):
        R, _ = d.ppr_batched(src, iters, 0.85)
        for c in range(len(src)):
            ref = orc.ppr(g, onehot(g.n, src[c]), iters, 0.85)
            assert rel_l1(R[c], ref) < TOL, (iters, c, int(src[c]))


# ---- multi-source BFS / SSSP -------------------------------------------------------------
def test_msbfs(pa, orc):
    say("case: msbfs 64 sources")
    g = orc.rmat(14, 1 << 18, seed=41)
    gs = g.symmetrized()
    d = pa.Graph.rmat(14, 1 << 18, seed=41, symmetrize=True)
    rng = np.random.default_rng(2)
    src = rng.choice(g.n, 64, replace=False)
    src[0] = 0
    L, st = d.msbfs(src)
    for c in range(64):
        assert np.array_equal(L[c], orc.bfs_levels(gs, int(src[c]))), c
    say("msbfs rounds %d, pull rounds %d" % (st.rounds, st.dense_rounds))
    assert 0 < st.dense_rounds < st.rounds  # small frontiers push, large ones pull
    # fewer sources, directed graph, one source without out-edges
    dd = pa.Graph.rmat(14, 1 << 18, seed=41)
    iso = int(np.nonzero(np.diff(g.off) == 0)[0][0])
    s7 = np.array(list(src[:6]) + [iso])
    L, _ = dd.msbfs(s7)
    for c in range(7):
        assert np.array_equal(L[c], orc.bfs_levels(g, int(s7[c]))), c


def test_mssssp(pa, orc):
    say("case: mssssp 64 sources")
    g = orc.rmat(13, 1 << 17, seed=43)
    gs = g.symmetrized()
    w = orc.weights_i32(gs, 7)
    d = pa.Graph.rmat(13, 1 << 17, seed=43, symmetrize=True, weights="i32", wseed=7)
    rng = np.random.default_rng(3)
    src = rng.choice(g.n, 64, replace=False)
    src[0] = 0
    D, st = d.mssssp(src)
    say("mssssp rounds", st.rounds)
    for c in range(64):
        assert np.array_equal(D[c], orc.dijkstra(gs, w, int(src[c]))), c


# ---- many graphs -------------------------------------------------------------------------
def make_many(orc, count, seed):
    rng = np.random.default_rng(seed)
    graphs, vb, offs, adjs, base, ebase = [], [0], [], [], 0, 0
    for i in range(count):
        scale = int(rng.integers(6, 12))
        h = orc.rmat(scale, 16 << scale, seed=1000 + i)
        graphs.append(h)
        offs.append(h.off[:-1] + ebase)
        adjs.append(h.adj + base)
        base += h.n
        ebase += h.m
        vb.append(base)
    off = np.concatenate(offs + [np.array([ebase])]).astype(np.int64)
    adj = np.concatenate(adjs).astype(np.int32)
    return graphs, np.array(vb, np.int64), off, adj


def test_multigraph(pa, orc):
    say("case: many graphs in one launch")
    graphs, vb, off, adj = make_many(orc, 40, 5)
    d = pa.Graph.multi(vb, off, adj)
    r1, st = d.pagerank_multigraph(20, 0.85)
    r2, _ = d.pagerank_multigraph(20, 0.85)            # replay of the captured graph
    r3, _ = d.pagerank_multigraph(20, 0.85, profile=True)  # launch by launch
    assert np.array_equal(r1, r2) and np.array_equal(r1, r3)
    worst = 0.0
    for i, h in enumerate(graphs):
        ref, _ = orc.pagerank(h, 20, 0.85)
        worst = max(worst, rel_l1(r1[vb[i]:vb[i + 1]], ref))
    say("multigraph worst relL1 %.3e" % worst)
    assert worst < TOL
    # a schedule change invalidates the capture; results stay the same
    d.set_schedule(8, 1024, 8192, 2048)
    r4, _ = d.pagerank_multigraph(20, 0.85)
    assert rel_l1(r4, r1) < 1e-7


def test_autotune(pa, orc):
    say("case: autotune")
    g = orc.rmat(15, 1 << 19, seed=3)
    d = pa.Graph.rmat(15, 1 << 19, seed=3)
    ipl, bt = d.autotune()
    say("autotune picked ipl", ipl, "block threshold", bt)
    r, _ = d.pagerank(20)
    ref, _ = orc.pagerank(g, 20, 0.85)
    assert rel_l1(r, ref) < TOL


# ---- weighted PageRank and its adjoint ----------------------------------------------------
def column_normalised(orc, g, seed):
    w = orc.values_f32(g, seed).astype(np.float64)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    tot = np.bincount(src, weights=w, minlength=g.n)
    return w / tot[src]


def test_wpr_adjoint_dense_and_fd(pa, orc):
    say("case: weighted pagerank adjoint")
    g = orc.rmat(12, 1 << 16, seed=17)
    wn = column_normalised(orc, g, 4)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    rng = np.random.default_rng(9)
    v = rng.random(g.n)
    v /= v.sum()
    gv = rng.standard_normal(g.n)
    dmp = 0.85
    W = np.zeros((g.n, g.n))
    W[g.adj, src] = wn
    A = np.eye(g.n) - dmp * W
    x_ref = np.linalg.solve(A, (1 - dmp) * v)
    lam_ref = np.linalg.solve(A.T, gv)
    gw_ref = dmp * lam_ref[g.adj] * x_ref[src]
    gt_ref = (1 - dmp) * lam_ref
    d = pa.Graph.from_csr(g.n, g.off, g.adj)
    x, stf = d.wpr_forward(wn.astype(np.float32), v.astype(np.float32), dmp, 500, 1e-9)
    lam, gw, gt, sta = d.wpr_adjoint(gv.astype(np.float32), 500, 1e-7)
    say("forward %d iters, adjoint %d iters" % (stf.rounds, sta.rounds))
    assert rel_l1(x, x_ref) < TOL
    assert rel_l1(lam, lam_ref) < TOL
    assert rel_l1(gt, gt_ref) < TOL
    assert rel_l1(gw, gw_ref) < 1e-5
    # fp64 central differences of the fp64 oracle's loss on a few edges and on v
    for e in (0, 5, g.m // 3, g.m - 1):
        h = 1e-6
        wp, wm = wn.copy(), wn.copy()
        wp[e] += h
        wm[e] -= h
        xp, _ = orc.wpr_forward(g, wp, v, dmp, 3000, 1e-15)
        xm, _ = orc.wpr_forward(g, wm, v, dmp, 3000, 1e-15)
        fd = (gv @ xp - gv @ xm) / (2 * h)
        assert abs(gw[e] - fd) <= 1e-4 * max(abs(fd), 1e-3), (e, gw[e], fd)
    for i in (1, 100):
        h = 1e-6
        vp, vm = v.copy(), v.copy()
        vp[i] += h
        vm[i] -= h
        xp, _ = orc.wpr_forward(g, wn, vp, dmp, 3000, 1e-15)
        xm, _ = orc.wpr_forward(g, wn, vm, dmp, 3000, 1e-15)
        fd = (gv @ xp - gv @ xm) / (2 * h)
        assert abs(gt[i] - fd) <= 1e-4 * max(abs(fd), 1e-3), (i, gt[i], fd)



def test_spmv_adjoint_vs_oracle(pa, orc):
    """pg_spmv_adjoint (numa-SPMV.C's iteration differentiated: a tiled pull over the
    CSR per step, the weight gradient in passes of up to 32 steps over the edges)
    against the oracle's fp64 adjoint, which tests/test_oracle.py pins by the
    dense-matrix gradient and central differences: fp32 and integer weights, the
    forward on the tiles and on the degree-binned pull, tiles with many parts and
    cold sources, 40 steps (two gradient passes), and a second loss on the same
    kept iterates."""
    say("case: spmv adjoint")
    g = orc.rmat(13, 1 << 17, seed=41)
    rng = np.random.default_rng(3)
    for wkind in ("f32", "i32"):
        if wkind == "f32":
            w = orc.values_f32(g, 5)
            d = pa.Graph.from_csr(g.n, g.off, g.adj, w=w)
        else:
            w = orc.weights_i32(g, 6)
            d = pa.Graph.from_csr(g.n, g.off, g.adj, w=w)
        x0 = rng.random(g.n).astype(np.float32)
        for tiles, te, cold in ((1, 0, 0), (0, 0, 0), (1, 4096, 1000)):
            d.set_tiles(tiles, te, cold)
            for T in ((1, 6, 40) if tiles and not te else (1, 6)):
                y_plain, _ = d.spmv(x0, T)
                y, _ = d.spmv(x0, T, save=True)
                assert np.array_equal(y, y_plain)  # keeping the iterates changes nothing
                for k in range(2):
                    gv = rng.standard_normal(g.n).astype(np.float32)
                    gx, gw, st = d.spmv_adjoint(gv)
                    rx, rw = orc.spmv_adjoint(g, w.astype(np.float64), x0.astype(np.float64),
                                              gv.astype(np.float64), T)
                    ex, ew = rel_l1(gx, rx), rel_l1(gw, rw)
                    say("%s weights, tiles %d %d %d, T %d: dL/dx0 relL1 %.3e, dL/dw relL1 %.3e" %
                        (wkind, tiles, te, cold, T, ex, ew))
                    assert ex < TOL and ew < TOL
                    assert st.rounds == T
        d.close()
    # no kept iterates: refused
    d = pa.Graph.from_csr(g.n, g.off, g.adj, w=orc.values_f32(g, 5))
    d.spmv(np.ones(g.n, np.float32), 3)
    with pytest.raises(pa.PolymerError):
        d.spmv_adjoint(np.ones(g.n, np.float32))
    d.close()

def test_wpr_torch_autograd(pa, orc):
    """A torch.autograd.Function over the C ABI (harness only): gradients flow
    through the column normalisation of raw weights done in torch."""
    torch = pytest.importorskip("torch")
    say("case: weighted pagerank through torch.autograd")
    g = orc.rmat(10, 1 << 14, seed=23)
    d = pa.Graph.from_csr(g.n, g.off, g.adj)
    src = torch.as_tensor(np.repeat(np.arange(g.n), np.diff(g.off)), dtype=torch.long)

    class WeightedPageRank(torch.autograd.Function):
        @staticmethod
        def forward(ctx, wn, v):
            x, _ = d.wpr_forward(wn.detach().numpy().astype(np.float32),
                                 v.detach().numpy().astype(np.float32), 0.85, 500, 1e-9)
            return torch.as_tensor(x, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gx):
            _, gw, gt, _ = d.wpr_adjoint(gx.numpy().astype(np.float32), 500, 1e-8)
            return torch.as_tensor(gw, dtype=torch.float64), torch.as_tensor(gt, dtype=torch.float64)

    raw = torch.as_tensor(orc.values_f32(g, 4), dtype=torch.float64).requires_grad_(True)
    v = torch.full((g.n,), 1.0 / g.n, dtype=torch.float64, requires_grad=True)
    gvec = torch.as_tensor(np.random.default_rng(0).standard_normal(g.n))
    tot = torch.zeros(g.n, dtype=torch.float64).index_add(0, src, raw)
    wn = raw / tot[src]
    loss = (gvec * WeightedPageRank.apply(wn, v)).sum()
    loss.backward()
    # fp64 central difference of the same loss on raw weights, through the oracle
    for e in (3, g.m // 2):
        h = 1e-6

        def L(rw):
            rw = torch.as_tensor(rw)
            t = torch.zeros(g.n, dtype=torch.float64).index_add(0, src, rw)
            x, _ = orc.wpr_forward(g, (rw / t[src]).numpy(), v.detach().numpy(), 0.85, 3000, 1e-15)
            return float(gvec.numpy() @ x)

        rp = raw.detach().numpy().copy()
        rm = rp.copy()
        rp[e] += h
        rm[e] -= h
        fd = (L(rp) - L(rm)) / (2 * h)
        got = float(raw.grad[e])
        assert abs(got - fd) <= 1e-4 * max(abs(fd), 1e-3), (e, got, fd)
