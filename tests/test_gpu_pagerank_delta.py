"""GPU parity of PageRankDelta (numa-PageRankDelta.C:63-330) through the C-ABI:
fp64 ranks within 1e-9 relative L1 of the oracle (fp64, push order; the GPU sums
in pull order for large frontiers, on the source-sorted tiles with fp64 LDS adds
when the graph has them, and with fp64 atomics for small frontiers), the frontier
sizes summed over the rounds equal, and the hand-derived 3-vertex known answer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel_l1(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).sum() / max(np.abs(b).sum(), 1e-300))


@pytest.mark.parametrize("sym", [False, True])
def test_pagerank_delta_vs_oracle(pa, orc, sym):
    g = orc.rmat(12, 1 << 16, seed=51)
    if sym:
        g = g.symmetrized()
    gg = pa.Graph.from_csr(g.n, g.off, g.adj)
    for iters, eps2 in ((0, 0.01), (1, 0.01), (10, 0.01), (10, 0.0), (25, 0.05)):
        p, st = gg.pagerank_delta(iters, 0.85, eps2)
        ref, act = orc.pagerank_delta(g, iters, 0.85, eps2)
        e = rel_l1(p, ref) if iters else float(np.abs(p).sum())
        say("sym %d iters %d eps2 %g: rel L1 %.3e, frontier total %d (oracle %d), %d pull rounds" %
            (sym, iters, eps2, e, st.active, int(act.sum()), st.dense_rounds))
        assert e < 1e-9, (iters, eps2)
        assert st.active == int(act.sum()), (iters, eps2)
        if iters:
            assert st.rounds == iters
            assert st.edges >= g.m  # round 1 pushes every edge
        if iters == 25:  # the frontier shrinks: later rounds take the push path
            assert st.dense_rounds < iters
    gg.close()


def test_pagerank_delta_tiles(pa, orc):
    """The pull rounds on the tiles: one-part and many-part tiles (the finishing pass
    adds fp64 partial sums), the cold pre-pass staging doubles in two halves of a
    source chunk, and the untiled pull, all against the oracle."""
    g = orc.rmat(16, 1 << 20, seed=8)
    iters, eps2 = 15, 0.01
    ref, act = orc.pagerank_delta(g, iters, 0.85, eps2)
    d = pa.Graph.rmat(16, 1 << 20, seed=8)
    d.set_tiles(max_degree=2048)
    for tiles, te, cold in ((1, 1 << 30, -1), (1, 7, -1), (1, 4096, 1000), (1, 262144, 5000),
                            (1, 7, 30000), (0, 0, -1)):
        d.set_tiles(tiles, te, cold)
        p, st = d.pagerank_delta(iters, 0.85, eps2)
        e = rel_l1(p, ref)
        say("tiles %d edges %d cold %d: %s rel L1 %.3e, frontier total %d (oracle %d), %d pull "
            "rounds" % (tiles, te, cold, d.tile_stats(), e, st.active, int(act.sum()),
                        st.dense_rounds))
        assert e < 1e-9, (tiles, te, cold)
        assert st.active == int(act.sum()), (tiles, te, cold)
        assert st.dense_rounds >= 2
        if tiles and cold >= 0:
            assert d.tile_stats()["cold_edges"] > 0
    d.close()


def test_pagerank_delta_known_answer(pa, orc):
    """0->1, 0->2, 1->2, 2->0 after three rounds (derived by hand)."""
    d = 0.85
    gg = pa.Graph.from_csr(3, [0, 2, 3, 4], [1, 2, 2, 0])
    p, st = gg.pagerank_delta(3, d, 0.01)
    want = [1 / 3 + d * d / 6 - d ** 3 / 6,
            d / 6 + 0.05 + d ** 3 / 12,
            d / 2 + 0.05 - d * d / 6 + d ** 3 / 12]
    assert np.allclose(p, want, rtol=0, atol=1e-15), (p, want)
    assert st.active == 3 + 2 + 2
    gg.close()
