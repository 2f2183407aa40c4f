// Self-checks of how include/polymer.h reports frontier counts, run on the GPU by
// tests/test_gpu_dropin.py (a device-resident build, like the other drivers):
//   * the count a dense edgeMap takes of its output (vertexCounter then reads it
//     instead of recounting) against a recount on the host, and vertexCounter after
//     the flags were changed with toDevice (it must count again);
//   * edgeMapSparseV3 with a functor for which every edge succeeds, from lists on
//     both sides of the push's small-list limit (1024 vertices): the successes, their
//     out-degree sum and the output list (as a multiset) against the host;
//   * vertexFilter into a frontier: its count against the host.
// Prints one line per failed check and "header_checks: ok" when all hold; the exit
// status is the number of failed checks.
#include <algorithm>

#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
struct Claim_F {  // the BFS claim of numa-BFS.C
    intT *Parents;
    inline bool update(intT s, intT d) {
        if (Parents[d] == -1) {
            Parents[d] = s;
            return true;
        }
        return false;
    }
    inline bool updateAtomic(intT s, intT d) { return CAS(&Parents[d], (intT)-1, s); }
    inline bool cond(intT d) { return Parents[d] == -1; }
};

struct Every_F {  // every edge succeeds: the successes are the edges pushed
    inline bool update(intT, intT) { return true; }
    inline bool updateAtomic(intT, intT) { return true; }
    inline bool cond(intT) { return true; }
};

struct Odd_F {
    inline bool operator()(intT i) { return (i & 1) != 0; }
};
POLYMER_FUNCTORS_END

static int fails = 0;
static void expect(bool ok, const char *what, long long got, long long want) {
    if (!ok) {
        std::printf("FAIL %s: %lld, expected %lld\n", what, got, want);
        fails++;
    }
}

int main() {
    graph<symmetricVertex> G = rmatGraph<symmetricVertex>(14, 1 << 18, 5, true);
    const intT n = (intT)G.n;
    std::vector<int32_t> off((size_t)n + 1);
    toHost(off.data(), G.view.out_off, (size_t)n + 1);
    auto deg = [&](intT v) { return (long long)(off[v + 1] - off[v]); };
    std::vector<int32_t> adj((size_t)off[n]);
    toHost(adj.data(), G.view.out_adj, adj.size());

    LocalFrontier *cur = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *next = new LocalFrontier(newFrontierArray(n), 0, n);
    vertices *V = new vertices(1);
    V->registerFrontier(0, cur);
    V->calculateOffsets();

    // a dense frontier from host flags; its counts by vertexCounter
    auto load = [&](const std::vector<char> &h) {
        if (!cur->isDense) cur->toDense();
        toDevice(cur->b, reinterpret_cast<const bool *>(h.data()), (size_t)n);
        cur->isDense = true;
        V->isDense = true;
        vertexCounter(G, cur, 0, 0, 1);
        V->m = -1;
        V->calculateNumOfNonZero(0);
    };
    auto host_count = [&](const std::vector<char> &h, long long *m, long long *e) {
        *m = *e = 0;
        for (intT v = 0; v < n; v++)
            if (h[v]) {
                (*m)++;
                *e += deg(v);
            }
    };

    // ---- a dense edgeMap's own count, then a count after the flags change ----
    {
        std::vector<char> h((size_t)n, 0);
        std::vector<intT> par((size_t)n, -1);
        for (intT v = 0; v < n; v += 7)
            if (deg(v) > 0) {
                h[v] = 1;
                par[v] = v;
            }
        intT sz = n;
        intT *Parents = static_cast<intT *>(mapDataArray(1, &sz, sizeof(intT)));
        toDevice(Parents, par.data(), (size_t)n);
        load(h);
        long long hm, he;
        host_count(h, &hm, &he);
        expect(cur->m == hm, "vertexCounter m", cur->m, hm);
        expect(cur->outEdgesCount == he, "vertexCounter out-edges", cur->outEdgesCount, he);
        edgeMap(G, V, Claim_F{Parents}, next, 0, DENSE);  // threshold 0: the dense pull
        expect(next->isDense, "dense edgeMap output dense", next->isDense, 1);
        vertexCounter(G, next, 0, 0, 1);
        std::vector<char> o((size_t)n);
        toHost(reinterpret_cast<bool *>(o.data()), next->b, (size_t)n);
        long long om, oe;
        host_count(o, &om, &oe);
        expect(next->m == om, "dense edgeMap count m", next->m, om);
        expect(next->outEdgesCount == oe, "dense edgeMap count out-edges", next->outEdgesCount, oe);
        // flags changed after the edgeMap: vertexCounter counts them again
        intT x = 0;
        while (x < n && (o[x] || deg(x) == 0)) x++;
        if (x < n) {
            const bool one = true;
            toDevice(next->b + x, &one, 1);
            vertexCounter(G, next, 0, 0, 1);
            expect(next->m == om + 1, "count after toDevice m", next->m, om + 1);
            expect(next->outEdgesCount == oe + deg(x), "count after toDevice out-edges",
                   next->outEdgesCount, oe + deg(x));
        }
        // a second vertexCounter with nothing in between gives the same
        const long long m2 = next->m;
        vertexCounter(G, next, 0, 0, 1);
        expect(next->m == m2, "repeated vertexCounter", next->m, m2);
    }

    // ---- pushes on both sides of the small-list limit ----
    for (int k : {1, 7, 1000, 1024, 1025, 5000}) {
        std::vector<char> h((size_t)n, 0);
        std::vector<intT> list;
        for (intT v = 0; v < n && (int)list.size() < k; v++)
            if (deg(v) > 0) {
                h[v] = 1;
                list.push_back(v);
            }
        load(h);
        V->toSparse();
        expect(cur->m == (long long)list.size(), "toSparse count", cur->m, (long long)list.size());
        edgeMapSparseV3(G, V, Every_F{}, next);
        std::vector<intT> want;
        long long wdeg = 0;
        for (intT s : list)
            for (int e = off[s]; e < off[s + 1]; e++) {
                want.push_back(adj[e]);
                wdeg += deg(adj[e]);
            }
        expect(next->m == (long long)want.size(), "push successes", next->m, (long long)want.size());
        expect(next->outEdgesCount == wdeg, "push successes' out-edges", next->outEdgesCount, wdeg);
        if (next->m == (long long)want.size()) {
            std::vector<intT> got((size_t)next->m);
            toHost(got.data(), next->s, got.size());
            std::sort(got.begin(), got.end());
            std::sort(want.begin(), want.end());
            expect(got == want, "push output list", (long long)k, (long long)k);
        }
    }

    // ---- vertexFilter into a frontier ----
    {
        std::vector<char> h((size_t)n, 0);
        for (intT v = 0; v < n; v += 3) h[v] = 1;
        load(h);
        long long want = 0;
        for (intT v = 0; v < n; v++) want += h[v] && (v & 1);
        clearLocalFrontier(next, 0, 0, 1);
        next->m = 0;
        vertexFilter(V, Odd_F{}, 0, 0, 1, next);
        expect(next->m == want, "vertexFilter count", next->m, want);
        vertexCounter(G, next, 0, 0, 1);
        expect(next->m == want, "vertexFilter then vertexCounter", next->m, want);
    }

    if (fails == 0) std::printf("header_checks: ok\n");
    return fails;
}
