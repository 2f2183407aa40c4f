// Bellman-Ford driver on polymer.h (numa-BellmanFord.C:166-211): relax
// SP[d] with writeMin(SP[s] + w), a vertex joins the next frontier once per
// round through a CAS on Visited; edgeMap dense-forward or sparse by
// |F| + sum of out-degrees against -thr (negative: m/20).  Visited is cleared
// on the frontier each round produced, so the distances are the exact
// shortest ones (the reference clears it on the round's input frontier,
// which can stop short of them).  Int32 weights.  Output: SP, n int32
// (unreached = 1073741823).
#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
struct BF_F {
    intT *SP;
    int *Visited;
    inline bool update(intT s, intT d, intE edgeLen) {
        const intT nd = SP[s] + edgeLen;
        if (nd < SP[d]) {
            SP[d] = nd;
            if (Visited[d] == 0) {
                Visited[d] = 1;
                return true;
            }
        }
        return false;
    }
    inline bool updateAtomic(intT s, intT d, intE edgeLen) {
        const intT nd = SP[s] + edgeLen;
        return writeMin(&SP[d], nd) && CAS(&Visited[d], 0, 1);
    }
    inline bool cond(intT d) { return true; }
};

struct BF_Vertex_F {
    int *Visited;
    inline bool operator()(intT i) {
        Visited[i] = 0;
        return true;
    }
};
POLYMER_FUNCTORS_END

int main(int argc, char **argv) {
    const AppArgs a = parse_args(argc, argv);
    const intT kInf = 1073741823;  // INT_MAX / 2
    wghGraph<asymmetricWghVertex> G =
        a.file.empty()
            ? rmatWghGraph<asymmetricWghVertex>(a.scale, a.edges, a.seed, a.symmetric,
                                                (uint64_t)(a.wseed < 0 ? 1 : a.wseed))
            : readWghGraph<asymmetricWghVertex>(a.file.c_str(), a.symmetric, a.binary);
    const intT n = (intT)G.n;
    if (a.source < 0 || a.source >= n) {
        std::fprintf(stderr, "source %ld out of range\n", a.source);
        return 2;
    }
    const intT src = (intT)a.source;
    const long threshold = a.threshold < 0 ? G.m / 20 : a.threshold;
    intT sz = n;
    intT *SP = static_cast<intT *>(mapDataArray(1, &sz, sizeof(intT)));
    int *Visited = static_cast<int *>(mapDataArray(1, &sz, sizeof(int)));
    std::vector<intT> init((size_t)n, kInf);
    init[src] = 0;
    toDevice(SP, init.data(), (size_t)n);

    LocalFrontier *current = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *next = new LocalFrontier(newFrontierArray(n), 0, n);
    const bool one = true;
    toDevice(current->b + src, &one, 1);
    vertexCounter(G, current, 0, 0, 1);
    vertices *Frontier = new vertices(1);
    Frontier->registerFrontier(0, current);
    Frontier->calculateOffsets();
    Frontier->calculateNumOfNonZero(0);  // the count isEmpty / edgeMap read

    intT round = 0;
    while (!Frontier->isEmpty()) {
        if (round == n) {
            std::fprintf(stderr, "negative cycle\n");
            return 1;
        }
        round++;
        edgeMap(G, Frontier, BF_F{SP, Visited}, next, (intT)threshold, DENSE_FORWARD);
        if (next->isDense) vertexCounter(G, next, 0, 0, 1);
        switchFrontier(0, Frontier, next);
        Frontier->calculateNumOfNonZero(0);
        vertexMap(Frontier, BF_Vertex_F{Visited}, 0);
    }
    polymerSync();
    std::vector<intT> out((size_t)n);
    toHost(out.data(), SP, (size_t)n);
    write_output(a.out, out);
    std::printf("bellmanford: n %d m %ld source %d rounds %d\n", n, G.m, src, round);
    return 0;
}
