// Connected components driver on polymer.h (numa-Components.C:233-281):
// labels start as the vertex ids; each round the active vertices remember
// their label, edgeMap lowers labels with writeMin and activates a vertex on
// the first decrease of the round; dense (pull) or sparse (push) by
// |F| + sum of out-degrees against -thr (negative: m/8).  Symmetric graphs.
// Output: labels, n int32 (the minimum id of each component).
#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
struct CC_F {
    intT *IDs, *prevIDs;
    inline bool update(intT s, intT d) {
        const intT orig = IDs[d];
        if (IDs[s] < orig) {
            IDs[d] = IDs[s];
            return orig == prevIDs[d];
        }
        return false;
    }
    inline bool updateAtomic(intT s, intT d) {
        const intT orig = IDs[d];
        return writeMin(&IDs[d], IDs[s]) && orig == prevIDs[d];
    }
    inline bool cond(intT d) { return true; }
};

struct CC_Vertex_F {
    intT *IDs, *prevIDs;
    inline bool operator()(intT i) {
        prevIDs[i] = IDs[i];
        return true;
    }
};

struct Id_F {
    intT *IDs;
    inline bool operator()(intT i) {
        IDs[i] = i;
        return true;
    }
};
POLYMER_FUNCTORS_END

int main(int argc, char **argv) {
    const AppArgs a = parse_args(argc, argv);
    graph<symmetricVertex> G = a.file.empty()
                                   ? rmatGraph<symmetricVertex>(a.scale, a.edges, a.seed, true)
                                   : readGraph<symmetricVertex>(a.file.c_str(), true, a.binary);
    const intT n = (intT)G.n;
    const long threshold = a.threshold < 0 ? G.m / 8 : a.threshold;
    intT sz = n;
    intT *IDs = static_cast<intT *>(mapDataArray(1, &sz, sizeof(intT)));
    intT *prevIDs = static_cast<intT *>(mapDataArray(1, &sz, sizeof(intT)));
    LocalFrontier *current = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *next = new LocalFrontier(newFrontierArray(n), 0, n);
    vertices *Frontier = new vertices(1);
    Frontier->registerFrontier(0, current);
    Frontier->calculateOffsets();
    fillLocalFrontier(current);
    vertexCounter(G, current, 0, 0, 1);
    Frontier->calculateNumOfNonZero(0);  // the count isEmpty / edgeMap read
    vertexMap(Frontier, Id_F{IDs}, 0);

    intT round = 0;
    while (!Frontier->isEmpty()) {
        round++;
        vertexMap(Frontier, CC_Vertex_F{IDs, prevIDs}, 0);
        edgeMap(G, Frontier, CC_F{IDs, prevIDs}, next, (intT)threshold, DENSE);
        if (next->isDense) vertexCounter(G, next, 0, 0, 1);
        switchFrontier(0, Frontier, next);
        Frontier->calculateNumOfNonZero(0);
    }
    polymerSync();
    std::vector<intT> out((size_t)n);
    toHost(out.data(), IDs, (size_t)n);
    write_output(a.out, out);
    std::printf("components: n %d m %ld rounds %d\n", n, G.m, round);
    return 0;
}
