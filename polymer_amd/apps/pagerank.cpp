// PageRank driver on polymer.h (the iteration of numa-PageRank.C:273-335):
// fixed rounds; the active sources' p_curr[s]/outdeg(s) are summed into p_next
// (pulled per destination with the functor's reduce trio); the
// active vertices then take p_next = d * p_next + (1-d)/n and reset p_curr; the
// next frontier is the set of out-neighbours of the active vertices.  fp64
// values and atomics, as the reference.  Output: p_curr, n doubles.
#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
template <class vertex>
struct PR_F {
    double *p_curr, *p_next;
    vertex *V;
    PR_F(double *c, double *n, vertex *v) : p_curr(c), p_next(n), V(v) {}
    inline bool update(intT s, intT d) {
        p_next[d] += p_curr[s] / V[s].getOutDegree();
        return true;
    }
    inline bool updateAtomic(intT s, intT d) {
        writeAdd(&p_next[d], p_curr[s] / V[s].getOutDegree());
        return true;
    }
    // the pull form (edgeMapDenseReduce): a private sum per destination, folded once
    inline void initFunc(void *data, intT d) { *(double *)data = 0.0; }
    inline bool reduceFunc(void *data, intT s) {
        *(double *)data += p_curr[s] / (double)V[s].getOutDegree();
        return true;
    }
    inline bool combineFunc(void *data, intT d) {
        writeAdd(&p_next[d], *(double *)data);
        return true;
    }
    inline bool cond(intT d) { return true; }
};

struct PR_Vertex_F {
    double damping, added;
    double *p_next;
    inline bool operator()(intT i) {
        p_next[i] = damping * p_next[i] + added;
        return true;
    }
};

struct PR_Vertex_Reset {
    double *p_curr;
    inline bool operator()(intT i) {
        p_curr[i] = 0.0;
        return true;
    }
};

struct Fill_F {
    double *p;
    double v;
    inline bool operator()(intT i) {
        p[i] = v;
        return true;
    }
};
POLYMER_FUNCTORS_END

int main(int argc, char **argv) {
    const AppArgs a = parse_args(argc, argv);
    graph<asymmetricVertex> G = a.file.empty()
                                    ? rmatGraph<asymmetricVertex>(a.scale, a.edges, a.seed, a.symmetric)
                                    : readGraph<asymmetricVertex>(a.file.c_str(), a.symmetric, a.binary);
    const intT n = (intT)G.n;
    const double damping = 0.85;
    intT sz = n;
    double *p_curr = static_cast<double *>(mapDataArray(1, &sz, sizeof(double)));
    double *p_next = static_cast<double *>(mapDataArray(1, &sz, sizeof(double)));
    LocalFrontier *current = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *next = new LocalFrontier(newFrontierArray(n), 0, n);
    vertices *Frontier = new vertices(1);
    Frontier->registerFrontier(0, current);
    Frontier->calculateOffsets();
    fillLocalFrontier(current);
    vertexMap(Frontier, Fill_F{p_curr, 1.0 / (double)n}, 0);
    long long edges = 0;
    polymerSync();
    const double t_start = wall_ms();
    for (int t = 0; t < a.rounds; t++) {
        vertexCounter(G, Frontier->getFrontier(0), 0, 0, 1);
        edges += Frontier->getEdgeStat();
        clearLocalFrontier(next, 0, 0, 1);
        // the active sources' contributions pulled per destination (the push of
        // numa-PageRank.C:297 with its reduce trio, :98-122)
        edgeMapDenseReduce(G, Frontier, PR_F<asymmetricVertex>(p_curr, p_next, G.V), next);
        vertexMap(Frontier, PR_Vertex_F{damping, (1.0 - damping) / (double)n, p_next}, 0);
        vertexMap(Frontier, PR_Vertex_Reset{p_curr}, 0);
        std::swap(p_curr, p_next);
        switchFrontier(0, Frontier, next);
    }
    polymerSync();
    const double t_loop = wall_ms() - t_start;
    std::vector<double> r((size_t)n);
    toHost(r.data(), p_curr, (size_t)n);
    write_output(a.out, r);
    std::printf("pagerank: n %d m %ld rounds %d edges %lld\n", n, G.m, a.rounds, edges);
    std::printf("time_ms %.3f\n", t_loop);
    return 0;
}
