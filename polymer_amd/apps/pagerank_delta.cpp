// PageRankDelta driver on polymer.h (numa-PageRankDelta.C:63-127 functors,
// :205-330 loop): p = 0, Delta = 1/n and every vertex in the frontier.  A round:
// the frontier's sources add Delta[s]/outdeg(s) into nghSum (edgeMap
// DENSE_FORWARD; threshold 0, the reference's, keeps every round dense, -thr > 0
// lets small frontiers take the sparse push); vertexFilter over all vertices
// updates p (round 1: Delta = d (p + nghSum) + (1-d)/n - p, then Delta -= 1/n;
// later Delta = d nghSum) and keeps |Delta| > epsilon2 p as the next frontier;
// nghSum is reset.  fp64 values and atomics, as the reference.  Output: p, n
// doubles.
#include <cmath>

#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
template <class vertex>
struct DeltaPush_F {
    vertex *V;
    double *Delta, *nghSum;
    inline bool update(intT s, intT d) {
        nghSum[d] += Delta[s] / V[s].getOutDegree();
        return true;
    }
    inline bool updateAtomic(intT s, intT d) {
        writeAdd(&nghSum[d], Delta[s] / V[s].getOutDegree());
        return true;
    }
    inline bool cond(intT) { return true; }
};

// round 1 starts from p = 0 and Delta = 1/n, so the 1/n is taken back out of Delta
struct FirstRound_F {
    double *p, *Delta, *nghSum;
    double damping, teleport, one_over_n, epsilon2;
    inline bool operator()(intT i) {
        const double next = damping * (p[i] + nghSum[i]) + teleport;
        Delta[i] = next - p[i];
        p[i] = p[i] + Delta[i];
        Delta[i] = Delta[i] - one_over_n;
        return fabs(Delta[i]) > epsilon2 * p[i];
    }
};

struct LaterRound_F {
    double *p, *Delta, *nghSum;
    double damping, epsilon2;
    inline bool operator()(intT i) {
        Delta[i] = damping * nghSum[i];
        p[i] = p[i] + Delta[i];
        return fabs(Delta[i]) > epsilon2 * p[i];
    }
};

struct Set_F {
    double *x;
    double v;
    inline bool operator()(intT i) {
        x[i] = v;
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
    const double damping = 0.85, epsilon2 = 0.01;  // numa-PageRankDelta.C:444-446
    const double one_over_n = 1.0 / (double)n;
    intT sz = n;
    double *p = static_cast<double *>(mapDataArray(1, &sz, sizeof(double)));
    double *Delta = static_cast<double *>(mapDataArray(1, &sz, sizeof(double)));
    double *nghSum = static_cast<double *>(mapDataArray(1, &sz, sizeof(double)));
    LocalFrontier *current = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *all = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *next = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *dummy = new LocalFrontier(newFrontierArray(n), 0, n);
    vertices *Frontier = new vertices(1);
    Frontier->registerFrontier(0, current);
    Frontier->calculateOffsets();
    vertices *All = new vertices(1);
    All->registerFrontier(0, all);
    All->calculateOffsets();
    fillLocalFrontier(current);
    fillLocalFrontier(all);
    vertexMap(All, Set_F{p, 0.0}, 0);
    vertexMap(All, Set_F{Delta, one_over_n}, 0);
    vertexMap(All, Set_F{nghSum, 0.0}, 0);
    const long threshold = a.threshold < 0 ? 0 : a.threshold;
    long long edges = 0, active = 0;
    int sparse_rounds = 0;
    polymerSync();
    const double t_start = wall_ms();
    for (int t = 0; t < a.rounds; t++) {
        vertexCounter(G, Frontier->getFrontier(0), 0, 0, 1);
        Frontier->calculateNumOfNonZero(0);
        edges += Frontier->getEdgeStat();
        active += Frontier->numNonzeros();
        edgeMap(G, Frontier, DeltaPush_F<asymmetricVertex>{G.V, Delta, nghSum}, dummy,
                (intT)threshold, DENSE_FORWARD);
        if (!dummy->isDense) sparse_rounds++;
        if (t == 0)
            vertexFilter(All, FirstRound_F{p, Delta, nghSum, damping, (1.0 - damping) * one_over_n,
                                           one_over_n, epsilon2},
                         0, 0, 1, next);
        else
            vertexFilter(All, LaterRound_F{p, Delta, nghSum, damping, epsilon2}, 0, 0, 1, next);
        vertexMap(All, Set_F{nghSum, 0.0}, 0);
        switchFrontier(0, Frontier, next);
    }
    polymerSync();
    const double t_loop = wall_ms() - t_start;
    std::vector<double> r((size_t)n);
    toHost(r.data(), p, (size_t)n);
    write_output(a.out, r);
    std::printf("pagerank_delta: n %d m %ld rounds %d frontier %lld edges %lld sparse %d\n", n, G.m,
                a.rounds, active, edges, sparse_rounds);
    std::printf("time_ms %.3f\n", t_loop);
    return 0;
}
