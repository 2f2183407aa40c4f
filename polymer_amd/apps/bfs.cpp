// BFS driver on polymer.h (numa-BFS.C:258-296): parents claimed with a CAS
// from -1, cond = "no parent yet", edgeMap switching between the dense pull and
// the sparse push by |F| + sum of out-degrees against -thr (0: every round
// dense, the reference's setting; negative: m/20).  Levels are recorded per
// round.  Output: parents (n int32) then levels (n int32; -1 = unreached).
#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
struct BFS_F {
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

struct Level_F {
    intT *Levels;
    intT round;
    inline bool operator()(intT i) {
        Levels[i] = round;
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
    if (a.source < 0 || a.source >= n) {
        std::fprintf(stderr, "source %ld out of range\n", a.source);
        return 2;
    }
    const intT src = (intT)a.source;
    const long threshold = a.threshold < 0 ? G.m / 20 : a.threshold;
    intT sz = n;
    intT *Parents = static_cast<intT *>(mapDataArray(1, &sz, sizeof(intT)));
    intT *Levels = static_cast<intT *>(mapDataArray(1, &sz, sizeof(intT)));
    std::vector<intT> init((size_t)n, -1);
    toDevice(Levels, init.data(), (size_t)n);
    init[src] = src;
    toDevice(Parents, init.data(), (size_t)n);
    const intT zero = 0;
    toDevice(Levels + src, &zero, 1);

    LocalFrontier *current = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *next = new LocalFrontier(newFrontierArray(n), 0, n);
    const bool one = true;
    toDevice(current->b + src, &one, 1);
    vertexCounter(G, current, 0, 0, 1);
    vertices *Frontier = new vertices(1);
    Frontier->registerFrontier(0, current);
    Frontier->calculateOffsets();
    Frontier->calculateNumOfNonZero(0);  // the count isEmpty / edgeMap read

    intT round = 0, dense_rounds = 0;
    long long reached = 1;
    polymerSync();
    const double t_start = wall_ms();
    while (!Frontier->isEmpty()) {
        round++;
        edgeMap(G, Frontier, BFS_F{Parents}, next, (intT)threshold, DENSE);
        if (next->isDense) {
            dense_rounds++;
            vertexCounter(G, next, 0, 0, 1);
        }
        switchFrontier(0, Frontier, next);
        Frontier->calculateNumOfNonZero(0);
        vertexMap(Frontier, Level_F{Levels, round}, 0);
        reached += Frontier->numNonzeros();
    }
    polymerSync();
    const double t_loop = wall_ms() - t_start;
    std::vector<intT> out((size_t)2 * n);
    toHost(out.data(), Parents, (size_t)n);
    toHost(out.data() + n, Levels, (size_t)n);
    write_output(a.out, out);
    std::printf("bfs: n %d m %ld source %d reached %lld rounds %d dense %d\n", n, G.m, src, reached,
                round, dense_rounds);
    std::printf("time_ms %.3f\n", t_loop);
    return 0;
}
