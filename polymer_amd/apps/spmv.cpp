// SpMV driver on polymer.h (numa-SPMV.C:168-193): every round all vertices
// push p_curr[s] * w(s,d) into p_next[d] (dense forward with writeAdd), then
// p_curr is reset and the arrays swap.  fp64 values, int32 weights; x0 = 1/n
// (the reference's start) or -in1 <n doubles>.  Output: p_curr, n doubles.
#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
struct SPMV_F {
    double *p_curr, *p_next;
    inline bool update(intT s, intT d, intE edgeLen) {
        p_next[d] += p_curr[s] * edgeLen;
        return true;
    }
    inline bool updateAtomic(intT s, intT d, intE edgeLen) {
        writeAdd(&p_next[d], p_curr[s] * edgeLen);
        return true;
    }
    inline bool cond(intT d) { return true; }
};

struct SPMV_Vertex_Reset {
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
    wghGraph<asymmetricWghVertex> G =
        a.file.empty()
            ? rmatWghGraph<asymmetricWghVertex>(a.scale, a.edges, a.seed, a.symmetric,
                                                (uint64_t)(a.wseed < 0 ? 1 : a.wseed))
            : readWghGraph<asymmetricWghVertex>(a.file.c_str(), a.symmetric, a.binary);
    const intT n = (intT)G.n;
    intT sz = n;
    double *p_curr = static_cast<double *>(mapDataArray(1, &sz, sizeof(double)));
    double *p_next = static_cast<double *>(mapDataArray(1, &sz, sizeof(double)));
    LocalFrontier *all = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *output = new LocalFrontier(newFrontierArray(n), 0, n);
    vertices *All = new vertices(1);
    All->registerFrontier(0, all);
    All->calculateOffsets();
    fillLocalFrontier(all);
    if (!a.in1.empty()) {
        const std::vector<double> x0 = read_input<double>(a.in1, (size_t)n);
        toDevice(p_curr, x0.data(), (size_t)n);
    } else {
        vertexMap(All, Fill_F{p_curr, 1.0 / (double)n}, 0);
    }
    for (int t = 0; t < a.rounds; t++) {
        clearLocalFrontier(output, 0, 0, 1);
        edgeMapDenseForward(G, All, SPMV_F{p_curr, p_next}, output);
        vertexMap(All, SPMV_Vertex_Reset{p_curr}, 0);
        std::swap(p_curr, p_next);
    }
    polymerSync();
    std::vector<double> r((size_t)n);
    toHost(r.data(), p_curr, (size_t)n);
    write_output(a.out, r);
    std::printf("spmv: n %d m %ld rounds %d\n", n, G.m, a.rounds);
    return 0;
}
