// Belief propagation driver on polymer.h (numa-BP.C:89-140, 280-316): each
// round the active vertices restart their product at (1, 1), then every
// out-edge e = (s -> d) of an active s multiplies P_next[d][k] by
//     b_e[k] = sum_l phi_d[l] * W_e[k][l] * P_curr[d][l]
// (edgeMapDenseBP, e = out-offset of s + j), d joins the next frontier, and
// the products swap.  fp64 state and messages (an atomic multiply per
// message, where the reference multiplies without one).  Inputs: -in1 vertex
// potentials (2n floats), -in2 edge potentials (4m floats, CSR order, W_e[k][l]
// at 4e+2k+l), -in3 initial products (2n doubles).  Output: P_curr, 2n doubles.
#include "app_common.h"
#include "polymer.h"

POLYMER_FUNCTORS_BEGIN
struct BP_F {
    const float *vpot, *epot;
    const double *P_curr;
    double *P_next;
    const intT *out_off;
    inline bool updateAtomic(intT s, intT d, intT j) {
        const long e = (long)out_off[s] + j;
        for (int k = 0; k < 2; k++) {
            double b = 0.0;
            for (int l = 0; l < 2; l++)
                b += (double)vpot[2 * d + l] * (double)epot[4 * e + 2 * k + l] * P_curr[2 * d + l];
            double *p = &P_next[2 * d + k];
            double c = *p;
            while (!CAS(p, c, c * b)) c = *(volatile double *)p;
        }
        return true;
    }
    inline bool cond(intT d) { return true; }
};

struct BP_Vertex_Reset {
    double *P_next;
    inline bool operator()(intT i) {
        P_next[2 * i] = 1.0;
        P_next[2 * i + 1] = 1.0;
        return true;
    }
};
POLYMER_FUNCTORS_END

int main(int argc, char **argv) {
    const AppArgs a = parse_args(argc, argv);
    if (a.in1.empty() || a.in2.empty() || a.in3.empty()) {
        std::fprintf(stderr, "bp needs -in1 <vertex potentials> -in2 <edge potentials> -in3 <products>\n");
        return 2;
    }
    graph<asymmetricVertex> G = a.file.empty()
                                    ? rmatGraph<asymmetricVertex>(a.scale, a.edges, a.seed, a.symmetric)
                                    : readGraph<asymmetricVertex>(a.file.c_str(), a.symmetric, a.binary);
    const intT n = (intT)G.n;
    const long m = G.m;
    const std::vector<float> vp = read_input<float>(a.in1, (size_t)2 * n);
    const std::vector<float> ep = read_input<float>(a.in2, (size_t)4 * m);
    const std::vector<double> p0 = read_input<double>(a.in3, (size_t)2 * n);
    intT sz2 = 2 * n, sz4 = (intT)(4 * m);
    float *vpot = static_cast<float *>(mapDataArray(1, &sz2, sizeof(float)));
    float *epot = static_cast<float *>(mapDataArray(1, &sz4, sizeof(float)));
    double *P_curr = static_cast<double *>(mapDataArray(1, &sz2, sizeof(double)));
    double *P_next = static_cast<double *>(mapDataArray(1, &sz2, sizeof(double)));
    toDevice(vpot, vp.data(), vp.size());
    toDevice(epot, ep.data(), ep.size());
    toDevice(P_curr, p0.data(), p0.size());

    LocalFrontier *current = new LocalFrontier(newFrontierArray(n), 0, n);
    LocalFrontier *next = new LocalFrontier(newFrontierArray(n), 0, n);
    vertices *Frontier = new vertices(1);
    Frontier->registerFrontier(0, current);
    Frontier->calculateOffsets();
    fillLocalFrontier(current);
    for (int t = 0; t < a.rounds; t++) {
        clearLocalFrontier(next, 0, 0, 1);
        vertexMap(Frontier, BP_Vertex_Reset{P_next}, 0);
        edgeMapDenseBP(G, Frontier, BP_F{vpot, epot, P_curr, P_next, G.view.out_off}, next);
        std::swap(P_curr, P_next);
        switchFrontier(0, Frontier, next);
    }
    polymerSync();
    std::vector<double> r((size_t)2 * n);
    toHost(r.data(), P_curr, r.size());
    write_output(a.out, r);
    std::printf("bp: n %d m %ld rounds %d\n", n, m, a.rounds);
    return 0;
}
