/*
 * polymer.h -- the edgeMap / vertexMap programming interface of Polymer on one
 * MI355X (gfx950).  Header-only C++ compiled by hipcc over the device graphs of
 * libpolymer_gpu.so: an app keeps the reference's types and calls
 *
 *     graph<vertex>, wghGraph<vertex>           graph.h:11-164
 *     LocalFrontier, vertices (vertexSubset)    polymer.h:527-881
 *     edgeMap, edgeMapDense, edgeMapDenseForward, edgeMapDenseReduce,
 *     edgeMapSparseV3, edgeMapDenseBP           polymer.h:924-1306, 1928-2003, 2241-2309
 *     vertexMap, vertexFilter, vertexCounter,
 *     clearLocalFrontier, switchFrontier        polymer.h:2224-2441
 *     CAS, writeMin, writeAdd                   utils.h:244-310
 *     mapDataArray                              polymer.h:499-519
 *
 * and its functors (update / updateAtomic / cond, the weighted (s, d, w) and
 * BP (s, d, j) forms, the initFunc / reduceFunc / combineFunc trio) run as
 * device code.  Functor classes written between POLYMER_FUNCTORS_BEGIN and
 * POLYMER_FUNCTORS_END (clang's force_cuda_host_device region) need no HIP
 * annotations: the reference's functor code compiles for the GPU as it is.
 *
 * What differs from the NUMA engine, by design:
 *   - one GPU replaces the NUMA nodes: every call is made once (the node /
 *     sub-worker arguments are accepted; only node 0, sub-worker 0 launches);
 *     a call returns when its kernels have been queued on the library stream
 *     (hipStream of the graph), and counts the host reads are synchronised;
 *   - value arrays from mapDataArray and frontier arrays from
 *     newFrontierArray are device memory: initialise them with vertexMap or
 *     copy with toDevice / toHost, never through host loops;
 *   - vertex ids are the ids of the input graph (the library's relabelling
 *     stays inside libpolymer_gpu.so).
 * The dense pull keeps the reference's per-destination order: one thread
 * scans a destination's in-neighbours in id order, calls updateAtomic and
 * stops when cond turns false (edgeMapDense polymer.h:924-968).
 */
#ifndef POLYMER_HIP_H
#define POLYMER_HIP_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "polymer_gpu.h"

typedef int intT;
typedef unsigned int uintT;
typedef int intE;
typedef unsigned int uintE;

#define PG_HD __host__ __device__

namespace polymer_detail {

inline void check(int rc, const char *what) {
    if (rc != PG_OK) {
        char msg[512];
        pg_last_error(msg, sizeof(msg));
        std::fprintf(stderr, "polymer: %s failed (%d): %s\n", what, rc, msg);
        std::abort();
    }
}

inline void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "polymer: %s failed: %s\n", what, hipGetErrorString(e));
        std::abort();
    }
}

// the stream of the library (set by the first graph; every call is ordered on it)
inline hipStream_t &lib_stream() {
    static hipStream_t s = nullptr;
    return s;
}

inline void ensure_init() {
    static bool done = false;
    if (!done) {
        check(pg_init(0), "pg_init");
        done = true;
    }
}

inline unsigned blocks_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (unsigned)b;
}

template <class T>
T *dev_alloc(size_t count) {
    void *p = nullptr;
    check_hip(hipMalloc(&p, (count ? count : 1) * sizeof(T)), "hipMalloc");
    return static_cast<T *>(p);
}

}  // namespace polymer_detail

/* ---- copies between host and the device arrays of an app ----------------- */
template <class T>
void toDevice(T *dev, const T *host, size_t count) {
    polymer_detail::check_hip(hipMemcpyAsync(dev, host, count * sizeof(T), hipMemcpyHostToDevice,
                                             polymer_detail::lib_stream()),
                              "hipMemcpyAsync");
    polymer_detail::check_hip(hipStreamSynchronize(polymer_detail::lib_stream()), "sync");
}
template <class T>
void toHost(T *host, const T *dev, size_t count) {
    polymer_detail::check_hip(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost,
                                             polymer_detail::lib_stream()),
                              "hipMemcpyAsync");
    polymer_detail::check_hip(hipStreamSynchronize(polymer_detail::lib_stream()), "sync");
}

/* ---- atomics (utils.h:244-310) --------------------------------------------- */
template <class ET>
PG_HD inline bool CAS(ET *ptr, ET oldv, ET newv) {
    static_assert(sizeof(ET) == 4 || sizeof(ET) == 8, "CAS on 4- or 8-byte values");
#if __HIP_DEVICE_COMPILE__
    if constexpr (sizeof(ET) == 4) {
        unsigned o = __builtin_bit_cast(unsigned, oldv), nv = __builtin_bit_cast(unsigned, newv);
        return atomicCAS(reinterpret_cast<unsigned *>(ptr), o, nv) == o;
    } else {
        unsigned long long o = __builtin_bit_cast(unsigned long long, oldv),
                           nv = __builtin_bit_cast(unsigned long long, newv);
        return atomicCAS(reinterpret_cast<unsigned long long *>(ptr), o, nv) == o;
    }
#else
    return __atomic_compare_exchange(ptr, &oldv, &newv, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
#endif
}

// *a = min(*a, b); true if this call lowered it
template <class ET>
PG_HD inline bool writeMin(ET *a, ET b) {
#if __HIP_DEVICE_COMPILE__
    if constexpr (std::is_integral<ET>::value && std::is_signed<ET>::value && sizeof(ET) == 4) {
        return atomicMin(reinterpret_cast<int *>(a), (int)b) > (int)b;
    } else {
        ET c = *a;
        while (c > b) {
            if (CAS(a, c, b)) return true;
            c = *(volatile ET *)a;
        }
        return false;
    }
#else
    ET c = *a;
    while (c > b) {
        if (CAS(a, c, b)) return true;
        c = *(volatile ET *)a;
    }
    return false;
#endif
}

template <class ET>
PG_HD inline void writeAdd(ET *a, ET b) {
#if __HIP_DEVICE_COMPILE__
    atomicAdd(a, b);
#else
    ET c = *a;
    while (!CAS(a, c, (ET)(c + b))) c = *(volatile ET *)a;
#endif
}

/* ---- vertices of the device graph (graph.h:11-164) -------------------------- */
struct symmetricVertex {
    const intE *neighbors;
    intT degree;
    PG_HD intE getInNeighbor(intT j) const { return neighbors[j]; }
    PG_HD intE getOutNeighbor(intT j) const { return neighbors[j]; }
    PG_HD const intE *getInNeighborPtr() const { return neighbors; }
    PG_HD const intE *getOutNeighborPtr() const { return neighbors; }
    PG_HD intT getInDegree() const { return degree; }
    PG_HD intT getOutDegree() const { return degree; }
    PG_HD intT getFakeInDegree() const { return degree; }
    PG_HD intT getFakeDegree() const { return degree; }
};

struct asymmetricVertex {
    const intE *inNeighbors, *outNeighbors;
    intT inDegree, outDegree;
    PG_HD intE getInNeighbor(intT j) const { return inNeighbors[j]; }
    PG_HD intE getOutNeighbor(intT j) const { return outNeighbors[j]; }
    PG_HD const intE *getInNeighborPtr() const { return inNeighbors; }
    PG_HD const intE *getOutNeighborPtr() const { return outNeighbors; }
    PG_HD intT getInDegree() const { return inDegree; }
    PG_HD intT getOutDegree() const { return outDegree; }
    PG_HD intT getFakeInDegree() const { return inDegree; }
    PG_HD intT getFakeDegree() const { return outDegree; }
};

// weighted forms: weights are separate arrays aligned with the neighbour lists
struct symmetricWghVertex : symmetricVertex {
    const intE *weights;
    PG_HD intE getInWeight(intT j) const { return weights[j]; }
    PG_HD intE getOutWeight(intT j) const { return weights[j]; }
};

struct asymmetricWghVertex : asymmetricVertex {
    const intE *inWeights, *outWeights;
    PG_HD intE getInWeight(intT j) const { return inWeights[j]; }
    PG_HD intE getOutWeight(intT j) const { return outWeights[j]; }
};

namespace polymer_detail {

template <class vertex>
__global__ void k_fill_vertices(pg_view v, vertex *V) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        vertex x{};
        const int ob = v.out_off[i], od = v.out_off[i + 1] - ob;
        const int ib = v.in_off[i], id = v.in_off[i + 1] - ib;
        if constexpr (std::is_base_of<asymmetricVertex, vertex>::value) {
            x.outNeighbors = v.out_adj + ob;
            x.inNeighbors = v.in_adj + ib;
            x.outDegree = od;
            x.inDegree = id;
            if constexpr (std::is_same<vertex, asymmetricWghVertex>::value) {
                x.outWeights = static_cast<const intE *>(v.out_w) + ob;
                x.inWeights = static_cast<const intE *>(v.in_w) + ib;
            }
        } else {
            x.neighbors = v.out_adj + ob;
            x.degree = od;
            if constexpr (std::is_same<vertex, symmetricWghVertex>::value)
                x.weights = static_cast<const intE *>(v.out_w) + ob;
        }
        V[i] = x;
    }
}

}  // namespace polymer_detail

/* ---- graphs ------------------------------------------------------------------ */
// graph<vertex> (graph.h:61-87): V is a device array of n vertices; the
// arrays behind it belong to the library graph handle.
template <class vertex>
struct graph {
    vertex *V = nullptr;
    long n = 0;
    long m = 0;
    pg_graph *handle = nullptr;
    pg_view view{};
    void del() {
        if (V) (void)hipFree(V);
        if (handle) pg_graph_destroy(handle);
        V = nullptr;
        handle = nullptr;
    }
};

template <class vertex>
struct wghGraph : graph<vertex> {};

namespace polymer_detail {

template <class G>
G wrap(pg_graph *h, bool weighted) {
    G g;
    g.handle = h;
    check(pg_graph_view(h, &g.view), "pg_graph_view");
    lib_stream() = static_cast<hipStream_t>(g.view.stream);
    if (weighted && g.view.w_dtype != PG_W_I32) {
        std::fprintf(stderr, "polymer: weighted graph without int32 weights\n");
        std::abort();
    }
    g.n = (long)g.view.n;
    g.m = (long)g.view.m;
    using vertex = typename std::remove_pointer<decltype(g.V)>::type;
    g.V = dev_alloc<vertex>((size_t)g.n);
    hipLaunchKernelGGL(k_fill_vertices<vertex>, dim3(blocks_for(g.n)), dim3(256), 0, lib_stream(),
                       g.view, g.V);
    check_hip(hipGetLastError(), "k_fill_vertices");
    check_hip(hipStreamSynchronize(lib_stream()), "sync");
    return g;
}

}  // namespace polymer_detail

// readGraph (IO-numa.h:73-190, 543-635): a PBBS AdjacencyGraph /
// WeightedAdjacencyGraph text file, or with binary the <fname>.config / .adj / .idx
// triple; isSymmetric marks an already symmetric graph (-s).
template <class vertex>
graph<vertex> readGraph(const char *fname, bool isSymmetric, bool binary = false) {
    polymer_detail::ensure_init();
    pg_graph *h = nullptr;
    const uint32_t flags = isSymmetric ? PG_GRAPH_SYMMETRIC : 0u;
    if (binary)
        polymer_detail::check(pg_graph_load_bin(fname, flags, &h), "pg_graph_load_bin");
    else
        polymer_detail::check(pg_graph_load_adj(fname, flags, &h), "pg_graph_load_adj");
    return polymer_detail::wrap<graph<vertex>>(h, false);
}

// readWghGraph (IO.h:186-272, IO-numa.h:637-720); a binary graph has unit weights
template <class vertex>
wghGraph<vertex> readWghGraph(const char *fname, bool isSymmetric, bool binary = false) {
    polymer_detail::ensure_init();
    pg_graph *h = nullptr;
    const uint32_t flags = isSymmetric ? PG_GRAPH_SYMMETRIC : 0u;
    if (binary)
        polymer_detail::check(pg_graph_load_bin(fname, flags | PG_GRAPH_WEIGHT_I32, &h),
                              "pg_graph_load_bin");
    else
        polymer_detail::check(pg_graph_load_adj(fname, flags, &h), "pg_graph_load_adj");
    return polymer_detail::wrap<wghGraph<vertex>>(h, true);
}

// A synthetic rMat graph (a/b/c = 0.57/0.19/0.19) generated on the device; the
// weighted form carries int32 weights U[1,255] (no counterpart in the reference,
// which reads files only).
template <class vertex>
graph<vertex> rmatGraph(int scale, long m, uint64_t seed, bool symmetrize) {
    polymer_detail::ensure_init();
    pg_graph *h = nullptr;
    polymer_detail::check(pg_graph_rmat(scale, m, 0.57, 0.19, 0.19, seed,
                                        symmetrize ? PG_GRAPH_SYMMETRIZE : 0u, 0, &h),
                          "pg_graph_rmat");
    return polymer_detail::wrap<graph<vertex>>(h, false);
}

template <class vertex>
wghGraph<vertex> rmatWghGraph(int scale, long m, uint64_t seed, bool symmetrize, uint64_t wseed) {
    polymer_detail::ensure_init();
    pg_graph *h = nullptr;
    polymer_detail::check(
        pg_graph_rmat(scale, m, 0.57, 0.19, 0.19, seed,
                      PG_GRAPH_WEIGHT_I32 | (symmetrize ? PG_GRAPH_SYMMETRIZE : 0u), wseed, &h),
        "pg_graph_rmat");
    return polymer_detail::wrap<wghGraph<vertex>>(h, true);
}

/* ---- value arrays (polymer.h:499-519) ------------------------------------------ */
// Device memory for sum(sizeArr[0..numOfShards)) elements of sizeOfOneEle bytes.
inline void *mapDataArray(int numOfShards, intT *sizeArr, int sizeOfOneEle) {
    polymer_detail::ensure_init();
    size_t total = 0;
    for (int i = 0; i < numOfShards; i++) total += (size_t)sizeArr[i];
    void *p = nullptr;
    polymer_detail::check_hip(hipMalloc(&p, (total ? total : 1) * (size_t)sizeOfOneEle), "hipMalloc");
    // zeroed on the library stream, where every later use of the array is ordered (a
    // null-stream memset is not ordered with that non-blocking stream and could land
    // after the app's first writes)
    polymer_detail::check_hip(hipMemsetAsync(p, 0, (total ? total : 1) * (size_t)sizeOfOneEle,
                                             polymer_detail::lib_stream()),
                              "hipMemsetAsync");
    polymer_detail::check_hip(hipStreamSynchronize(polymer_detail::lib_stream()), "sync");
    return p;
}

inline void freeDataArray(void *p) {
    if (p) (void)hipFree(p);
}

// A device bool array of n entries, all false (the frontier storage the
// reference takes from numa_alloc_local).
inline bool *newFrontierArray(intT n) {
    intT sz = n;
    return static_cast<bool *>(mapDataArray(1, &sz, (int)sizeof(bool)));
}

/* ---- frontiers (polymer.h:527-639, 642-881) ------------------------------------ */
namespace polymer_detail {

// ordered compaction of b[0, n) into s (ascending ids + offset); count -> *m
__global__ void k_count_flags(const bool *b, int64_t n, int64_t per_block, unsigned *counts) {
    __shared__ unsigned part[256];
    const int64_t base = (int64_t)blockIdx.x * per_block;
    unsigned c = 0;
    for (int64_t i = base + threadIdx.x; i < base + per_block && i < n; i += blockDim.x) c += b[i] ? 1u : 0u;
    part[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = part[0];
}

__global__ void k_write_flags(const bool *b, int64_t n, int64_t per_block, const unsigned *starts,
                              intT offset, intT *s) {
    __shared__ unsigned scan[256];
    const int64_t base = (int64_t)blockIdx.x * per_block;
    unsigned out = starts[blockIdx.x];
    for (int64_t c0 = base; c0 < base + per_block && c0 < n; c0 += blockDim.x) {
        const int64_t i = c0 + threadIdx.x;
        const bool f = i < n && i < base + per_block && b[i];
        scan[threadIdx.x] = f ? 1u : 0u;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {  // inclusive scan
            const unsigned add = (int)threadIdx.x >= o ? scan[threadIdx.x - o] : 0u;
            __syncthreads();
            scan[threadIdx.x] += add;
            __syncthreads();
        }
        if (f) s[out + scan[threadIdx.x] - 1] = (intT)i + offset;
        out += scan[255];
        __syncthreads();
    }
}

__global__ void k_set_list(bool *b, const intT *s, int64_t m, intT offset) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        b[s[i] - offset] = true;
}

template <class vertex>
__global__ void k_counter(const bool *b, int64_t n, intT offset, const vertex *V,
                          unsigned long long *out) {
    unsigned long long c = 0, e = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (b[i]) {
            c++;
            e += (unsigned long long)V[i + offset].getOutDegree();
        }
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o, 64);
        e += __shfl_xor(e, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && (c || e)) {
        atomicAdd(&out[0], c);
        atomicAdd(&out[1], e);
    }
}

}  // namespace polymer_detail

struct LocalFrontier {
    intT n;
    intT m;
    intT outEdgesCount;
    intT startID;
    intT endID;
    bool *b;          // device, endID - startID entries
    intT *s;          // device list (valid when !isDense), capacity sCap
    size_t sCap;
    bool isDense;
    unsigned long long *counters;  // device scratch: [0] count, [1] edges

    LocalFrontier(bool *_b, intT start, intT end)
        : n(end - start), m(0), outEdgesCount(0), startID(start), endID(end), b(_b), s(nullptr),
          sCap(0), isDense(true), counters(nullptr) {
        counters = polymer_detail::dev_alloc<unsigned long long>(4);
    }
    ~LocalFrontier() {
        if (s) (void)hipFree(s);
        if (counters) (void)hipFree(counters);
    }
    LocalFrontier(const LocalFrontier &) = delete;
    LocalFrontier &operator=(const LocalFrontier &) = delete;

    bool inRange(intT index) const { return startID <= index && index < endID; }
    // device-side bit access (b is device memory)
    PG_HD void setBit(intT index, bool val) { b[index - startID] = val; }
    PG_HD bool getBit(intT index) const { return b[index - startID]; }

    void reserveSparse(size_t cap) {
        if (cap <= sCap) return;
        if (s) (void)hipFree(s);
        s = polymer_detail::dev_alloc<intT>(cap);
        sCap = cap;
    }

    // packIndex (polymer.h:552-567): ascending ids of the set bits
    void toSparse() {
        if (isDense) {
            using namespace polymer_detail;
            reserveSparse((size_t)n + 1);
            const int64_t per_block = 8192;
            const int64_t nb = (n + per_block - 1) / per_block;
            std::vector<unsigned> cnt((size_t)(nb > 0 ? nb : 1));
            unsigned *d = dev_alloc<unsigned>(cnt.size() + 1);
            if (nb > 0) {
                hipLaunchKernelGGL(k_count_flags, dim3((unsigned)nb), dim3(256), 0, lib_stream(), b,
                                   (int64_t)n, per_block, d);
                toHost(cnt.data(), d, (size_t)nb);
            }
            unsigned tot = 0;
            for (auto &c : cnt) {
                const unsigned x = c;
                c = tot;
                tot += x;
            }
            if (nb > 0) {
                toDevice(d, cnt.data(), (size_t)nb);
                hipLaunchKernelGGL(k_write_flags, dim3((unsigned)nb), dim3(256), 0, lib_stream(), b,
                                   (int64_t)n, per_block, d, startID, s);
                check_hip(hipGetLastError(), "k_write_flags");
            }
            check_hip(hipStreamSynchronize(lib_stream()), "sync");
            (void)hipFree(d);
            m = (intT)tot;
        }
        isDense = false;
    }

    // toDense (polymer.h:594-601)
    void toDense() {
        if (!isDense) {
            using namespace polymer_detail;
            check_hip(hipMemsetAsync(b, 0, (size_t)n, lib_stream()), "hipMemsetAsync");
            if (m > 0)
                hipLaunchKernelGGL(k_set_list, dim3(blocks_for(m)), dim3(256), 0, lib_stream(), b, s,
                                   (int64_t)m, startID);
            check_hip(hipGetLastError(), "k_set_list");
        }
        isDense = true;
    }

    void setSparse(intT _m, intT *_s) {  // takes a device list
        if (s) (void)hipFree(s);
        m = _m;
        s = _s;
        sCap = (size_t)_m;
        isDense = false;
    }

    bool *swapBitVector(bool *newB) {
        bool *t = b;
        b = newB;
        return t;
    }

    void clearFrontier() {
        m = 0;
        outEdgesCount = 0;
    }
};

struct vertices {
    intT n, m;
    int numOfNodes;
    intT numOfVertices;
    std::vector<intT> numOfVertexOnNode, offsets, numOfNonZero;
    std::vector<LocalFrontier *> frontiers, nextFrontiers;
    bool isDense;
    bool firstSparse;

    explicit vertices(int _numOfNodes)
        : n(0), m(-1), numOfNodes(_numOfNodes), numOfVertices(0),
          numOfVertexOnNode(_numOfNodes, 0), offsets(_numOfNodes + 1, 0),
          numOfNonZero(_numOfNodes, 0), frontiers(_numOfNodes, nullptr),
          nextFrontiers(_numOfNodes, nullptr), isDense(true), firstSparse(false) {}

    void registerFrontier(int nodeNum, LocalFrontier *frontier) {
        frontiers[nodeNum] = frontier;
        numOfVertexOnNode[nodeNum] = frontier->n;
        isDense = frontier->isDense;
    }
    void calculateOffsets() {
        numOfVertices = 0;
        for (int i = 0; i < numOfNodes; i++) numOfVertices += numOfVertexOnNode[i];
        offsets[0] = 0;
        for (int i = 1; i <= numOfNodes; i++) offsets[i] = offsets[i - 1] + numOfVertexOnNode[i - 1];
        n = numOfVertices;
    }
    intT getSize(int nodeNum) const { return numOfVertexOnNode[nodeNum]; }
    intT getOffset(int nodeNum) const { return offsets[nodeNum]; }
    intT getSparseSize(int nodeNum) const { return numOfNonZero[nodeNum]; }
    void calculateNumOfNonZero(int nodeNum) { numOfNonZero[nodeNum] = frontiers[nodeNum]->m; }
    intT numNonzeros() {
        if (m < 0) {
            intT sum = 0;
            for (int i = 0; i < numOfNodes; i++) sum += frontiers[i]->m;
            m = sum;
        }
        return m;
    }
    bool isEmpty() { return numNonzeros() == 0; }
    long long getEdgeStat() const {
        long long sum = 0;
        for (int i = 0; i < numOfNodes; i++) sum += frontiers[i]->outEdgesCount;
        return sum;
    }
    void toDense() {
        if (isDense) return;
        isDense = true;
        for (auto *f : frontiers) f->toDense();
    }
    void toSparse() {
        if (!isDense) {
            firstSparse = false;
            return;
        }
        isDense = false;
        firstSparse = true;
        for (auto *f : frontiers) f->toSparse();
    }
    bool *getArr(int nodeNum) { return frontiers[nodeNum]->b; }
    bool *getNextArr(int nodeNum) { return nextFrontiers[nodeNum] ? nextFrontiers[nodeNum]->b : nullptr; }
    intT *getSparseArr(int nodeNum) { return frontiers[nodeNum]->s; }
    LocalFrontier *getFrontier(int nodeNum) { return frontiers[nodeNum]; }
    LocalFrontier *swapFrontier(int nodeNum, LocalFrontier *newOne) {
        if (nodeNum == 0) isDense = newOne->isDense;
        LocalFrontier *old = frontiers[nodeNum];
        frontiers[nodeNum] = newOne;
        m = -1;
        return old;
    }
    void del() {}
};

// Subworker_Partitioner (polymer.h:75-104): the SPMD roles of the NUMA engine.
// On the GPU there is one caller; the object keeps the interface.
struct Subworker_Partitioner {
    int tid = 0, subTid = 0, numOfSub = 1;
    intT dense_start = 0, dense_end = 0;
    void *global_barr = nullptr, *local_barr = nullptr;
    explicit Subworker_Partitioner(int nSub) : numOfSub(nSub) {}
    bool isMaster() const { return tid + subTid == 0; }
    bool isSubMaster() const { return subTid == 0; }
    void globalWait() {}
    void localWait() {}
    intT getStartPos(intT m) const { return subTid * (m / numOfSub); }
    intT getEndPos(intT m) const { return subTid == numOfSub - 1 ? m : (subTid + 1) * (m / numOfSub); }
};

inline Subworker_Partitioner &polymerDummyPartitioner() {
    static Subworker_Partitioner p(1);
    return p;
}

// edgeMap options (polymer.h:917-919)
enum options { DENSE, DENSE_PARALLEL, DENSE_FORWARD };

/* ---- edge kernels ---------------------------------------------------------------- */
namespace polymer_detail {

template <class G>
struct is_weighted : std::false_type {};
template <class V>
struct is_weighted<wghGraph<V>> : std::true_type {};

template <bool W, class F>
__device__ __forceinline__ bool call_update_atomic(F &f, intT s, intT d, const intE *w, int64_t e) {
    if constexpr (W)
        return f.updateAtomic(s, d, w[e]);
    else
        return f.updateAtomic(s, d);
}

// Push over sources [0, n) (dense: bit set) or list[0, count) (sparse).  A block
// takes 256 sources, scans their out-degrees in LDS and spreads its edges over
// its threads.  Mode 0: set next bits (edgeMapDenseForward); 1: append to the
// next list (edgeMapSparseV3); 2: set next bits, functor gets the edge's
// index in the source's list (edgeMapDenseBP).
template <bool W, int Mode, class F>
__global__ __launch_bounds__(256) void k_push(F f, pg_view v, const bool *cur, const intT *list,
                                              int64_t count, bool *next, intT *out_list,
                                              unsigned long long *counters) {
    __shared__ int64_t start[257];
    __shared__ intT srcs[256];
    const int64_t base = (int64_t)blockIdx.x * 256;
    const int t = threadIdx.x;
    intT s = -1;
    int64_t deg = 0;
    if (base + t < count) {
        s = list ? list[base + t] : (intT)(base + t);
        if (list || cur[s]) deg = v.out_off[s + 1] - v.out_off[s];
        else s = -1;
    }
    srcs[t] = s;
    start[t + 1] = deg;
    if (t == 0) start[0] = 0;
    __syncthreads();
    if (t == 0)
        for (int i = 1; i <= 256; i++) start[i] += start[i - 1];
    __syncthreads();
    const int64_t total = start[256];
    const intE *w = W ? static_cast<const intE *>(v.out_w) : nullptr;
    for (int64_t k = t; k < total; k += 256) {
        int lo = 0, hi = 255;  // last i with start[i] <= k
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= k) lo = mid; else hi = mid - 1;
        }
        const intT src = srcs[lo];
        const int64_t j = k - start[lo];
        const int64_t e = v.out_off[src] + j;
        const intT ngh = v.out_adj[e];
        bool act;
        if constexpr (Mode == 2)
            act = f.cond(ngh) && f.updateAtomic(src, ngh, (intT)j);
        else
            act = f.cond(ngh) && call_update_atomic<W>(f, src, ngh, w, e);
        if (act) {
            if constexpr (Mode == 1) {
                const unsigned long long pos = atomicAdd(&counters[0], 1ull);
                out_list[pos] = ngh;
                atomicAdd(&counters[1], (unsigned long long)(v.out_off[ngh + 1] - v.out_off[ngh]));
            } else {
                next[ngh] = true;
            }
        }
    }
}

// Dense pull in the reference's order: one thread per destination d with
// cond(d) scans its in-neighbours, calls updateAtomic for the active ones, marks
// d in next on success and stops when cond(d) turns false.
template <bool W, class F>
__global__ void k_pull(F f, pg_view v, const bool *cur, bool *next) {
    const intE *w = W ? static_cast<const intE *>(v.in_w) : nullptr;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const intT d = (intT)i;
        if (!f.cond(d)) continue;
        for (int64_t e = v.in_off[d]; e < v.in_off[d + 1]; e++) {
            const intT ngh = v.in_adj[e];
            if (cur[ngh] && call_update_atomic<W>(f, ngh, d, w, e)) next[d] = true;
            if (!f.cond(d)) break;
        }
    }
}

// edgeMapDenseReduce (polymer.h:1097-1168): initFunc / reduceFunc per active
// in-neighbour / combineFunc when the destination has in-edges.
template <bool W, class F>
__global__ void k_pull_reduce(F f, pg_view v, const bool *cur, bool *next) {
    const intE *w = W ? static_cast<const intE *>(v.in_w) : nullptr;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const intT d = (intT)i;
        if (!f.cond(d)) continue;
        double data[2];
        f.initFunc((void *)data, d);
        const int64_t b = v.in_off[d], e1 = v.in_off[d + 1];
        for (int64_t e = b; e < e1; e++) {
            const intT ngh = v.in_adj[e];
            bool r = false;
            if (cur[ngh]) {
                if constexpr (W) r = f.reduceFunc((void *)data, ngh, w[e]);
                else r = f.reduceFunc((void *)data, ngh);
            }
            if (r) next[d] = true;
            if (!f.cond(d)) break;
        }
        if (e1 > b) f.combineFunc((void *)data, d);
    }
}

template <class F>
__global__ void k_vertex_map_dense(F f, const bool *b, int64_t n, intT offset) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (b[i]) f((intT)i + offset);
}

template <class F>
__global__ void k_vertex_map_sparse(F f, const intT *s, int64_t m) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        f(s[i]);
}

template <class F>
__global__ void k_vertex_filter(F f, const bool *b, int64_t n, intT offset, bool *dst,
                                unsigned long long *count) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        bool r = false;
        if (b[i]) r = f((intT)i + offset);
        dst[i] = r;
        c += r ? 1ull : 0ull;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

inline void launched(const char *what) { check_hip(hipGetLastError(), what); }

// one 256-thread block per 256 sources
inline unsigned push_blocks(int64_t count) { return (unsigned)(count > 0 ? (count + 255) / 256 : 1); }

}  // namespace polymer_detail

/* ---- edge maps ---------------------------------------------------------------------- */
// edgeMapDenseForward (polymer.h:970-1032): for every active s and out-edge
// (s, ngh): if cond(ngh) && updateAtomic(s, ngh) then ngh joins next.
template <class F, class G>
bool *edgeMapDenseForward(G &GA, vertices *frontier, F f, LocalFrontier *next, bool part = false,
                          intT start = 0, intT end = 0) {
    using namespace polymer_detail;
    (void)part; (void)start; (void)end;
    constexpr bool W = is_weighted<G>::value;
    hipLaunchKernelGGL((k_push<W, 0, F>), dim3(push_blocks(GA.n)), dim3(256), 0, lib_stream(), f,
                       GA.view, (const bool *)frontier->getArr(0), (const intT *)nullptr,
                       (int64_t)GA.n, next->b, (intT *)nullptr, next->counters);
    launched("edgeMapDenseForward");
    return nullptr;
}

// edgeMapDenseBP (polymer.h:1263-1306): the forward push with the edge's index
// j in the source's out-list passed to updateAtomic(s, ngh, j).
template <class F, class G>
bool *edgeMapDenseBP(G &GA, vertices *frontier, F f, LocalFrontier *next, bool part = false,
                     intT start = 0, intT end = 0) {
    using namespace polymer_detail;
    (void)part; (void)start; (void)end;
    hipLaunchKernelGGL((k_push<false, 2, F>), dim3(push_blocks(GA.n)), dim3(256), 0, lib_stream(), f,
                       GA.view, (const bool *)frontier->getArr(0), (const intT *)nullptr,
                       (int64_t)GA.n, next->b, (intT *)nullptr, next->counters);
    launched("edgeMapDenseBP");
    return nullptr;
}

// edgeMapDense / edgeMapDenseDynamic (polymer.h:924-968, 1170-1261)
template <class F, class G>
bool *edgeMapDense(G &GA, vertices *frontier, F f, LocalFrontier *next, bool parallel = 0,
                   Subworker_Partitioner &subworker = polymerDummyPartitioner()) {
    using namespace polymer_detail;
    (void)parallel; (void)subworker;
    constexpr bool W = is_weighted<G>::value;
    frontier->nextFrontiers[0] = next;
    hipLaunchKernelGGL((k_pull<W, F>), dim3(blocks_for(GA.n)), dim3(256), 0, lib_stream(), f,
                       GA.view, (const bool *)frontier->getArr(0), next->b);
    launched("edgeMapDense");
    return nullptr;
}

// edgeMapDenseReduce (polymer.h:1097-1168)
template <class F, class G>
bool *edgeMapDenseReduce(G &GA, vertices *frontier, F f, LocalFrontier *next, bool parallel = 0,
                         Subworker_Partitioner &subworker = polymerDummyPartitioner()) {
    using namespace polymer_detail;
    (void)parallel; (void)subworker;
    constexpr bool W = is_weighted<G>::value;
    frontier->nextFrontiers[0] = next;
    hipLaunchKernelGGL((k_pull_reduce<W, F>), dim3(blocks_for(GA.n)), dim3(256), 0, lib_stream(),
                       f, GA.view, (const bool *)frontier->getArr(0), next->b);
    launched("edgeMapDenseReduce");
    return nullptr;
}

// edgeMapSparseV3 (polymer.h:1928-2003): the active list pushes; each success
// appends ngh to next's list (duplicates possible, as in the reference) and
// adds outdeg(ngh) to next->outEdgesCount.
template <class F, class G>
void edgeMapSparseV3(G &GA, vertices *frontier, F f, LocalFrontier *next, bool part = false,
                     Subworker_Partitioner &subworker = polymerDummyPartitioner()) {
    using namespace polymer_detail;
    (void)part; (void)subworker;
    constexpr bool W = is_weighted<G>::value;
    LocalFrontier *cur = frontier->getFrontier(0);
    const long long bufferLen = frontier->getEdgeStat();
    next->reserveSparse((size_t)(bufferLen > 0 ? bufferLen : 1));
    check_hip(hipMemsetAsync(next->counters, 0, 2 * sizeof(unsigned long long), lib_stream()),
              "hipMemsetAsync");
    const int64_t count = cur->m;
    if (count > 0) {
        hipLaunchKernelGGL((k_push<W, 1, F>), dim3(push_blocks(count)), dim3(256), 0,
                           lib_stream(), f, GA.view, (const bool *)nullptr, (const intT *)cur->s, count,
                           (bool *)nullptr, next->s, next->counters);
        launched("edgeMapSparseV3");
    }
    unsigned long long c[2];
    toHost(c, next->counters, 2);
    next->m = (intT)c[0];
    next->outEdgesCount = (intT)c[1];
    next->isDense = false;
}

void clearLocalFrontier(LocalFrontier *next, int nodeNum, int subNum, int totalSub);

// edgeMap (polymer.h:2241-2309): dense iff |F| + sum of out-degrees of F >= threshold
template <class F, class G>
void edgeMap(G &GA, vertices *V, F f, LocalFrontier *next, intT threshold = -1, char option = DENSE,
             bool remDups = false, bool part = false,
             Subworker_Partitioner &subworker = polymerDummyPartitioner()) {
    (void)remDups;
    const long long m = (long long)V->numNonzeros() + V->getEdgeStat();
    if (m >= threshold) {
        V->toDense();
        clearLocalFrontier(next, 0, 0, 1);
        if (option == DENSE_FORWARD)
            edgeMapDenseForward(GA, V, f, next, part, subworker.dense_start, subworker.dense_end);
        else
            edgeMapDense(GA, V, f, next, option, subworker);
        next->isDense = true;
    } else {
        V->toSparse();
        edgeMapSparseV3(GA, V, f, next, part, subworker);
        next->isDense = false;
    }
}

/* ---- vertex functions ------------------------------------------------------------------ */
// vertexCounter (polymer.h:2312-2338): m and outEdgesCount of a dense frontier
template <class G>
void vertexCounter(G &GA, LocalFrontier *frontier, int nodeNum, int subNum, int totalSub) {
    using namespace polymer_detail;
    (void)totalSub;
    if (!frontier->isDense || nodeNum != 0 || subNum != 0) return;
    check_hip(hipMemsetAsync(frontier->counters, 0, 2 * sizeof(unsigned long long), lib_stream()),
              "hipMemsetAsync");
    using vertex = typename std::remove_pointer<decltype(GA.V)>::type;
    hipLaunchKernelGGL(k_counter<vertex>, dim3(blocks_for(frontier->n)), dim3(256), 0, lib_stream(),
                       (const bool *)frontier->b, (int64_t)frontier->n, frontier->startID,
                       (const vertex *)GA.V, frontier->counters);
    launched("vertexCounter");
    unsigned long long c[2];
    toHost(c, frontier->counters, 2);
    frontier->m = (intT)c[0];
    frontier->outEdgesCount = (intT)c[1];
}

// vertexMap (polymer.h:2340-2381): f(i) for every active vertex
template <class F>
void vertexMap(vertices *V, F add, int nodeNum, int subNum = 0, int totalSub = 1) {
    using namespace polymer_detail;
    (void)totalSub;
    if (subNum != 0) return;
    LocalFrontier *fr = V->getFrontier(nodeNum);
    if (V->isDense) {
        hipLaunchKernelGGL(k_vertex_map_dense<F>, dim3(blocks_for(fr->n)), dim3(256), 0, lib_stream(),
                           add, (const bool *)fr->b, (int64_t)fr->n, V->getOffset(nodeNum));
    } else if (fr->m > 0) {
        hipLaunchKernelGGL(k_vertex_map_sparse<F>, dim3(blocks_for(fr->m)), dim3(256), 0, lib_stream(),
                           add, (const intT *)fr->s, (int64_t)fr->m);
    }
    launched("vertexMap");
}

// clearLocalFrontier (polymer.h:2383-2397)
inline void clearLocalFrontier(LocalFrontier *next, int nodeNum, int subNum, int totalSub) {
    (void)nodeNum; (void)totalSub;
    if (subNum != 0) return;
    polymer_detail::check_hip(hipMemsetAsync(next->b, 0, (size_t)next->n, polymer_detail::lib_stream()),
                              "hipMemsetAsync");
}

// Every vertex of the frontier's range active (the apps' "all true" start,
// e.g. numa-PageRank.C:427-429); m and outEdgesCount come from vertexCounter.
inline void fillLocalFrontier(LocalFrontier *f) {
    polymer_detail::check_hip(hipMemsetAsync(f->b, 1, (size_t)f->n, polymer_detail::lib_stream()),
                              "hipMemsetAsync");
    f->isDense = true;
    f->m = f->n;
}

// vertexFilter (polymer.h:2411-2441): result bit i = active(i) && filter(i); m counted
template <class F>
void vertexFilter(vertices *V, F filter, int nodeNum, int subNum, int totalSub, LocalFrontier *result) {
    using namespace polymer_detail;
    (void)totalSub;
    if (subNum != 0) return;
    V->toDense();
    LocalFrontier *fr = V->getFrontier(nodeNum);
    check_hip(hipMemsetAsync(result->counters, 0, sizeof(unsigned long long), lib_stream()),
              "hipMemsetAsync");
    hipLaunchKernelGGL(k_vertex_filter<F>, dim3(blocks_for(fr->n)), dim3(256), 0, lib_stream(), filter,
                       (const bool *)fr->b, (int64_t)fr->n, V->getOffset(nodeNum), result->b,
                       result->counters);
    launched("vertexFilter");
    unsigned long long c = 0;
    toHost(&c, result->counters, 1);
    result->m = (intT)c;
    result->isDense = true;
}

// switchFrontier (polymer.h:2224-2236): next becomes the current frontier; the
// old current one is returned in next, cleared
inline void switchFrontier(int nodeNum, vertices *V, LocalFrontier *&next) {
    LocalFrontier *old = V->swapFrontier(nodeNum, next);
    next = old;
    next->clearFrontier();
}

// Waits for the queued work (the reference's calls are synchronous).
inline void polymerSync() {
    polymer_detail::check_hip(hipStreamSynchronize(polymer_detail::lib_stream()), "sync");
}

// Functor classes written between these two lines compile for the host and the
// device without annotations (the app's host code, main() included, stays
// outside: a host+device main() would be compiled for the GPU as well).
#define POLYMER_FUNCTORS_BEGIN _Pragma("clang force_cuda_host_device begin")
#define POLYMER_FUNCTORS_END _Pragma("clang force_cuda_host_device end")

#endif /* POLYMER_HIP_H */
