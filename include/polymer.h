/*
 * polymer.h -- the edgeMap / vertexMap programming interface of Polymer on one
 * MI355X (gfx950).  Header-only C++ compiled by hipcc over the device graphs of
 * libpolymer_gpu.so.  An application keeps the reference's types and calls
 *
 *     graph<vertex>, wghGraph<vertex>, vertex structs     graph.h:11-164
 *     LocalFrontier, vertices (vertexSubset)              polymer.h:521-881
 *     edgeMap, edgeMapDense, edgeMapDenseForward, edgeMapDenseReduce,
 *     edgeMapDenseBP, edgeMapSparseV3, edgeMapSparseAsync(Pipe), ...
 *                                                         polymer.h:924-2309
 *     vertexMap, vertexFilter, vertexCounter,
 *     clearLocalFrontier, switchFrontier                  polymer.h:2224-2441
 *     CAS, writeMin, writeAdd, fetchAndAdd                utils.h:244-310
 *     mapDataArray, partitionByDegree, graphFilter, graphHasher,
 *     Subworker_Partitioner, Custom_barrier, Default_Hash_F,
 *     Default_worker_arg, dummyPartitioner                polymer.h:66-519, 883-919;
 *                                                         custom-barrier.h:20-37
 *     parallel_for, parallel_main, intT, intE             parallel.h
 *
 * so the reference's programs compile with this header in place of its own.
 * Their functors (update / updateAtomic / cond, the weighted (s, d, w) and BP
 * (s, d, j) forms, the initFunc / reduceFunc / combineFunc trio) run as device
 * code: either compile with hipcc --hipstdpar (unannotated functions called from
 * kernels are compiled for the GPU), or write the functor classes between
 * POLYMER_FUNCTORS_BEGIN and POLYMER_FUNCTORS_END.
 *
 * One GPU replaces the NUMA machine:
 *   - numa_num_configured_nodes() and numa_num_configured_cpus() report one node
 *     with one core, so an SPMD program runs one worker thread and every
 *     collective call is made once (its barriers count one thread);
 *     graphFilter / graphFilter2Direction of the one node's range return the
 *     whole graph, which stays in HBM;
 *   - memory from mapDataArray and numa_alloc_local is visible to the host and to
 *     the GPU at one address (managed memory), so the programs' host loops over
 *     value and frontier arrays work unchanged; on a GPU without page migration
 *     the kernels then reach it over the host link.  Compiled with
 *     POLYMER_DEVICE_RESIDENT defined, those arrays and graph.V live in HBM
 *     instead (initialise them with vertexMap or toDevice; host loops over them
 *     must go);
 *   - every call returns when its kernels are done (the reference's calls are
 *     synchronous), on the library stream of libpolymer_gpu.so;
 *   - vertex ids are the ids of the input graph.
 * The edge maps run on the engine's schedules: the dense pull binned by
 * in-degree (a thread, 16 lanes, a wavefront or a workgroup per destination,
 * stopping when cond turns false), the pushes edge-balanced over the active
 * sources' out-edges (a frontier's hub is spread over many workgroups), the
 * sparse push appending its successes with one atomic per wavefront.
 */
#ifndef POLYMER_HIP_H
#define POLYMER_HIP_H

#include <hip/hip_runtime.h>
#include <limits.h>
#include <pthread.h>
#include <sys/mman.h>
#include <sys/time.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#if __has_include(<numa.h>)
#include <numa.h>
#define POLYMER_HAVE_LIBNUMA 1
#endif

#include "polymer_gpu.h"

/* ---- parallel.h (its sequential branch: one host thread per worker) ------- */
#ifndef parallel_for
#define cilk_spawn
#define cilk_sync
#define parallel_main main
#define parallel_for for
#define parallel_for_1 for
#define parallel_for_256 for
#define cilk_for for
#endif
typedef int intT;
typedef unsigned int uintT;
typedef int intE;
typedef unsigned int uintE;
#ifndef INT_T_MAX
#define INT_T_MAX INT_MAX
#define PRIintT "d"
#define PRIuintT "u"
#define PRIintE "d"
#define PRIuintE "u"
#endif
#ifndef PAGESIZE
#define PAGESIZE (4096)
#endif
#ifndef MIN
#define MIN(x, y) ((x > y) ? (y) : (x))
#endif

#define PG_HD __host__ __device__

/* ---- gettime.h: the apps' phase timer ("<phase> : <seconds>") -------------- */
#ifndef _BENCH_GETTIME_INCLUDED
#define _BENCH_GETTIME_INCLUDED
struct timer {
    double totalTime = 0.0, lastTime = 0.0, totalWeight = 0.0;
    bool on = false;
    double getTime() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
            .count();
    }
    void start() {
        on = true;
        lastTime = getTime();
    }
    double stop() {
        on = false;
        const double d = getTime() - lastTime;
        totalTime += d;
        return d;
    }
    double stop(double weight) {
        on = false;
        totalWeight += weight;
        const double d = getTime() - lastTime;
        totalTime += weight * d;
        return d;
    }
    double total() { return on ? totalTime + getTime() - lastTime : totalTime; }
    double next() {
        if (!on) return 0.0;
        const double t = getTime(), d = t - lastTime;
        totalTime += d;
        lastTime = t;
        return d;
    }
    void reportT(double time) { std::cout << std::setprecision(3) << time << std::endl; }
    void reportTime(double time) { reportT(time); }
    void reportStop(double weight, std::string str) {
        std::cout << str << " :" << weight << ": ";
        reportTime(stop(weight));
    }
    void reportTotal() {
        reportTime(totalWeight > 0.0 ? total() / totalWeight : total());
        totalTime = 0.0;
        totalWeight = 0.0;
    }
    void reportTotal(std::string str) {
        std::cout << str << " : ";
        reportTotal();
    }
    void reportNext() { reportTime(next()); }
    void reportNext(std::string str) {
        std::cout << str << " : ";
        reportNext();
    }
};
static timer _tm;
#define timeStatement(_A, _string) \
    _tm.start();                   \
    _A;                            \
    _tm.reportNext(_string);
#define startTime() _tm.start();
#define stopTime(_weight, _str) _tm.reportStop(_weight, _str);
#define reportTime(_str) _tm.reportTotal(_str);
#define nextTime(_string) _tm.reportNext(_string);
#define nextTimeN() _tm.reportT(_tm.next());
#endif

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

[[noreturn]] inline void fail(const char *msg) {
    std::fprintf(stderr, "polymer: %s\n", msg);
    std::abort();
}

// the library stream: every call of this header is ordered on it
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

inline void sync() { check_hip(hipStreamSynchronize(lib_stream()), "hipStreamSynchronize"); }

// The end of a call whose results the program may read directly: with managed
// arrays the host loops read them, so the call waits for its kernels; with
// device-resident arrays the program reaches them only through toHost (which
// waits) or further calls on the same stream, so the call returns at once.
inline void sync_visible() {
#ifndef POLYMER_DEVICE_RESIDENT
    sync();
#endif
}

inline void launched(const char *what) { check_hip(hipGetLastError(), what); }

// Bumped by every call that may change a frontier's flags or a program's device
// data (copies in, vertex maps and filters, the edge maps, clearing and filling):
// counts a dense edgeMap took of its output stay valid for vertexCounter only while
// nothing else has run.
inline unsigned long long &frontier_gen() {
    static unsigned long long g = 1;
    return g;
}
inline void frontiers_changed() { ++frontier_gen(); }

inline unsigned blocks_for(int64_t n, int threads = 256, int64_t cap = 65536) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

template <class T>
T *dev_alloc(size_t count) {
    void *p = nullptr;
    check_hip(hipMalloc(&p, (count ? count : 1) * sizeof(T)), "hipMalloc");
    return static_cast<T *>(p);
}

// Named device buffers kept for the life of the process, grown on demand (an
// allocation only when a call needs more than any call before it).
inline void *workspace(const char *name, size_t bytes) {
    static std::map<std::string, std::pair<void *, size_t>> ws;
    auto &e = ws[name];
    if (e.second < bytes) {
        if (e.first) check_hip(hipFree(e.first), "hipFree");
        e.first = nullptr;
        check_hip(hipMalloc(&e.first, bytes + 64), "hipMalloc");
        e.second = bytes;
    }
    return e.first;
}
template <class T>
T *workspace_t(const char *name, size_t count) {
    return static_cast<T *>(workspace(name, (count ? count : 1) * sizeof(T)));
}

// Memory the application's host loops and the kernels both address: managed
// memory (one address on the host and the GPU), zero-filled like the reference's
// anonymous mappings.  With POLYMER_DEVICE_RESIDENT: device memory.
inline void *shared_alloc(size_t bytes, bool zero = true) {
    ensure_init();
    void *p = nullptr;
    const size_t b = bytes ? bytes : 1;
#ifdef POLYMER_DEVICE_RESIDENT
    check_hip(hipMalloc(&p, b), "hipMalloc");
    if (zero) {
        check_hip(hipMemsetAsync(p, 0, b, lib_stream()), "hipMemsetAsync");
        sync();
    }
#else
    check_hip(hipMallocManaged(&p, b, hipMemAttachGlobal), "hipMallocManaged");
    if (zero) std::memset(p, 0, b);
#endif
    return p;
}

inline void shared_free(void *p) {
    if (p) (void)hipFree(p);
}

// A device-readable view of count elements at p: p itself when the GPU can address
// it (device, managed or pinned memory), else a staged copy (a list the program
// built in malloc'd memory, e.g. numa-BFS.C:380).
template <class T>
const T *device_readable(const T *p, size_t count, const char *stage) {
    if (!p || count == 0) return p;
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e == hipSuccess && a.type != hipMemoryTypeUnregistered &&
        (a.type != hipMemoryTypeHost || a.devicePointer != nullptr))
        return p;
    (void)hipGetLastError();  // clear the lookup's error state
    T *d = workspace_t<T>(stage, count);
    check_hip(hipMemcpyAsync(d, p, count * sizeof(T), hipMemcpyHostToDevice, lib_stream()),
              "hipMemcpyAsync");
    sync();  // the program may reuse its list once the call returns
    return d;
}

}  // namespace polymer_detail

/* ---- libnuma: one node with one core --------------------------------------- */
namespace polymer_detail {
inline int numa_nodes() { return 1; }
inline int numa_cpus() { return 1; }
}  // namespace polymer_detail
#define numa_num_configured_nodes() (polymer_detail::numa_nodes())
#define numa_num_configured_cpus() (polymer_detail::numa_cpus())
#define numa_alloc_local(sz) (polymer_detail::shared_alloc(sz))
#define numa_alloc_onnode(sz, node) (polymer_detail::shared_alloc(sz))
#define numa_alloc_interleaved(sz) (polymer_detail::shared_alloc(sz))
#define numa_free(p, sz) (polymer_detail::shared_free(p))
#ifndef POLYMER_HAVE_LIBNUMA
struct bitmask;
inline bitmask *numa_parse_nodestring(const char *) { return nullptr; }
inline void numa_bind(bitmask *) {}
inline void numa_set_interleave_mask(bitmask *) {}
static bitmask *numa_all_nodes_ptr = nullptr;
#endif

/* ---- copies between host and the device arrays of an app ----------------- */
template <class T>
void toDevice(T *dev, const T *host, size_t count) {
    polymer_detail::frontiers_changed();
    polymer_detail::check_hip(hipMemcpyAsync(dev, host, count * sizeof(T), hipMemcpyHostToDevice,
                                             polymer_detail::lib_stream()),
                              "hipMemcpyAsync");
    polymer_detail::sync();
}
template <class T>
void toHost(T *host, const T *dev, size_t count) {
    polymer_detail::check_hip(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost,
                                             polymer_detail::lib_stream()),
                              "hipMemcpyAsync");
    polymer_detail::sync();
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

// On device memory (POLYMER_DEVICE_RESIDENT) floating-point additions use the
// hardware's atomic add; memory the host also maps takes the compare-and-swap form.
template <class ET>
PG_HD inline void writeAdd(ET *a, ET b) {
#if __HIP_DEVICE_COMPILE__
#ifdef POLYMER_DEVICE_RESIDENT
    if constexpr (std::is_same<ET, double>::value || std::is_same<ET, float>::value)
        unsafeAtomicAdd(a, b);
    else
        atomicAdd(a, b);
#else
    atomicAdd(a, b);
#endif
#else
    ET c = *a;
    while (!CAS(a, c, (ET)(c + b))) c = *(volatile ET *)a;
#endif
}

template <class ET>
PG_HD inline ET fetchAndAdd(ET *a, ET b) {
#if __HIP_DEVICE_COMPILE__
    return atomicAdd(a, b);
#else
    return __atomic_fetch_add(a, b, __ATOMIC_SEQ_CST);
#endif
}

/* ---- custom-barrier.h: a counting barrier over two shared words ------------ */
#ifndef CUSTOM_BARRIER
#define CUSTOM_BARRIER
// Custom_barrier (custom-barrier.h:20-37): count threads meet; the last one to
// arrive resets the counter and flips the toggle word the others wait on.
struct Custom_barrier {
    volatile int *counter = nullptr;
    volatile int *toggleVar = nullptr;
    int count = 0;
    Custom_barrier() {}
    Custom_barrier(volatile int *_counter, volatile int *_toggle, int _count)
        : counter(_counter), toggleVar(_toggle), count(_count) {}
    inline void wait() {
        if (!counter || !toggleVar || count <= 1) return;  // one worker: nothing to meet
        const int before = __atomic_load_n(toggleVar, __ATOMIC_ACQUIRE);
        if (__atomic_add_fetch(counter, 1, __ATOMIC_ACQ_REL) >= count) {
            __atomic_store_n(counter, 0, __ATOMIC_RELAXED);
            __atomic_store_n(toggleVar, before ^ 1, __ATOMIC_RELEASE);
        } else {
            while (__atomic_load_n(toggleVar, __ATOMIC_ACQUIRE) == before) __builtin_ia32_pause();
        }
    }
};
#endif

/* ---- SPMD roles and worker arguments (polymer.h:66-129, 883-919) ----------- */
inline long long roundUp(double x) {
    long long ones = (long long)x;
    return (x - (double)ones > 0) ? ones + 1 : ones;
}

struct Subworker_Partitioner {
    int tid = 0;
    int subTid = 0;
    int numOfSub = 1;
    intT dense_start = 0;
    intT dense_end = 0;
    pthread_barrier_t *global_barr = nullptr;
    pthread_barrier_t *local_barr = nullptr;
    pthread_barrier_t *leader_barr = nullptr;
    Custom_barrier local_custom;
    Custom_barrier subMaster_custom;
    explicit Subworker_Partitioner(int nSub) : numOfSub(nSub) {}
    inline bool isMaster() const { return tid + subTid == 0; }
    inline bool isSubMaster() const { return subTid == 0; }
    inline intT getStartPos(intT m) const { return subTid * (m / numOfSub); }
    inline intT getEndPos(intT m) const {
        return subTid == numOfSub - 1 ? m : (subTid + 1) * (m / numOfSub);
    }
    inline void localWait() { local_custom.wait(); }
    inline void globalWait() {
        local_custom.wait();
        if (isSubMaster()) subMaster_custom.wait();
        local_custom.wait();
    }
};

// the shard-interleaving relabel of the apps (identity for one shard)
struct Default_Hash_F {
    int shardNum;
    intT vertPerShard;
    intT n;
    Default_Hash_F(intT _n, int _shardNum) : shardNum(_shardNum), vertPerShard(_n / _shardNum), n(_n) {}
    inline intT hashFunc(intT index) {
        if (index >= shardNum * vertPerShard) return index;
        return (index % shardNum) * vertPerShard + index / shardNum;
    }
    inline intT hashBackFunc(intT index) {
        if (index >= shardNum * vertPerShard) return index;
        return index / vertPerShard + (index % vertPerShard) * shardNum;
    }
};

struct Default_worker_arg {
    void *GA;
    int maxIter;
    int tid;
    int numOfNode;
    intT rangeLow;
    intT rangeHi;
};

struct LocalFrontier;
struct Default_subworker_arg {
    void *GA;
    void *Global_G;
    int maxIter;
    int tid;
    int subTid;
    intT startPos;
    intT endPos;
    intT rangeLow;
    intT rangeHi;
    pthread_barrier_t *global_barr;
    pthread_barrier_t *node_barr;
    pthread_barrier_t *master_barr;
    LocalFrontier *localFrontier;
    volatile int *local_custom_counter;
    volatile int *local_custom_toggle;
};

struct nonNegF {
    bool operator()(intT a) { return a >= 0; }
};

// edgeMap options (polymer.h:917-919)
enum options { DENSE, DENSE_PARALLEL, DENSE_FORWARD };

inline Subworker_Partitioner dummyPartitioner(1);

inline Subworker_Partitioner &polymerDummyPartitioner() { return dummyPartitioner; }

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

// Destinations of a graph binned by in-degree for the dense pull (built on first
// use, dropped with the graph): bin g is ids[start[g], start[g] + count[g]), g = 0
// a thread per destination (in-degree 1-16), 1 sixteen lanes (17-256), 2 a
// wavefront (257-4096), 3 a workgroup (more); heavier bins come first in ids.
struct Bins {
    intT *ids = nullptr;
    int64_t start[4] = {0};
    int64_t count[4] = {0};
    int64_t nonzero = 0;  // destinations with in-edges
    intT *first_in = nullptr;  // first in-neighbour of every vertex (-1: none)
};
inline std::map<const void *, Bins> &bins_cache() {
    static std::map<const void *, Bins> cache;
    return cache;
}

// Owner of one library graph and the vertex array over it; graph<vertex> copies
// share it (a program's filtered "local" graph of the one node is the graph itself)
struct GraphOwner {
    pg_graph *handle = nullptr;
    void *V = nullptr;
    const void *key = nullptr;  // the graph's in-offsets, key of its bins
    ~GraphOwner() {
        auto it = bins_cache().find(key);
        if (it != bins_cache().end()) {
            if (it->second.ids) (void)hipFree(it->second.ids);
            if (it->second.first_in) (void)hipFree(it->second.first_in);
            bins_cache().erase(it);
        }
        if (V) (void)hipFree(V);
        if (handle) pg_graph_destroy(handle);
    }
};

}  // namespace polymer_detail

/* ---- graphs ------------------------------------------------------------------ */
// graph<vertex> (graph.h:61-87).  V holds n vertices whose neighbour pointers
// address the device adjacency of the library graph; V itself is host-visible
// (device-resident under POLYMER_DEVICE_RESIDENT).
template <class vertex>
struct graph {
    vertex *V = nullptr;
    long n = 0;
    long m = 0;
    pg_graph *handle = nullptr;
    pg_view view{};
    std::shared_ptr<polymer_detail::GraphOwner> own;
    // this copy lets go of the graph; the last copy to do so frees it
    void del() {
        own.reset();
        V = nullptr;
        handle = nullptr;
    }
};

template <class vertex>
struct wghGraph : graph<vertex> {};

namespace polymer_detail {

template <class G>
struct is_weighted : std::false_type {};
template <class V>
struct is_weighted<wghGraph<V>> : std::true_type {};

inline const Bins &bins_of(const pg_view &v);
inline void presize_workspaces(int64_t n);

template <class G>
void wrap_into(G &g, pg_graph *h, bool weighted) {
    g.handle = h;
    check(pg_graph_view(h, &g.view), "pg_graph_view");
    lib_stream() = static_cast<hipStream_t>(g.view.stream);
    if (weighted && g.view.w_dtype != PG_W_I32) fail("weighted graph without int32 weights");
    g.n = (long)g.view.n;
    g.m = (long)g.view.m;
    using vertex = typename std::remove_pointer<decltype(g.V)>::type;
    g.own = std::make_shared<GraphOwner>();
    g.own->handle = h;
    g.own->key = g.view.in_off;
    g.V = static_cast<vertex *>(shared_alloc(sizeof(vertex) * (size_t)g.n));
    g.own->V = g.V;
    hipLaunchKernelGGL(k_fill_vertices<vertex>, dim3(blocks_for(g.n)), dim3(256), 0, lib_stream(),
                       g.view, g.V);
    launched("k_fill_vertices");
    sync();
    (void)bins_of(g.view);  // the dense pull's in-degree bins, built with the graph
    presize_workspaces((int64_t)g.n);
}

template <class G>
G wrap(pg_graph *h, bool weighted) {
    G g;
    wrap_into(g, h, weighted);
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

// readGraphSkipRing (IO-numa.h:192-315): the same without self-loops
template <class vertex>
graph<vertex> readGraphSkipRing(const char *fname, bool isSymmetric, bool binary = false) {
    polymer_detail::ensure_init();
    pg_graph *h = nullptr;
    const uint32_t flags = (isSymmetric ? PG_GRAPH_SYMMETRIC : 0u) | PG_GRAPH_SKIP_SELF_LOOPS;
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

/* ---- NUMA partitioning of the graph (polymer.h:131-497) --------------------- */
namespace polymer_detail {

// degrees of the graph in input ids (in- or out-), on the host
template <class G>
std::vector<int32_t> host_degrees(const G &GA, bool out) {
    std::vector<int32_t> off((size_t)GA.n + 1);
    toHost(off.data(), out ? GA.view.out_off : GA.view.in_off, off.size());
    std::vector<int32_t> deg((size_t)GA.n);
    for (long i = 0; i < GA.n; i++) deg[i] = off[i + 1] - off[i];
    return deg;
}

// cut [lo, hi) into k consecutive ranges of about equal degree sum, each a whole
// number of pages of eleSize-byte elements except the last
inline void split_by_degree(const std::vector<int32_t> &deg, intT lo, intT hi, int k,
                            intT *sizeArr, int eleSize) {
    if (k <= 1) {
        sizeArr[0] = hi - lo;
        return;
    }
    long long total = 0;
    for (intT i = lo; i < hi; i++) total += deg[i];
    const intT page = std::max<intT>(1, PAGESIZE / std::max(1, eleSize));
    intT start = lo;
    long long acc = 0;
    int shard = 0;
    for (intT i = lo; i < hi && shard < k - 1; i++) {
        acc += deg[i];
        if (acc * k >= total * (shard + 1)) {
            intT end = std::min<intT>(hi, lo + ((i + 1 - lo + page - 1) / page) * page);
            if (end < start) end = start;
            sizeArr[shard++] = end - start;
            start = end;
            i = end - 1;
        }
    }
    while (shard < k - 1) sizeArr[shard++] = 0;
    sizeArr[k - 1] = hi - start;
}

}  // namespace polymer_detail

// partitionByDegree (polymer.h:131-192): node ranges of equal in- (or out-)degree
template <class G>
void partitionByDegree(const G &GA, int numOfShards, intT *sizeArr, intT sizeOfOneEle,
                       bool useOutDegree = false) {
    if (numOfShards <= 1) {
        sizeArr[0] = (intT)GA.n;
        return;
    }
    polymer_detail::split_by_degree(polymer_detail::host_degrees(GA, useOutDegree), 0, (intT)GA.n,
                                    numOfShards, sizeArr, sizeOfOneEle);
}

// subPartitionByDegree (polymer.h:194-282): the sub-worker ranges of one node
template <class G>
void subPartitionByDegree(const G &GA, int numOfShards, intT *sizeArr, int sizeOfOneEle,
                          bool useOutDegree = false, bool useFakeDegree = false) {
    (void)useFakeDegree;
    if (numOfShards <= 1) {
        sizeArr[0] = (intT)GA.n;
        return;
    }
    polymer_detail::split_by_degree(polymer_detail::host_degrees(GA, useOutDegree), 0, (intT)GA.n,
                                    numOfShards, sizeArr, sizeOfOneEle);
}

template <class G>
void subPartitionByDegree(const G &GA, int numOfShards, intT *sizeArr, int sizeOfOneEle,
                          intT subStart, intT subEnd, bool useOutDegree = false,
                          bool useFakeDegree = false) {
    (void)useFakeDegree;
    if (numOfShards <= 1) {
        sizeArr[0] = subEnd - subStart;
        return;
    }
    polymer_detail::split_by_degree(polymer_detail::host_degrees(GA, useOutDegree), subStart,
                                    subEnd, numOfShards, sizeArr, sizeOfOneEle);
}

namespace polymer_detail {

// relabel the graph by hash.hashFunc (a permutation of [0, n)); nothing to do
// for the identity, which is what one shard gives
template <class G, class Hash_F>
void relabel(G &GA, Hash_F &hash) {
    const long n = GA.n;
    std::vector<int32_t> h((size_t)n);
    bool identity = true;
    for (long i = 0; i < n; i++) {
        h[i] = (int32_t)hash.hashFunc((intT)i);
        identity = identity && h[i] == i;
    }
    if (identity) return;
    std::vector<int64_t> off((size_t)n + 1);
    std::vector<int32_t> adj((size_t)GA.m);
    std::vector<int32_t> w;
    const bool weighted = GA.view.w_dtype == PG_W_I32;
    if (weighted) w.resize((size_t)GA.m);
    check(pg_graph_download(GA.handle, off.data(), adj.data(), weighted ? w.data() : nullptr),
          "pg_graph_download");
    std::vector<int64_t> noff((size_t)n + 1, 0);
    for (long u = 0; u < n; u++) noff[(size_t)h[u] + 1] = off[u + 1] - off[u];
    for (long v = 0; v < n; v++) noff[v + 1] += noff[v];
    std::vector<int32_t> nadj((size_t)GA.m), nw(weighted ? (size_t)GA.m : 0);
    for (long u = 0; u < n; u++) {
        int64_t p = noff[(size_t)h[u]];
        for (int64_t e = off[u]; e < off[u + 1]; e++, p++) {
            nadj[(size_t)p] = h[(size_t)adj[(size_t)e]];
            if (weighted) nw[(size_t)p] = w[(size_t)e];
        }
    }
    pg_graph *nh = nullptr;
    check(pg_graph_create(n, GA.m, noff.data(), nadj.data(), weighted ? nw.data() : nullptr,
                          weighted ? PG_W_I32 : PG_W_NONE, GA.view.symmetric ? PG_GRAPH_SYMMETRIC : 0u,
                          &nh),
          "pg_graph_create");
    GA.del();
    wrap_into(GA, nh, weighted);
}

}  // namespace polymer_detail

// graphHasher / graphInEdgeHasher / graphAllEdgeHasher (polymer.h:284-344): vertex
// ids relabelled by hash.hashFunc, every edge list with them
template <class G, class Hash_F>
void graphHasher(G &GA, Hash_F hash) {
    polymer_detail::relabel(GA, hash);
}
template <class G, class Hash_F>
void graphInEdgeHasher(G &GA, Hash_F hash) {
    polymer_detail::relabel(GA, hash);
}
template <class G, class Hash_F>
void graphAllEdgeHasher(G &GA, Hash_F hash) {
    polymer_detail::relabel(GA, hash);
}

// graphFilter / graphFilter2Direction (polymer.h:346-497): the edges of one node's
// range.  One GPU is one node: its range is the whole graph, which is returned.
template <class G>
G graphFilter(G &GA, intT rangeLow, intT rangeHi, bool useOutEdge = true) {
    (void)useOutEdge;
    if (rangeLow != 0 || rangeHi != (intT)GA.n)
        polymer_detail::fail("graphFilter: a range other than the whole graph (the GPU is one node)");
    return GA;
}
template <class G>
G graphFilter2Direction(G &GA, intT rangeLow, intT rangeHi) {
    return graphFilter(GA, rangeLow, rangeHi);
}

/* ---- value arrays (polymer.h:499-519) ------------------------------------------ */
// sum(sizeArr[0..numOfShards)) elements of sizeOfOneEle bytes, zeroed, visible to
// the host and the GPU (device memory under POLYMER_DEVICE_RESIDENT)
inline void *mapDataArray(int numOfShards, intT *sizeArr, int sizeOfOneEle) {
    size_t total = 0;
    for (int i = 0; i < numOfShards; i++) total += (size_t)sizeArr[i];
    return polymer_detail::shared_alloc(total * (size_t)sizeOfOneEle);
}

inline void freeDataArray(void *p) { polymer_detail::shared_free(p); }

// zeroed device memory for n elements of T (HBM whatever the mode)
template <class T>
T *mapDeviceArray(size_t n) {
    polymer_detail::ensure_init();
    T *p = polymer_detail::dev_alloc<T>(n);
    polymer_detail::check_hip(hipMemsetAsync(p, 0, (n ? n : 1) * sizeof(T), polymer_detail::lib_stream()),
                              "hipMemsetAsync");
    polymer_detail::sync();
    return p;
}

// A bool array of n entries, all false: the frontier storage the reference takes
// from numa_alloc_local.
inline bool *newFrontierArray(intT n) {
    return static_cast<bool *>(polymer_detail::shared_alloc((size_t)n));
}

/* ---- frontier kernels ---------------------------------------------------------- */
namespace polymer_detail {

constexpr int kT = 256;          // threads of most kernels
constexpr int kChunk = 16 * kT;  // bytes of a frontier per compaction block

__device__ __forceinline__ int lane() { return (int)__lane_id(); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int l = lane();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (l >= o) v += y;
    }
    return v;
}

// exclusive scan over a 256-thread block; *tot = the sum; ends with a barrier
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sm, uint32_t *tot) {
    const uint32_t inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6, l = lane();
    if (l == 63) sm[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < kT / 64; i++) {
            const uint32_t x = sm[i];
            sm[i] = run;
            run += x;
        }
        sm[kT / 64] = run;
    }
    __syncthreads();
    const uint32_t r = sm[w] + inc - v;
    *tot = sm[kT / 64];
    __syncthreads();
    return r;
}

// ---- totals a launch reports to the host ---------------------------------------------
// The blocks of a counting launch add partial sums into kTotSlots spread slot pairs
// of a device buffer that holds zeros between launches; one single-block kernel
// after it adds the slots up into pinned host memory and clears them.  A call that
// reports counts thus needs no clearing copy before its kernel and no copy back
// after it (a copy into pageable memory measured 22 us against 12 us for a launch
// and a synchronisation), and no block waits for the others (a last-block ticket
// serialised 2048 blocks on one word).
constexpr int kTotSlots = 32;  // slot pairs
struct Totals {
    unsigned long long *slots;  // device, 2 * kTotSlots words
    unsigned long long *host;   // the kernels' address of the pinned words
    unsigned long long *cpu;    // the host's address: the two totals and a sequence number
    unsigned long long seq;     // the sequence number of the last finishing launch
};
inline Totals &totals() {
    static Totals t{nullptr, nullptr, nullptr, 0};
    if (!t.slots) {
        ensure_init();
        const size_t bytes = 2 * kTotSlots * sizeof(unsigned long long);
        check_hip(hipMalloc(&t.slots, bytes), "hipMalloc");
        check_hip(hipMemsetAsync(t.slots, 0, bytes, lib_stream()), "hipMemsetAsync");
        check_hip(hipHostMalloc(&t.cpu, 4 * sizeof(unsigned long long), hipHostMallocDefault),
                  "hipHostMalloc");
        check_hip(hipHostGetDevicePointer((void **)&t.host, t.cpu, 0), "hipHostGetDevicePointer");
        t.cpu[0] = t.cpu[1] = t.cpu[2] = 0;
        sync();
    }
    return t;
}
// The two totals into the pinned words, then (after a system-scope fence) the
// sequence number that tells the host they are there.
__device__ __forceinline__ void publish_totals(unsigned long long *host, unsigned long long a,
                                               unsigned long long b, unsigned long long seq) {
    host[0] = a;
    host[1] = b;
    __threadfence_system();
    host[2] = seq;
    __threadfence_system();
}
// the slot pairs' sums published; the slots cleared for the next launch
__global__ void k_totals(unsigned long long *slots, unsigned long long *host,
                         unsigned long long seq) {
    __shared__ unsigned long long s[2 * kTotSlots];
    const int i = threadIdx.x;  // 2 * kTotSlots threads
    s[i] = slots[i];
    slots[i] = 0ull;
    __syncthreads();
    if (i == 0) {
        unsigned long long a = 0, b = 0;
        for (int k = 0; k < kTotSlots; k++) {
            a += s[2 * k];
            b += s[2 * k + 1];
        }
        publish_totals(host, a, b, seq);
    }
}
__global__ void k_publish_count(const uint32_t *count, unsigned long long *host,
                                unsigned long long seq) {
    publish_totals(host, *count, 0ull, seq);
}
inline void finish_totals() {
    Totals &t = totals();
    hipLaunchKernelGGL(k_totals, dim3(1), dim3(2 * kTotSlots), 0, lib_stream(), t.slots, t.host,
                       ++t.seq);
    launched("totals");
}
// The totals of the last publishing launch (finish_totals or the compaction's
// scan).  Everything queued before that launch has finished once its sequence
// number is visible, so the host spins on that word (a wake-up from a stream
// synchronisation measured 12 us for a launch against 6 us for the spin) and
// falls back to the synchronisation, which reports a failed kernel, after 2 ms.
inline void read_totals(unsigned long long *a, unsigned long long *b) {
    Totals &t = totals();
    volatile unsigned long long *h = t.cpu;
    const auto t0 = std::chrono::steady_clock::now();
    while (h[2] != t.seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
            sync();
            break;
        }
    }
    *a = h[0];
    if (b) *b = h[1];
}

// The set flags among b[i, i + 16) as the bits of a mask (bit j = flag i + j; i a
// multiple of 16): one 16-byte load when b is 16-byte aligned and the block lies
// inside [0, n), byte loads otherwise (a program's own offset array, the tail).
__device__ __forceinline__ uint32_t flag_mask16(const bool *b, int64_t i, int64_t n) {
    uint32_t m = 0;
    if ((((uintptr_t)b) & 15u) == 0u && i + 16 <= n) {
        const uint4 x = *reinterpret_cast<const uint4 *>(b + i);
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((w[k] >> (8 * j)) & 0xFFu) m |= 1u << (4 * k + j);
    } else {
        for (int j = 0; j < 16; j++)
            if (i + j < n && b[i + j]) m |= 1u << j;
    }
    return m;
}

// count of set flags per kChunk bytes (16 per thread)
__global__ __launch_bounds__(kT) void k_flag_counts(const bool *b, int64_t n, uint32_t *cnt) {
    __shared__ uint32_t sm[kT / 64 + 1];
    const int64_t i = (int64_t)blockIdx.x * kChunk + 16 * (int64_t)threadIdx.x;
    const uint32_t c = i < n ? (uint32_t)__popc(flag_mask16(b, i, n)) : 0u;
    uint32_t tot = 0;
    (void)block_excl_scan(c, sm, &tot);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// out[i] = in[0] + ... + in[i-1] for i < n, out[n] = total; one workgroup (n small)
// exclusive scan of in[0, n) into out, out[n] = the total, by one block: each pass
// takes 16 * kT entries, 16 consecutive ones per thread, all loaded before the
// block's scan (a pass of kT entries per block scan was a chain of dependent
// load-and-scan steps: 30 us for 4096 entries)
__global__ __launch_bounds__(kT) void k_scan_small(const uint32_t *in, uint32_t *out, int64_t n) {
    __shared__ uint32_t sm[kT / 64 + 1];
    uint32_t carry = 0;
    for (int64_t base = 0; base < n; base += 16 * kT) {  // block-uniform
        const int64_t i0 = base + 16 * (int64_t)threadIdx.x;
        uint32_t v[16], run = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = i0 + k < n ? in[i0 + k] : 0u;
#pragma unroll
        for (int k = 0; k < 16; k++) run += v[k];
        uint32_t tot = 0;
        uint32_t x = carry + block_excl_scan(run, sm, &tot);
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (i0 + k < n) out[i0 + k] = x;
            x += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// ordered compaction: the ids (+ offset) of the set flags, block by block; a
// thread writes the ids of its 16 flags in order after its prefix
__global__ __launch_bounds__(kT) void k_flag_write(const bool *b, int64_t n, const uint32_t *starts,
                                                   intT offset, intT *s) {
    __shared__ uint32_t sm[kT / 64 + 1];
    const int64_t i = (int64_t)blockIdx.x * kChunk + 16 * (int64_t)threadIdx.x;
    uint32_t mk = i < n ? flag_mask16(b, i, n) : 0u;
    uint32_t tot = 0;
    uint32_t out = starts[blockIdx.x] + block_excl_scan((uint32_t)__popc(mk), sm, &tot);
    while (mk) {
        const int j = __ffs(mk) - 1;
        mk &= mk - 1u;
        s[out++] = (intT)(i + j) + offset;
    }
}

// the frontier's flags into an ordered list; returns the count (host)
inline intT compact_flags(const bool *b, int64_t n, intT offset, intT *s) {
    const int64_t nb = (n + kChunk - 1) / kChunk;
    if (nb == 0) return 0;
    uint32_t *cnt = workspace_t<uint32_t>("compact_cnt", (size_t)nb + 1);
    uint32_t *st = workspace_t<uint32_t>("compact_start", (size_t)nb + 1);
    hipLaunchKernelGGL(k_flag_counts, dim3((unsigned)nb), dim3(kT), 0, lib_stream(), b, n, cnt);
    Totals &t = totals();
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kT), 0, lib_stream(), (const uint32_t *)cnt, st, nb);
    hipLaunchKernelGGL(k_flag_write, dim3((unsigned)nb), dim3(kT), 0, lib_stream(), b, n,
                       (const uint32_t *)st, offset, s);
    // the count, published once the list is written
    hipLaunchKernelGGL(k_publish_count, dim3(1), dim3(1), 0, lib_stream(), (const uint32_t *)(st + nb),
                       t.host, ++t.seq);
    launched("compaction");
    unsigned long long total = 0;
    read_totals(&total, nullptr);
    return (intT)total;
}

// scan over the out-degrees of list[0, count): deg scan in tiles of 4096
constexpr int kScanTile = 16 * kT;

// the out-degrees of list[i0, i0 + 16) (0 past count), every load issued before
// the first is used
__device__ __forceinline__ void degrees16(const intT *list, int64_t count, const int32_t *off,
                                          int64_t i0, uint32_t (&d)[16]) {
    intT v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = i0 + k < count ? list[i0 + k] : -1;
    int32_t a[16], b[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        a[k] = v[k] >= 0 ? off[v[k]] : 0;
        b[k] = v[k] >= 0 ? off[v[k] + 1] : 0;
    }
#pragma unroll
    for (int k = 0; k < 16; k++) d[k] = (uint32_t)(b[k] - a[k]);
}

// tiles of kScanTile entries, 16 consecutive ones per thread
__global__ __launch_bounds__(kT) void k_deg_tile_sums(const intT *list, int64_t count,
                                                      const int32_t *off, uint32_t *sums) {
    __shared__ uint32_t sm[kT / 64 + 1];
    uint32_t d[16], c = 0;
    degrees16(list, count, off, (int64_t)blockIdx.x * kScanTile + 16 * (int64_t)threadIdx.x, d);
#pragma unroll
    for (int k = 0; k < 16; k++) c += d[k];
    uint32_t tot = 0;
    (void)block_excl_scan(c, sm, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kT) void k_deg_tiles(const intT *list, int64_t count, const int32_t *off,
                                                  const uint32_t *tile_start, uint32_t *scan) {
    __shared__ uint32_t sm[kT / 64 + 1];
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + 16 * (int64_t)threadIdx.x;
    uint32_t d[16], run = 0;
    degrees16(list, count, off, i0, d);
#pragma unroll
    for (int k = 0; k < 16; k++) run += d[k];
    uint32_t tot = 0;
    const uint32_t start = tile_start[blockIdx.x];
    uint32_t x = start + block_excl_scan(run, sm, &tot);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (i0 + k < count) scan[i0 + k] = x;
        x += d[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) scan[count] = start + tot;
}

// scan[i] = sum of outdeg(list[0..i)), scan[count] = total, on the device
inline void degree_scan(const intT *list, int64_t count, const int32_t *off, uint32_t *scan) {
    const int64_t nt = (count + kScanTile - 1) / kScanTile;
    uint32_t *sums = workspace_t<uint32_t>("deg_sums", (size_t)nt + 1);
    uint32_t *starts = workspace_t<uint32_t>("deg_starts", (size_t)nt + 1);
    hipLaunchKernelGGL(k_deg_tile_sums, dim3((unsigned)nt), dim3(kT), 0, lib_stream(), list, count,
                       off, sums);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kT), 0, lib_stream(), (const uint32_t *)sums,
                       starts, nt);
    hipLaunchKernelGGL(k_deg_tiles, dim3((unsigned)nt), dim3(kT), 0, lib_stream(), list, count, off,
                       (const uint32_t *)starts, scan);
    launched("degree scan");
}

// The per-round workspaces of the edge maps sized for a graph of n vertices when
// the graph is made (a list longer than n, possible with repeated successes, still
// grows them on use), so a program's loop of rounds does not allocate.
inline void presize_workspaces(int64_t n) {
    const size_t nw = (size_t)(n + 31) / 32 + 1;
    (void)workspace_t<uint32_t>("frontier_bits", nw);
    (void)workspace_t<uint32_t>("pull_marks", nw);
    (void)workspace_t<intT>("active_list", (size_t)n + 1);
    (void)workspace_t<uint32_t>("push_scan", (size_t)n + 1);
    const size_t nb = (size_t)(n + kChunk - 1) / kChunk + 1;
    (void)workspace_t<uint32_t>("compact_cnt", nb);
    (void)workspace_t<uint32_t>("compact_start", nb);
    const size_t nt = (size_t)(n + kScanTile - 1) / kScanTile + 1;
    (void)workspace_t<uint32_t>("deg_sums", nt);
    (void)workspace_t<uint32_t>("deg_starts", nt);
}

template <bool W, class F>
__device__ __forceinline__ bool call_update_atomic(F &f, intT s, intT d, const intE *w, int64_t e) {
    if constexpr (W)
        return f.updateAtomic(s, d, w[e]);
    else
        return f.updateAtomic(s, d);
}

// Edge-balanced push over list[0, count) with degree scan `scan` (scan[count] =
// the total, read on the device, so the launch waits for no count): a fixed grid,
// block b taking the frontier's edges [b*epb, (b+1)*epb), epb = total / grid,
// in chunks of at most kPushEdgesMax.  Mode 0 sets the next flags
// (edgeMapDenseForward), 1 appends to the next list and adds the destinations'
// out-degrees to ctr[1..32] (edgeMapSparseV3), 2 is mode 0 with the edge's index in
// its source's list passed to updateAtomic (edgeMapDenseBP).
constexpr int kWin = 1024;
constexpr int kPushEdgesMax = 2048;  // edges per chunk (the block's append stage)
constexpr int kPushBlocks = 2048;    // grid of the push (8 workgroups per CU)

// The last j in [0, count] with scan[j] <= key (scan non-decreasing, scan[count]
// > key), found by the whole block: each step samples the remaining range at kT
// evenly spaced entries with one load per thread and keeps the interval between
// the last sample <= key and the next (three steps for a million entries, where a
// single thread's binary search is a chain of twenty dependent loads).  Every
// thread of the block calls it and gets the result.
__device__ __forceinline__ int64_t block_search(const uint32_t *scan, int64_t count, uint32_t key) {
    int64_t lo = 0, hi = count + 1;  // scan[lo] <= key < scan[hi] (scan[count + 1] taken as infinite)
    while (hi - lo > 1) {  // block-uniform
        const int64_t step = (hi - lo + kT - 1) / kT;
        const int64_t p = lo + (int64_t)threadIdx.x * step;
        const int below = p < hi && scan[p] <= key;  // true for a prefix of the threads
        const int c = __syncthreads_count(below);    // at least 1: scan[lo] <= key
        const int64_t nlo = lo + (int64_t)(c - 1) * step;
        hi = min(hi, nlo + step);
        lo = nlo;
    }
    return lo;
}

template <bool W, int Mode, bool Small, class F>
__global__ __launch_bounds__(kT) void k_push_edges(F f, pg_view v, const intT *list,
                                                   const uint32_t *scan, int64_t count, bool *next,
                                                   intT *out_list, size_t cap,
                                                   unsigned long long *ctr) {
    __shared__ uint32_t s_scan[kWin + 1];
    __shared__ int32_t s_beg[kWin];
    __shared__ int64_t s_j0, s_j1;
    // Mode 1: a chunk's appends are staged here and placed with one atomic on
    // ctr[0] per chunk (a counter shared by every wave of the grid serialises)
    __shared__ intT s_out[Mode == 1 ? kPushEdgesMax : 1];
    __shared__ uint32_t s_cnt;
    __shared__ unsigned long long s_base;
    uint32_t total;
    if constexpr (Small) {
        // a list of at most kWin vertices: every block forms the degree prefix of
        // the whole list in LDS (kWin / kT consecutive entries per thread), so no
        // scan kernels run before the push
        __shared__ uint32_t sm[kT / 64 + 1];
        uint32_t d[kWin / kT], run = 0;
#pragma unroll
        for (int k = 0; k < kWin / kT; k++) {
            const int64_t i = (int64_t)threadIdx.x * (kWin / kT) + k;
            int32_t b = 0, e = 0;
            if (i < count) {
                const intT u = list[i];
                b = v.out_off[u];
                e = v.out_off[u + 1];
                s_beg[i] = b;
            }
            d[k] = (uint32_t)(e - b);
            run += d[k];
        }
        uint32_t tot = 0;
        uint32_t x = block_excl_scan(run, sm, &tot);
#pragma unroll
        for (int k = 0; k < kWin / kT; k++) {
            const int64_t i = (int64_t)threadIdx.x * (kWin / kT) + k;
            if (i < count) s_scan[i] = x;
            x += d[k];
        }
        if (threadIdx.x == 0) s_scan[count] = tot;
        __syncthreads();
        total = tot;
    } else {
        total = scan[count];
    }
    const uint32_t epb = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t B0 = min(total, (uint32_t)blockIdx.x * epb);
    const uint32_t B1 = min(total, B0 + epb);
    const intE *w = W ? static_cast<const intE *>(v.out_w) : nullptr;
    unsigned long long outdeg = 0;
    for (uint32_t E0 = B0; E0 < B1; E0 += kPushEdgesMax) {  // block-uniform
        const uint32_t E1 = min(B1, E0 + kPushEdgesMax);
        if constexpr (Small) {
            if (threadIdx.x == 0) {
                s_cnt = 0;
                s_j0 = 0;
                s_j1 = count - 1;
            }
        } else {
            // the list entries holding edges E0 and E1 - 1
            const int64_t j0 = block_search(scan, count, E0);
            const int64_t j1 = block_search(scan, count, E1 - 1);
            if (threadIdx.x == 0) {
                s_cnt = 0;
                s_j0 = j0;
                s_j1 = j1;
            }
        }
        __syncthreads();
        const int64_t j0 = s_j0, nwin = s_j1 - s_j0 + 1;
        const bool staged = Small || nwin <= kWin;
        if (!Small && staged) {
            for (int64_t i = threadIdx.x; i < nwin; i += kT) {
                s_scan[i] = scan[j0 + i];
                s_beg[i] = v.out_off[list[j0 + i]];
            }
        }
        __syncthreads();
        for (uint32_t k = E0 + threadIdx.x; k < E1; k += kT) {
            int64_t j;
            int32_t beg;
            uint32_t sc;
            if (staged) {
                int lo = 0, hi = (int)nwin - 1;  // last i with s_scan[i] <= k
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_scan[mid] <= k) lo = mid; else hi = mid - 1;
                }
                j = j0 + lo;
                beg = s_beg[lo];
                sc = s_scan[lo];
            } else {
                int64_t lo = j0, hi = s_j1;
                while (lo < hi) {
                    const int64_t mid = (lo + hi + 1) >> 1;
                    if (scan[mid] <= k) lo = mid; else hi = mid - 1;
                }
                j = lo;
                sc = scan[lo];
                beg = v.out_off[list[lo]];
            }
            const intT src = list[j];
            const int64_t e = (int64_t)beg + (k - sc);
            const intT ngh = v.out_adj[e];
            bool act;
            if constexpr (Mode == 2)
                act = f.cond(ngh) && f.updateAtomic(src, ngh, (intT)(k - sc));
            else
                act = f.cond(ngh) && call_update_atomic<W>(f, src, ngh, w, e);
            if constexpr (Mode == 1) {
                const unsigned long long mask = __ballot(act);
                if (mask) {
                    const int l = lane(), leader = __ffsll((long long)mask) - 1;
                    uint32_t at = 0;
                    if (l == leader) at = atomicAdd(&s_cnt, (uint32_t)__popcll(mask));
                    at = (uint32_t)__shfl((int)at, leader, 64);
                    if (act) s_out[at + __popcll(mask & ((1ull << l) - 1ull))] = ngh;
                }
                if (act) outdeg += (unsigned long long)(v.out_off[ngh + 1] - v.out_off[ngh]);
            } else {
                if (act) next[ngh] = true;
            }
        }
        if constexpr (Mode == 1) {
            __syncthreads();
            if (threadIdx.x == 0)
                s_base = s_cnt ? atomicAdd(&ctr[0], (unsigned long long)s_cnt) : 0ull;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < s_cnt; i += kT)
                if (s_base + i < cap) out_list[s_base + i] = s_out[i];
        }
        __syncthreads();  // the chunk's LDS is reused by the next one
    }
    if constexpr (Mode == 1) {
        // the successes' out-degree sum into slot pair blockIdx mod kTotSlots (ctr[0],
        // the append position, is the first word of pair 0)
        for (int o = 32; o >= 1; o >>= 1) outdeg += __shfl_xor(outdeg, o, 64);
        if (lane() == 0 && outdeg) atomicAdd(&ctr[2 * (blockIdx.x % kTotSlots) + 1], outdeg);
    }
}

template <bool W, int Mode, class F>
inline void push_list(F &f, const pg_view &v, const intT *list, int64_t count, bool *next,
                      intT *out_list, size_t cap, long long edges_hint = -1) {
    if (count <= 0) return;
    const Totals &t = totals();  // Mode 1: successes and their out-degree sum in the slots
    if (count <= kWin) {
        // the degree prefix in each block's LDS; the grid from the frontier's
        // out-degree sum when the caller knows it (a stale hint costs time only)
        unsigned grid = kPushBlocks;
        if (edges_hint > 0)
            grid = (unsigned)std::min<long long>(kPushBlocks, (edges_hint + 1023) / 1024);
        hipLaunchKernelGGL((k_push_edges<W, Mode, true, F>), dim3(grid), dim3(kT), 0, lib_stream(), f,
                           v, list, (const uint32_t *)nullptr, count, next, out_list, cap, t.slots);
    } else {
        uint32_t *scan = workspace_t<uint32_t>("push_scan", (size_t)count + 1);
        degree_scan(list, count, v.out_off, scan);
        hipLaunchKernelGGL((k_push_edges<W, Mode, false, F>), dim3(kPushBlocks), dim3(kT), 0,
                           lib_stream(), f, v, list, (const uint32_t *)scan, count, next, out_list,
                           cap, t.slots);
    }
    launched("edge push");
    if constexpr (Mode == 1) finish_totals();
}

// Dense pull over the in-degree bins of the graph (Bins).  A destination with
// cond(d) scans its in-neighbours; an active one calling updateAtomic(s, d) with
// success marks d in next; the scan stops once cond(d) is false (checked after
// every neighbour for a thread, after every group of G edges for a group).
__device__ __forceinline__ int bin_of(int deg) {
    return deg == 0 ? 4 : deg <= 16 ? 0 : deg <= 256 ? 1 : deg <= 4096 ? 2 : 3;
}

// per-bin counts: reduced in the block, four atomics per block
__global__ __launch_bounds__(kT) void k_bin_count(const int32_t *in_off, int64_t n,
                                                  unsigned long long *cnt) {
    __shared__ unsigned long long sm[4];
    if (threadIdx.x < 4) sm[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long c[4] = {0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int b = bin_of(in_off[i + 1] - in_off[i]);
        if (b < 4) c[b]++;
    }
    for (int b = 0; b < 4; b++)
        if (c[b]) atomicAdd(&sm[b], c[b]);
    __syncthreads();
    if (threadIdx.x < 4 && sm[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], sm[threadIdx.x]);
}

// scatter into the bins with one atomic per wavefront and bin
__global__ __launch_bounds__(kT) void k_bin_fill(const int32_t *in_off, int64_t n,
                                                 unsigned long long *pos, intT *ids) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t rounds = (n + stride - 1) / stride;  // the same for every lane
    for (int64_t r = 0; r < rounds; r++) {
        const int64_t i = first + r * stride;
        const int b = i < n ? bin_of(in_off[i + 1] - in_off[i]) : 4;
        for (int k = 0; k < 4; k++) {
            const unsigned long long mask = __ballot(b == k);
            if (!mask) continue;
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            if (lane() == leader) base = atomicAdd(&pos[k], (unsigned long long)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (b == k) ids[base + __popcll(mask & ((1ull << lane()) - 1ull))] = (intT)i;
        }
    }
}

__global__ void k_first_in(const int32_t *in_off, const intE *in_adj, int64_t n, intT *first_in) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        first_in[v] = in_off[v] < in_off[v + 1] ? in_adj[in_off[v]] : -1;
}

inline const Bins &bins_of(const pg_view &v) {
    auto it = bins_cache().find(v.in_off);
    if (it != bins_cache().end()) return it->second;
    Bins b;
    unsigned long long *cnt = workspace_t<unsigned long long>("bin_cnt", 8);
    check_hip(hipMemsetAsync(cnt, 0, 8 * sizeof(unsigned long long), lib_stream()), "hipMemsetAsync");
    hipLaunchKernelGGL(k_bin_count, dim3(blocks_for(v.n, kT, 4096)), dim3(kT), 0, lib_stream(),
                       v.in_off, (int64_t)v.n, cnt);
    unsigned long long h[8];
    toHost(h, cnt, 8);
    // heaviest bins first, so the workgroup-per-destination bin starts first
    int64_t s = 0;
    unsigned long long pos[8] = {0};
    for (int g = 3; g >= 0; g--) {
        b.start[g] = s;
        b.count[g] = (int64_t)h[g];
        pos[g] = (unsigned long long)s;
        s += (int64_t)h[g];
    }
    b.nonzero = s;
    b.ids = dev_alloc<intT>((size_t)s + 1);
    toDevice(cnt, pos, 8);
    hipLaunchKernelGGL(k_bin_fill, dim3(blocks_for(v.n, kT, 4096)), dim3(kT), 0, lib_stream(),
                       v.in_off, (int64_t)v.n, cnt, b.ids);
    b.first_in = dev_alloc<intT>((size_t)v.n + 1);
    hipLaunchKernelGGL(k_first_in, dim3(blocks_for(v.n)), dim3(kT), 0, lib_stream(), v.in_off,
                       v.in_adj, (int64_t)v.n, b.first_in);
    launched("k_bin_fill");
    sync();
    return bins_cache().emplace(v.in_off, b).first->second;
}

template <int G>
__device__ __forceinline__ bool group_any(bool p) {
    const unsigned long long bal = __ballot(p);
    if constexpr (G >= 64) {
        return bal != 0ull;
    } else {
        const int base = lane() & ~(G - 1);
        return ((bal >> base) & ((1ull << G) - 1ull)) != 0ull;
    }
}

// the lanes of this lane's group of G whose predicate holds, as a mask in the group
template <int G>
__device__ __forceinline__ unsigned long long group_ballot(bool p) {
    const unsigned long long bal = __ballot(p);
    if constexpr (G >= 64) return bal;
    else return (bal >> (lane() & ~(G - 1))) & ((1ull << G) - 1ull);
}

// The current frontier as a bitmap (n/8 bytes, held in L2) for the dense pulls'
// per-edge source test: the frontier's flag bytes (n bytes) made that test a third
// random line fetch per in-edge beside the functor's own gathers.
struct FrontierBits {
    const uint32_t *w;
    __device__ bool operator[](intT s) const { return (w[s >> 5] >> (s & 31)) & 1u; }
};

// w[k] = the flags b[32k .. 32k+31], a thread per word (two 16-flag loads)
__global__ void k_flag_bits(const bool *b, int64_t n, uint32_t *w) {
    const int64_t nw = (n + 31) / 32;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nw;
         k += (int64_t)gridDim.x * blockDim.x)
        w[k] = flag_mask16(b, 32 * k, n) | (flag_mask16(b, 32 * k + 16, n) << 16);
}

template <bool W, class F>
__device__ __forceinline__ bool call_update(F &f, intT s, intT d, const intE *w, int64_t e) {
    if constexpr (W)
        return f.update(s, d, w[e]);
    else
        return f.update(s, d);
}

// Dense pull (edgeMapDense, polymer.h:924-968) in id order.  A lane owns a
// destination d with in-edges and cond(d): it tries d's first in-neighbour (from
// first_in, read like cond's and the offsets' loads, coalesced), then, while cond(d)
// holds, the rest of d's in-list in list order, four ids per step (a step's loads
// issued together; the scan stops at the first step after cond(d) turns false),
// when at most kLaneRest remain; a longer rest is marked in qbits (each wave writes its two
// words whole: no atomics, nothing to clear) for k_pull_queue.  The owning lane calls
// update (the reference's non-atomic form: no other lane touches d here) and stops
// as soon as cond(d) is false, so a functor whose first success ends the scan (the
// BFS claim) updates d from the first active in-neighbour in list order.
constexpr int kLaneRest = 32;

// (Count: the destinations this pass marks in next and their out-degree sum go to
// the totals slots, for an output the caller cleared before)
template <bool Count>
__device__ __forceinline__ void count_to_slots(unsigned long long c, unsigned long long e,
                                               unsigned long long *slots) {
    if constexpr (Count) {
        __shared__ unsigned long long sc[kT / 64], se[kT / 64];
        for (int o = 32; o > 0; o >>= 1) {
            c += __shfl_xor(c, o, 64);
            e += __shfl_xor(e, o, 64);
        }
        if (lane() == 0) {
            sc[threadIdx.x >> 6] = c;
            se[threadIdx.x >> 6] = e;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kT / 64; w++) {
                c += sc[w];
                e += se[w];
            }
            unsigned long long *slot = slots + 2 * (blockIdx.x % kTotSlots);
            if (c) atomicAdd(&slot[0], c);
            if (e) atomicAdd(&slot[1], e);
        }
    }
}

template <bool W, bool Count, class F>
__global__ __launch_bounds__(kT) void k_pull_first(F f, pg_view v, const intT *first_in,
                                                   FrontierBits cur, bool *next, uint32_t *qbits,
                                                   unsigned long long *slots) {
    const int64_t n = v.n;
    unsigned long long cnt = 0, deg = 0;
    const intE *w = W ? static_cast<const intE *>(v.in_w) : nullptr;
    const int64_t stride = (int64_t)gridDim.x * kT;
    for (int64_t base = (int64_t)blockIdx.x * kT; base < n; base += stride) {  // block-uniform
        const int64_t d64 = base + threadIdx.x;
        const intT d = (intT)d64;
        const intT s0 = d64 < n ? first_in[d64] : -1;
        bool queued = false;
        if (s0 >= 0 && f.cond(d)) {
            // the offsets are loaded only when the rest of the list is needed (a hit
            // on the first in-neighbour usually ends the scan), or first when the
            // update takes the first edge's weight
            int beg = W ? v.in_off[d] : 0, end = 0;
            bool hit = cur[s0] && call_update<W>(f, s0, d, w, beg);
            const bool more = f.cond(d);
            if (more) {
                beg = v.in_off[d];
                end = v.in_off[d + 1];
            }
            if (more && end - beg > 1) {
                if (end - beg - 1 > kLaneRest) {
                    queued = true;
                } else {
                    bool go = true;
                    for (int e0 = beg + 1; go && e0 < end; e0 += 4) {
                        intT sv[4];
#pragma unroll
                        for (int k = 0; k < 4; k++) sv[k] = e0 + k < end ? v.in_adj[e0 + k] : -1;
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            if (go && sv[k] >= 0 && cur[sv[k]]) {
                                if (call_update<W>(f, sv[k], d, w, e0 + k)) hit = true;
                                if (!f.cond(d)) go = false;
                            }
                    }
                }
            }
            if (hit) {
                next[d] = true;
                if constexpr (Count) {
                    cnt++;
                    deg += (unsigned long long)(v.out_off[d + 1] - v.out_off[d]);
                }
            }
        }
        const unsigned long long qm = __ballot(queued);
        const int64_t wd = (base + (threadIdx.x & ~63)) >> 5;  // this wave's first word
        const int64_t nq = (n + 31) / 32;
        if (lane() == 0 && wd < nq) qbits[wd] = (uint32_t)qm;
        if (lane() == 1 && wd + 1 < nq) qbits[wd + 1] = (uint32_t)(qm >> 32);
    }
    count_to_slots<Count>(cnt, deg, slots);
}

// The destinations k_pull_first marked, a wave each (a wave takes the marks of 64
// destinations at a time), over their in-lists from the second entry, 64 entries
// per step in list order: the first active entry of a step is applied alone, then,
// if cond(d) still holds, the step's other active entries together; the scan stops
// once cond(d) is false.  Lanes of a wave may update d at once here, so this part
// calls updateAtomic.
template <bool W, bool Count, class F>
__global__ __launch_bounds__(kT) void k_pull_queue(F f, pg_view v, FrontierBits cur, bool *next,
                                                   const uint32_t *qbits,
                                                   unsigned long long *slots) {
    const int64_t ngroups = (v.n + 63) / 64, nq = (v.n + 31) / 32;
    unsigned long long cnt = 0, deg = 0;
    const int l = lane();
    const intE *w = W ? static_cast<const intE *>(v.in_w) : nullptr;
    const int64_t nwaves = ((int64_t)gridDim.x * kT) >> 6;
    for (int64_t g = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6; g < ngroups; g += nwaves)
    for (unsigned long long qm = (unsigned long long)qbits[2 * g] |
                                 (2 * g + 1 < nq ? (unsigned long long)qbits[2 * g + 1] << 32 : 0ull);
         qm; qm &= qm - 1ull) {
        const intT d = (intT)(64 * g + __ffsll((long long)qm) - 1);
        const int beg = v.in_off[d] + 1, end = v.in_off[d + 1];
        bool hit = false;
        for (int base = beg; base < end; base += 64) {
            const int e = base + l;
            intT s = 0;
            bool a = false;
            if (e < end) {
                s = v.in_adj[e];
                a = cur[s];
            }
            const unsigned long long am = __ballot(a);
            if (am) {
                const int first = __ffsll((long long)am) - 1;
                if (l == first && call_update_atomic<W>(f, s, d, w, e)) hit = true;
                if (!group_any<64>(f.cond(d))) break;
                if (a && l != first && call_update_atomic<W>(f, s, d, w, e)) hit = true;
            }
            if (!group_any<64>(f.cond(d))) break;
        }
        if (group_any<64>(hit) && l == 0) {
            next[d] = true;
            if constexpr (Count) {
                cnt++;
                deg += (unsigned long long)(v.out_off[d + 1] - v.out_off[d]);
            }
        }
    }
    count_to_slots<Count>(cnt, deg, slots);
}

// (counted: next was cleared before, and the number of destinations the pull marks
// and their out-degree sum are published with finish_totals)
template <bool W, class F>
inline void pull_dense(F &f, const pg_view &v, FrontierBits cur, bool *next, bool counted = false) {
    const Bins &b = bins_of(v);
    uint32_t *qbits = workspace_t<uint32_t>("pull_marks", (size_t)(v.n + 31) / 32 + 1);
    const unsigned gq = blocks_for((v.n + 63) / 64 * 64, kT, 8192);
    if (counted) {
        unsigned long long *slots = totals().slots;
        // at most 8192 blocks, so the slots take few atomics
        hipLaunchKernelGGL((k_pull_first<W, true, F>), dim3(blocks_for(v.n, kT, 8192)), dim3(kT), 0,
                           lib_stream(), f, v, (const intT *)b.first_in, cur, next, qbits, slots);
        hipLaunchKernelGGL((k_pull_queue<W, true, F>), dim3(gq), dim3(kT), 0, lib_stream(), f, v, cur,
                           next, (const uint32_t *)qbits, slots);
        launched("dense pull");
        finish_totals();
    } else {
        hipLaunchKernelGGL((k_pull_first<W, false, F>), dim3(blocks_for(v.n)), dim3(kT), 0, lib_stream(),
                           f, v, (const intT *)b.first_in, cur, next, qbits,
                           (unsigned long long *)nullptr);
        hipLaunchKernelGGL((k_pull_queue<W, false, F>), dim3(gq), dim3(kT), 0, lib_stream(), f, v, cur,
                           next, (const uint32_t *)qbits, (unsigned long long *)nullptr);
        launched("dense pull");
    }
}

// edgeMapDenseReduce (polymer.h:1097-1168) over the in-degree bins: the G lanes of
// a destination split its in-list, each reducing its part into its own 16-byte
// scratch (initFunc, then reduceFunc per active in-neighbour, an activation marks d
// in next), stopping once cond(d) is false; the lanes that saw an edge then fold
// their scratches into d with combineFunc one after another, in lane order.  The
// reference folds one scratch per destination; folding partial scratches gives the
// same result for combine functions that accumulate into d (its PageRank and SpMV
// add with writeAdd, its Components lowers with writeMin).
template <bool W, class F>
__device__ __forceinline__ bool reduce_edge(F &f, void *data, intT s, const intE *w, int64_t e) {
    if constexpr (W)
        return f.reduceFunc(data, s, w[e]);
    else
        return f.reduceFunc(data, s);
}

template <bool W, int G, class F>
__global__ __launch_bounds__(kT) void k_reduce_group(F f, pg_view v, FrontierBits cur, bool *next,
                                                     const intT *ids, int64_t count) {
    const int64_t slot = ((int64_t)blockIdx.x * kT + threadIdx.x) / G;
    if (slot >= count) return;  // uniform within the group
    const int l = lane() & (G - 1);
    const intT d = ids[slot];
    if (!group_any<G>(f.cond(d))) return;
    const intE *w = W ? static_cast<const intE *>(v.in_w) : nullptr;
    const int beg = v.in_off[d], end = v.in_off[d + 1];
    double data[2];
    f.initFunc((void *)data, d);
    bool hit = false, saw = false;
    if constexpr (G == 1) {
        for (int e = beg; e < end; e++) {
            saw = true;
            const intT s = v.in_adj[e];
            if (cur[s] && reduce_edge<W>(f, (void *)data, s, w, e)) hit = true;
            if (!f.cond(d)) break;
        }
    } else {
        // four chunks of G edges per cond test: their neighbour ids and flags are
        // loaded together, so four gathers per lane are in flight
        constexpr int U = 4;
        for (int base = beg; base < end; base += U * G) {
            intT s[U];
            bool a[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int e = base + k * G + l;
                s[k] = e < end ? v.in_adj[e] : 0;
                saw = saw || e < end;
            }
#pragma unroll
            for (int k = 0; k < U; k++) a[k] = base + k * G + l < end && cur[s[k]];
#pragma unroll
            for (int k = 0; k < U; k++)
                if (a[k] && reduce_edge<W>(f, (void *)data, s[k], w, base + k * G + l)) hit = true;
            if (!group_any<G>(f.cond(d))) break;
        }
    }
    if (group_any<G>(hit) && l == 0) next[d] = true;
    if constexpr (G == 1) {
        if (saw) f.combineFunc((void *)data, d);
    } else {
        for (int r = 0; r < G; r++)
            if (l == r && saw) f.combineFunc((void *)data, d);
    }
}

// a 256-thread workgroup per destination; the scratches fold wave after wave
template <bool W, class F>
__global__ __launch_bounds__(kT) void k_reduce_block(F f, pg_view v, FrontierBits cur, bool *next,
                                                     const intT *ids) {
    __shared__ int s_stop, s_hit;
    const intT d = ids[blockIdx.x];
    if (threadIdx.x == 0) {
        s_stop = f.cond(d) ? 0 : 1;
        s_hit = 0;
    }
    __syncthreads();
    if (s_stop) return;  // uniform over the block
    const intE *w = W ? static_cast<const intE *>(v.in_w) : nullptr;
    const int beg = v.in_off[d], end = v.in_off[d + 1];
    double data[2];
    f.initFunc((void *)data, d);
    bool saw = false;
    for (int base = beg; base < end; base += kT) {
        const int e = base + (int)threadIdx.x;
        if (e < end) {
            saw = true;
            const intT s = v.in_adj[e];
            if (cur[s] && reduce_edge<W>(f, (void *)data, s, w, e)) s_hit = 1;
        }
        __syncthreads();
        if (threadIdx.x == 0 && !f.cond(d)) s_stop = 1;
        __syncthreads();
        if (s_stop) break;
    }
    const int wv = threadIdx.x >> 6, l = lane();
    for (int k = 0; k < kT / 64; k++) {
        if (wv == k)
            for (int r = 0; r < 64; r++)
                if (l == r && saw) f.combineFunc((void *)data, d);
        __syncthreads();
    }
    if (threadIdx.x == 0 && s_hit) next[d] = true;
}

template <bool W, class F>
inline void reduce_dense(F &f, const pg_view &v, FrontierBits cur, bool *next) {
    const Bins &b = bins_of(v);
    const int64_t nblk = b.count[3], nwave = b.count[2], n16 = b.count[1], n1 = b.count[0];
    if (nblk > 0)
        hipLaunchKernelGGL((k_reduce_block<W, F>), dim3((unsigned)nblk), dim3(kT), 0, lib_stream(), f,
                           v, cur, next, (const intT *)(b.ids + b.start[3]));
    if (nwave > 0)
        hipLaunchKernelGGL((k_reduce_group<W, 64, F>), dim3(blocks_for(nwave * 64, kT, 1ll << 31)),
                           dim3(kT), 0, lib_stream(), f, v, cur, next,
                           (const intT *)(b.ids + b.start[2]), nwave);
    if (n16 > 0)
        hipLaunchKernelGGL((k_reduce_group<W, 16, F>), dim3(blocks_for(n16 * 16, kT, 1ll << 31)),
                           dim3(kT), 0, lib_stream(), f, v, cur, next,
                           (const intT *)(b.ids + b.start[1]), n16);
    if (n1 > 0)
        hipLaunchKernelGGL((k_reduce_group<W, 1, F>), dim3(blocks_for(n1, kT, 1ll << 31)), dim3(kT), 0,
                           lib_stream(), f, v, cur, next, (const intT *)(b.ids + b.start[0]), n1);
    launched("dense reduce");
}

// A wave's chunk of 1024 flags is loaded 16 per lane (lane l holds flags 16l ..
// 16l + 15 as a mask) and visited lane-consecutive: in step r lane l takes vertex
// 64r + l, its bit fetched from lane 4r + l/16, so the program's arrays indexed by
// the vertex are read and written coalesced.
__device__ __forceinline__ bool chunk_bit(uint32_t mk, int r) {
    const int l = lane();
    return (__shfl((int)mk, 4 * r + (l >> 4), 64) >> (l & 15)) & 1;
}

template <class F>
__global__ void k_vertex_map_dense(F f, const bool *b, int64_t n, intT offset) {
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c * 1024 < n; c += nw) {
        const int64_t base = c * 1024;  // wave-uniform
        const int64_t mine = base + 16 * lane();
        const uint32_t mk = mine < n ? flag_mask16(b, mine, n) : 0u;
        if (!__ballot(mk != 0u)) continue;
#pragma unroll 4
        for (int r = 0; r < 16; r++)
            if (chunk_bit(mk, r)) f((intT)(base + 64 * r + lane()) + offset);
    }
}

template <class F>
__global__ void k_vertex_map_sparse(F f, const intT *s, int64_t m) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        f(s[i]);
}

template <class F>
__global__ __launch_bounds__(kT) void k_vertex_filter(F f, const bool *b, int64_t lo, int64_t hi,
                                                      intT offset, bool *dst,
                                                      unsigned long long *count) {
    __shared__ unsigned long long sc[kT / 64];
    unsigned long long c = 0;
    for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi;
         i += (int64_t)gridDim.x * blockDim.x) {
        bool r = false;
        if (b[i]) r = f((intT)i + offset);
        dst[i] = r;
        c += r ? 1ull : 0ull;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane() == 0) sc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / 64; w++) c += sc[w];
        if (c && count) atomicAdd(&count[2 * (blockIdx.x % kTotSlots)], c);  // spread slots
    }
}

// |F| and the out-degree sum of a flag frontier: reduced per block, two atomics
// per block (a grid of at most 2048 blocks) into one of 16 slot pairs
__global__ __launch_bounds__(kT) void k_counter(const bool *b, int64_t n, intT offset,
                                                const int32_t *out_off, unsigned long long *out) {
    __shared__ unsigned long long sc[kT / 64], se[kT / 64];
    unsigned long long c = 0, e = 0;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t ch = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; ch * 1024 < n; ch += nw) {
        const int64_t base = ch * 1024;  // wave-uniform
        const int64_t mine = base + 16 * lane();
        const uint32_t mk = mine < n ? flag_mask16(b, mine, n) : 0u;
        c += (unsigned long long)__popc(mk);
        if (!__ballot(mk != 0u)) continue;
#pragma unroll 4
        for (int r = 0; r < 16; r++)
            if (chunk_bit(mk, r)) {
                const int64_t v = base + 64 * r + lane() + offset;
                e += (unsigned long long)(out_off[v + 1] - out_off[v]);
            }
    }
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o, 64);
        e += __shfl_xor(e, o, 64);
    }
    if (lane() == 0) {
        sc[threadIdx.x >> 6] = c;
        se[threadIdx.x >> 6] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / 64; w++) {
            c += sc[w];
            e += se[w];
        }
        // kTotSlots slot pairs (count, out-degree sum): a pair shared by every block
        // of the grid serialises its atomics
        unsigned long long *slot = out + 2 * (blockIdx.x % kTotSlots);
        if (c) atomicAdd(&slot[0], c);
        if (e) atomicAdd(&slot[1], e);
    }
}

__global__ void k_set_list(bool *b, const intT *s, int64_t m, intT offset) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        b[s[i] - offset] = true;
}

// b[0, n) = v: 16-byte stores over the aligned body, byte stores for the rest
__global__ void k_fill_bytes(bool *b, int64_t n, bool v) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t gs = (int64_t)gridDim.x * blockDim.x;
    const int64_t nv = ((((uintptr_t)b) & 15u) == 0u) ? n / 16 : 0;
    const uint32_t w = v ? 0x01010101u : 0u;
    for (int64_t i = gt; i < nv; i += gs) reinterpret_cast<uint4 *>(b)[i] = make_uint4(w, w, w, w);
    for (int64_t i = 16 * nv + gt; i < n; i += gs) b[i] = v;
}

// grid of the 16-flags-per-thread kernels over n flags
inline unsigned blocks_for_flags(int64_t n, int64_t cap = 65536) {
    return blocks_for((n + 15) / 16, kT, cap);
}

}  // namespace polymer_detail

/* ---- frontiers (polymer.h:521-881) ------------------------------------------ */
struct AsyncChunk {
    intT accessCounter;
    intT m;
    intT *s;
};

// a chunk of blockSize queue entries (polymer.h:1350-1356); the asynchronous edge
// maps here seed from the frontier's flags, so chunks only carry the program's own
// bookkeeping
inline AsyncChunk *newChunk(intT blockSize) {
    AsyncChunk *c = static_cast<AsyncChunk *>(std::malloc(sizeof(AsyncChunk)));
    c->s = static_cast<intT *>(std::malloc(sizeof(intT) * (size_t)(blockSize > 0 ? blockSize : 1)));
    c->m = 0;
    c->accessCounter = 0;
    return c;
}

struct LocalFrontier {
    intT n;
    intT m;
    intT head = 0;
    intT tail = 0;
    intT insertTail = 0;
    intT emptySignal = 0;
    intT outEdgesCount;
    intT startID;
    intT endID;
    bool *b;
    intT *s;
    intT sparseCounter = 0;
    intT **sparseChunks = nullptr;
    intT *chunkSizes = nullptr;
    intT *tmp = nullptr;
    AsyncChunk **localQueue = nullptr;
    bool isDense;
    size_t sCap = 0;        // capacity of s when the header allocated it
    bool sOwned = false;    // s was allocated here
    // a dense edgeMap's own count of this frontier (read by vertexCounter while the
    // publication with this sequence number is the last and nothing changed since)
    unsigned long long pendingSeq = 0, pendingGen = 0;

    LocalFrontier(bool *_b, intT start, intT end)
        : n(end - start), m(0), outEdgesCount(0), startID(start), endID(end), b(_b), s(nullptr),
          isDense(true) {
        polymer_detail::ensure_init();
        (void)polymer_detail::totals();
        reserveSparse((size_t)n + 1);  // made here, before a program's loop, not in a round
    }
    ~LocalFrontier() {
        if (s && sOwned) polymer_detail::shared_free(s);
    }
    LocalFrontier(const LocalFrontier &) = delete;
    LocalFrontier &operator=(const LocalFrontier &) = delete;

    bool inRange(intT index) const { return startID <= index && index < endID; }
    PG_HD inline void setBit(intT index, bool val) { b[index - startID] = val; }
    PG_HD inline bool getBit(intT index) const { return b[index - startID]; }

    // a list of at least cap entries owned by the frontier, kept across calls: at
    // least n + 1 entries (every frontier fits), doubled when a call needs more, so
    // a loop of rounds allocates once (the reference mallocs and frees one per call)
    void reserveSparse(size_t cap) {
        if (s && sOwned && cap <= sCap) return;
        const size_t grown = std::max(std::max(cap, (size_t)n + 1), s && sOwned ? 2 * sCap : 0);
        if (s && sOwned) {
            polymer_detail::sync();  // kernels queued before may still read the old list
            polymer_detail::shared_free(s);
        }
        s = static_cast<intT *>(polymer_detail::shared_alloc(sizeof(intT) * grown, false));
        sCap = grown;
        sOwned = true;
    }

    // packIndex (polymer.h:552-567): ascending ids of the set flags
    void toSparse() {
        if (isDense) {
            reserveSparse((size_t)n + 1);
            m = polymer_detail::compact_flags(b, n, startID, s);
        }
        isDense = false;
    }
    void toSparseAsync(int nextID, LocalFrontier *next) {
        (void)nextID;
        (void)next;
        toSparse();
    }

    // toDense (polymer.h:594-601)
    void toDense() {
        if (!isDense) {
            using namespace polymer_detail;
            frontiers_changed();
            hipLaunchKernelGGL(k_fill_bytes, dim3(blocks_for_flags(n)), dim3(256), 0, lib_stream(), b,
                               (int64_t)n, false);
            if (m > 0) {
                const intT *sl = device_readable(s, (size_t)m, "todense_list");
                hipLaunchKernelGGL(k_set_list, dim3(blocks_for(m)), dim3(256), 0, lib_stream(), b, sl,
                                   (int64_t)m, startID);
            }
            launched("toDense");
            sync_visible();
        }
        isDense = true;
    }
    void toDenseWithMerge(int numOfSub) {
        (void)numOfSub;
        toDense();
    }

    // takes a list the caller allocated (host-visible, device or malloc'd memory)
    void setSparse(intT _m, intT *_s) {
        if (s && sOwned) polymer_detail::shared_free(s);
        m = _m;
        s = _s;
        sCap = (size_t)_m;
        sOwned = false;
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
    std::vector<bool *> d;
    std::vector<LocalFrontier *> frontiers, nextFrontiers;
    bool isDense;
    bool firstSparse;
    AsyncChunk **asyncQueue = nullptr;
    intT asyncEndSignal = 0;
    intT readerTail = 0;
    intT insertTail = 0;

    explicit vertices(int _numOfNodes)
        : n(0), m(-1), numOfNodes(_numOfNodes), numOfVertices(0),
          numOfVertexOnNode(_numOfNodes, 0), offsets(_numOfNodes + 1, 0),
          numOfNonZero(_numOfNodes, 0), d(_numOfNodes, nullptr), frontiers(_numOfNodes, nullptr),
          nextFrontiers(_numOfNodes, nullptr), isDense(true), firstSparse(false) {}

    void registerFrontier(int nodeNum, LocalFrontier *frontier) {
        frontiers[nodeNum] = frontier;
        numOfVertexOnNode[nodeNum] = frontier->n;
        isDense = frontier->isDense;
    }
    void toDense() {
        if (isDense) return;
        isDense = true;
        for (auto *f : frontiers) f->toDense();
    }
    void toDenseWithMerge(int numOfSub) {
        (void)numOfSub;
        toDense();
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
    void toSparseAsync() { toSparse(); }
    long long getEdgeStat() const {
        long long sum = 0;
        for (int i = 0; i < numOfNodes; i++) sum += frontiers[i]->outEdgesCount;
        return sum;
    }
    void calculateOffsets() {
        numOfVertices = 0;
        for (int i = 0; i < numOfNodes; i++) numOfVertices += numOfVertexOnNode[i];
        offsets[0] = 0;
        for (int i = 1; i <= numOfNodes; i++) offsets[i] = offsets[i - 1] + numOfVertexOnNode[i - 1];
        n = numOfVertices;
    }
    intT getSize(int nodeNum) const { return numOfVertexOnNode[nodeNum]; }
    intT getSparseSize(int nodeNum) const { return numOfNonZero[nodeNum]; }
    void calculateNumOfNonZero(int nodeNum) { numOfNonZero[nodeNum] = frontiers[nodeNum]->m; }
    intT numNonzeros() {
        if (m < 0) {
            intT sum = 0;
            for (int i = 0; i < numOfNodes; i++) sum += numOfNonZero[i];
            m = sum;
        }
        return m;
    }
    bool isEmpty() { return numNonzeros() == 0; }
    int getNodeNumOfIndex(intT index) const {
        int r = 0;
        while (r < numOfNodes && offsets[r] <= index) r++;
        return r - 1;
    }
    int getNodeNumOfSparseIndex(intT index) const {
        int r = 0;
        intT acc = 0;
        while (r < numOfNodes && acc <= index) acc += numOfNonZero[r++];
        return r - 1;
    }
    intT getOffset(int nodeNum) const { return offsets[nodeNum]; }
    // host access to one flag of the (host-visible) frontier arrays
    void setBit(intT index, bool bit) {
        const int i = std::max(0, getNodeNumOfIndex(index));
        frontiers[i]->b[index - offsets[i]] = bit;
    }
    bool getBit(intT index) const {
        const int i = std::max(0, getNodeNumOfIndex(index));
        return frontiers[i]->b[index - offsets[i]];
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
    bool eq(vertices &) { return false; }
    void print() {}
    void del() {}
};

/* ---- edge maps ---------------------------------------------------------------------- */
namespace polymer_detail {

// the frontier's flags as the kernels read them (the calls' node is 0: one GPU)
inline const bool *cur_flags(vertices *frontier) { return frontier->getArr(0); }

// the frontier's flags as a bitmap (a workspace rebuilt on each dense call)
inline FrontierBits cur_bits(vertices *frontier, int64_t n) {
    const int64_t nw = (n + 31) / 32;
    uint32_t *w = workspace_t<uint32_t>("frontier_bits", (size_t)nw + 1);
    hipLaunchKernelGGL(k_flag_bits, dim3(blocks_for(nw, kT)), dim3(kT), 0, lib_stream(),
                       cur_flags(frontier), n, w);
    launched("frontier bitmap");
    return FrontierBits{w};
}

// the active sources of a dense frontier as an ordered list (kept in a workspace)
inline intT list_active(vertices *frontier, intT **list) {
    LocalFrontier *f = frontier->getFrontier(0);
    intT *l = workspace_t<intT>("active_list", (size_t)f->n + 1);
    *list = l;
    return compact_flags(f->b, f->n, f->startID, l);
}

template <class G, class F, int Mode>
void dense_push(G &GA, vertices *frontier, F &f, LocalFrontier *next) {
    constexpr bool W = is_weighted<G>::value && Mode != 2;
    frontiers_changed();
    intT *list = nullptr;
    const intT count = list_active(frontier, &list);
    push_list<W, Mode>(f, GA.view, list, count, next->b, nullptr, 0,
                       frontier->getFrontier(0)->outEdgesCount);
    sync_visible();
}

}  // namespace polymer_detail

// edgeMapDenseForward (polymer.h:970-1032): for every active s and out-edge
// (s, ngh): if cond(ngh) && updateAtomic(s, ngh) then ngh joins next.
template <class F, class G>
bool *edgeMapDenseForward(G GA, vertices *frontier, F f, LocalFrontier *next, bool part = false,
                          intT start = 0, intT end = 0) {
    (void)part; (void)start; (void)end;
    polymer_detail::dense_push<G, F, 0>(GA, frontier, f, next);
    return nullptr;
}

// edgeMapDenseForwardDynamic (polymer.h:1036-1095): the same push
template <class F, class G>
bool *edgeMapDenseForwardDynamic(G GA, vertices *frontier, F f, LocalFrontier *next,
                                 Subworker_Partitioner &subworker = dummyPartitioner) {
    (void)subworker;
    polymer_detail::dense_push<G, F, 0>(GA, frontier, f, next);
    return nullptr;
}

// edgeMapDenseBP (polymer.h:1263-1306): the forward push with the edge's index
// j in the source's out-list passed to updateAtomic(s, ngh, j).
template <class F, class G>
bool *edgeMapDenseBP(G GA, vertices *frontier, F f, LocalFrontier *next, bool part = false,
                     intT start = 0, intT end = 0) {
    (void)part; (void)start; (void)end;
    polymer_detail::dense_push<G, F, 2>(GA, frontier, f, next);
    return nullptr;
}

// edgeMapDense / edgeMapDenseDynamic (polymer.h:924-968, 1170-1261)
template <class F, class G>
bool *edgeMapDense(G GA, vertices *frontier, F f, LocalFrontier *next, bool parallel = 0,
                   Subworker_Partitioner &subworker = dummyPartitioner) {
    (void)parallel; (void)subworker;
    constexpr bool W = polymer_detail::is_weighted<G>::value;
    polymer_detail::frontiers_changed();
    frontier->nextFrontiers[0] = next;
    polymer_detail::pull_dense<W>(f, GA.view, polymer_detail::cur_bits(frontier, GA.n), next->b);
    polymer_detail::sync_visible();
    return nullptr;
}

template <class F, class G>
bool *edgeMapDenseDynamic(G GA, vertices *frontier, F f, LocalFrontier *next,
                          Subworker_Partitioner &subworker = dummyPartitioner) {
    return edgeMapDense(GA, frontier, f, next, 0, subworker);
}

// edgeMapDenseReduce (polymer.h:1097-1168)
template <class F, class G>
bool *edgeMapDenseReduce(G GA, vertices *frontier, F f, LocalFrontier *next, bool parallel = 0,
                         Subworker_Partitioner &subworker = dummyPartitioner) {
    using namespace polymer_detail;
    (void)parallel; (void)subworker;
    constexpr bool W = is_weighted<G>::value;
    frontiers_changed();
    frontier->nextFrontiers[0] = next;
    reduce_dense<W>(f, GA.view, cur_bits(frontier, GA.n), next->b);
    sync_visible();
    return nullptr;
}

// edgeMapSparseV3 (polymer.h:1928-2003): the active list pushes; each success
// appends ngh to next's list (duplicates possible, as in the reference) and adds
// outdeg(ngh) to next->outEdgesCount.  The list holds the frontier's out-degree
// sum (getEdgeStat); a stale count that would overflow it is an error.
template <class F, class G>
void edgeMapSparseV3(G GA, vertices *frontier, F f, LocalFrontier *next, bool part = false,
                     Subworker_Partitioner &subworker = dummyPartitioner) {
    using namespace polymer_detail;
    (void)part; (void)subworker;
    constexpr bool W = is_weighted<G>::value;
    frontiers_changed();
    LocalFrontier *cur = frontier->getFrontier(0);
    const long long bufferLen = frontier->getEdgeStat();
    next->reserveSparse((size_t)(bufferLen > 0 ? bufferLen : 1));
    const int64_t count = cur->m;
    unsigned long long c = 0, e = 0;
    if (count > 0) {
        const intT *list = device_readable(cur->s, (size_t)count, "sparse_in");
        push_list<W, 1>(f, GA.view, list, count, nullptr, next->s, next->sCap, bufferLen);
        read_totals(&c, &e);
    }
    if (c > next->sCap) {
        std::fprintf(stderr,
                     "polymer: edgeMapSparseV3: %llu successes exceed the list of %zu entries sized "
                     "from getEdgeStat()\n",
                     c, next->sCap);
        std::abort();
    }
    next->m = (intT)c;
    next->outEdgesCount = (intT)e;
    next->isDense = false;
}

template <class F, class G>
void edgeMapSparseV2(G GA, vertices *frontier, F f, LocalFrontier *next, bool part = false,
                     Subworker_Partitioner &subworker = dummyPartitioner) {
    edgeMapSparseV3(GA, frontier, f, next, part, subworker);
}
template <class F, class G>
void edgeMapSparseV4(G GA, vertices *frontier, F f, LocalFrontier *next, bool firstTime = false,
                     Subworker_Partitioner &subworker = dummyPartitioner) {
    (void)firstTime;
    edgeMapSparseV3(GA, frontier, f, next, false, subworker);
}
template <class F, class G>
void edgeMapSparseV5(G GA, vertices *frontier, F f, LocalFrontier *next,
                     Subworker_Partitioner &subworker = dummyPartitioner) {
    edgeMapSparseV3(GA, frontier, f, next, false, subworker);
}

// edgeMapSparseAsync / edgeMapSparseAsyncPipe (polymer.h:1358-1751): every vertex
// the push activates is expanded within the same call, until no activation is
// left (the reference passes discoveries between nodes' queues and ends with an
// end-game chunk).  Here the discoveries are pushed again, list after list, on the
// GPU; the call ends with next empty.  For functors whose updates are monotone
// (BFS claims, Bellman-Ford's writeMin) the end state is the reference's; which
// claimer wins a CAS is not fixed in either.
template <class F, class G>
void edgeMapSparseAsyncPipe(G GA, vertices *frontier, F f, LocalFrontier *next,
                            Subworker_Partitioner &subworker = dummyPartitioner) {
    using namespace polymer_detail;
    (void)subworker;
    constexpr bool W = is_weighted<G>::value;
    frontiers_changed();
    LocalFrontier *cur = frontier->getFrontier(0);
    if (cur->isDense) cur->toSparse();
    const size_t cap = (size_t)std::max(GA.m, GA.n) + 1;
    intT *qa = workspace_t<intT>("async_a", cap);
    intT *qb = workspace_t<intT>("async_b", cap);
    int64_t count = cur->m;
    if (count > 0)
        check_hip(hipMemcpyAsync(qa, device_readable(cur->s, (size_t)count, "async_in"),
                                 sizeof(intT) * (size_t)count, hipMemcpyDefault, lib_stream()),
                  "hipMemcpyAsync");
    while (count > 0) {
        push_list<W, 1>(f, GA.view, qa, count, nullptr, qb, cap);
        unsigned long long c = 0;
        read_totals(&c, nullptr);
        if (c > cap) fail("edgeMapSparseAsyncPipe: more activations in one pass than edges");
        count = (int64_t)c;
        std::swap(qa, qb);
    }
    next->m = 0;
    next->outEdgesCount = 0;
    next->isDense = false;
}

template <class F, class G>
void edgeMapSparseAsync(G GA, vertices *frontier, F f, LocalFrontier *next,
                        Subworker_Partitioner &subworker = dummyPartitioner) {
    edgeMapSparseAsyncPipe(GA, frontier, f, next, subworker);
}

void clearLocalFrontier(LocalFrontier *next, int nodeNum, int subNum, int totalSub);

// edgeMap (polymer.h:2241-2309): dense iff |F| + sum of out-degrees of F >= threshold
template <class F, class G>
void edgeMap(G GA, vertices *V, F f, LocalFrontier *next, intT threshold = -1, char option = DENSE,
             bool remDups = false, bool part = false,
             Subworker_Partitioner &subworker = dummyPartitioner) {
    (void)remDups;
    const long long m = (long long)V->numNonzeros() + V->getEdgeStat();
    if (m >= threshold) {
        V->toDense();
        next->sparseCounter = 0;
        clearLocalFrontier(next, 0, 0, 1);
        if (option == DENSE_FORWARD) {
            edgeMapDenseForward(GA, V, f, next, part, subworker.dense_start, subworker.dense_end);
        } else {
#ifdef POLYMER_DEVICE_RESIDENT
            // next was just cleared: the pull counts what it marks, for vertexCounter
            // (with managed arrays the program's host code may write next's flags
            // before it counts them, so the count is taken there instead)
            using namespace polymer_detail;
            constexpr bool W = is_weighted<G>::value;
            V->nextFrontiers[0] = next;
            pull_dense<W>(f, GA.view, cur_bits(V, GA.n), next->b, true);
            next->pendingSeq = totals().seq;
            next->pendingGen = frontier_gen();
#else
            edgeMapDense(GA, V, f, next, option, subworker);
#endif
        }
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
    const unsigned long long pending = frontier->pendingSeq;
    frontier->pendingSeq = 0;
    if (pending && pending == totals().seq && frontier->pendingGen == frontier_gen()) {
        unsigned long long cnt = 0, e = 0;  // the dense edgeMap that made it counted it
        read_totals(&cnt, &e);
        frontier->m = (intT)cnt;
        frontier->outEdgesCount = (intT)e;
        return;
    }
    hipLaunchKernelGGL(k_counter, dim3(blocks_for_flags(frontier->n, 2048)), dim3(256), 0, lib_stream(),
                       (const bool *)frontier->b, (int64_t)frontier->n, frontier->startID,
                       GA.view.out_off, totals().slots);
    launched("vertexCounter");
    finish_totals();
    unsigned long long cnt = 0, e = 0;
    read_totals(&cnt, &e);
    frontier->m = (intT)cnt;
    frontier->outEdgesCount = (intT)e;
}

// vertexMap (polymer.h:2340-2381): f(i) for every active vertex
template <class F>
void vertexMap(vertices *V, F add, int nodeNum, int subNum = 0, int totalSub = 1) {
    using namespace polymer_detail;
    (void)totalSub;
    if (subNum != 0) return;
    frontiers_changed();  // the functor may write anything
    LocalFrontier *fr = V->getFrontier(nodeNum);
    if (V->isDense) {
        hipLaunchKernelGGL(k_vertex_map_dense<F>, dim3(blocks_for_flags(fr->n)), dim3(256), 0, lib_stream(),
                           add, (const bool *)fr->b, (int64_t)fr->n, V->getOffset(nodeNum));
    } else if (fr->m > 0) {
        const intT *sl = device_readable(fr->s, (size_t)fr->m, "vmap_list");
        hipLaunchKernelGGL(k_vertex_map_sparse<F>, dim3(blocks_for(fr->m)), dim3(256), 0, lib_stream(),
                           add, sl, (int64_t)fr->m);
    }
    launched("vertexMap");
    sync_visible();
}

// clearLocalFrontier (polymer.h:2383-2397)
inline void clearLocalFrontier(LocalFrontier *next, int nodeNum, int subNum, int totalSub) {
    using namespace polymer_detail;
    (void)nodeNum; (void)totalSub;
    if (subNum != 0) return;
    frontiers_changed();
    hipLaunchKernelGGL(k_fill_bytes, dim3(blocks_for_flags(next->n)), dim3(256), 0, lib_stream(), next->b,
                       (int64_t)next->n, false);
    launched("clearLocalFrontier");
    sync_visible();
}

// Every vertex of the frontier's range active (the apps' "all true" start,
// e.g. numa-PageRank.C:427-429); m and outEdgesCount come from vertexCounter.
inline void fillLocalFrontier(LocalFrontier *f) {
    using namespace polymer_detail;
    frontiers_changed();
    hipLaunchKernelGGL(k_fill_bytes, dim3(blocks_for_flags(f->n)), dim3(256), 0, lib_stream(), f->b,
                       (int64_t)f->n, true);
    launched("fillLocalFrontier");
    sync_visible();
    f->isDense = true;
    f->m = f->n;
}

// vertexFilter (polymer.h:2399-2409): result[i] = active(i) && filter(i) over the
// node's range (a host-visible or device bool array)
template <class F>
void vertexFilter(vertices *V, F filter, int nodeNum, bool *result) {
    using namespace polymer_detail;
    frontiers_changed();
    V->toDense();
    LocalFrontier *fr = V->getFrontier(nodeNum);
    hipLaunchKernelGGL(k_vertex_filter<F>, dim3(blocks_for(fr->n, 256, 2048)), dim3(256), 0, lib_stream(), filter,
                       (const bool *)fr->b, (int64_t)0, (int64_t)fr->n, V->getOffset(nodeNum), result,
                       (unsigned long long *)nullptr);  // no count
    launched("vertexFilter");
    sync_visible();
}

// vertexFilter (polymer.h:2411-2441): the same into a frontier; the count of
// passing vertices is added to result->m
template <class F>
void vertexFilter(vertices *V, F filter, int nodeNum, int subNum, int totalSub, LocalFrontier *result) {
    using namespace polymer_detail;
    (void)totalSub;
    if (subNum != 0) return;
    frontiers_changed();
    V->toDense();
    LocalFrontier *fr = V->getFrontier(nodeNum);
    hipLaunchKernelGGL(k_vertex_filter<F>, dim3(blocks_for(fr->n, 256, 2048)), dim3(256), 0, lib_stream(), filter,
                       (const bool *)fr->b, (int64_t)0, (int64_t)fr->n, V->getOffset(nodeNum), result->b,
                       totals().slots);
    launched("vertexFilter");
    finish_totals();
    unsigned long long tot = 0;
    read_totals(&tot, nullptr);
    result->m += (intT)tot;
    result->isDense = true;
}

// switchFrontier (polymer.h:2224-2236): next becomes the current frontier; the
// old current one is returned in next, cleared
inline void switchFrontier(int nodeNum, vertices *V, LocalFrontier *&next) {
    LocalFrontier *old = V->swapFrontier(nodeNum, next);
    next = old;
    next->clearFrontier();
}

// Waits for the queued work (every call above already does).
inline void polymerSync() { polymer_detail::sync(); }

// Functor classes written between these two lines compile for the host and the
// device without annotations when the program is not compiled with --hipstdpar
// (the program's host code, main() included, stays outside: a host+device main()
// would be compiled for the GPU as well).
#define POLYMER_FUNCTORS_BEGIN _Pragma("clang force_cuda_host_device begin")
#define POLYMER_FUNCTORS_END _Pragma("clang force_cuda_host_device end")

#endif /* POLYMER_HIP_H */
