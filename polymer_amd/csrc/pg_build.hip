// pg_build.hip -- library state, graph ingest and the HBM layout.
//
// Replaces the reference's NUMA layer (polymer.h:66-519: partitionByDegree,
// graphFilter2Direction, graph*Hasher, mapDataArray) and its in-edge build
// (IO-numa.h:145-183, a quickSort of m (dst,src) pairs) with one device
// resident CSR + CSC:
//   * vertices relabelled by in-degree, descending (ties by input id), so a
//     pull over the CSC visits vertices in non-increasing degree order and
//     the degree bins of the pull schedule are contiguous id ranges;
//   * both directions sorted by 64-bit (vertex, neighbour) keys with a device
//     radix sort; the payload keeps each edge's index in the input CSR so that
//     edge data (weights, BP potentials, gradients) can be given and returned
//     in the caller's order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <fstream>
#include <mutex>
#include <string>
#include <thread>

#include "pg_engine.h"
#include "pg_internal.h"
#include "pg_kernels_common.h"

namespace pg {

static thread_local char g_err[1024] = "";
static hipStream_t g_stream = nullptr;
static hipStream_t g_side = nullptr;  // work overlapped with the library stream's
static int g_device = -1;
static bool g_init = false;

// Device memory of pg_subset_to_sparse_noalloc: a file-scope array, so the entry point
// allocates nothing.  Its host-side name is not a device address: the address is taken
// with hipGetSymbolAddress at pg_init, once per device (each device holds its own copy),
// and only that pointer is used.  Layout: the input-order bitmap (kNoAllocMaxN / 32 + 1
// words), then per-block counts and their offsets (kNoAllocBlocks words each).  About
// 4.03 MiB per device, held until the process ends.
constexpr int64_t kNoAllocMaxN = (int64_t)1 << 25;
constexpr int64_t kNoAllocWords = kNoAllocMaxN / 32 + 1;
constexpr int64_t kNoAllocBlocks = (kNoAllocWords + 255) / 256 + 1;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords + 2 * kNoAllocBlocks];
constexpr int kMaxDevices = 64;
static uint32_t *g_scratch[kMaxDevices] = {};

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

hipStream_t stream() { return g_stream; }
hipStream_t side_stream() { return g_side; }
bool initialised() { return g_init; }

// ---- temporary device buffers (graph build only) ------------------------------
// Frees wait for the library stream first: kernels enqueued earlier may still
// read the buffer.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) {
            (void)hipStreamSynchronize(g_stream);
            (void)hipFree(p);
        }
    }
    int alloc(size_t bytes) {
        if (p) {
            PG_CHECK_HIP(hipStreamSynchronize(g_stream));
            (void)hipFree(p);
            p = nullptr;
        }
        PG_CHECK_HIP(hipMalloc(&p, bytes ? bytes : 16));
        return PG_OK;
    }
    template <class T>
    T *as() const {
        return reinterpret_cast<T *>(p);
    }
};

int galloc(pg_graph *g, void **p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return e == hipErrorOutOfMemory ? PG_ERR_OOM : PG_ERR_HIP;
    }
    g->owned.push_back(*p);
    return PG_OK;
}

// Captured solves keep raw device pointers: any buffer they could reference that
// moves (a workspace grows, a schedule is rebuilt) drops them; they are captured
// again on next use.
void drop_captures(pg_graph *g) {
    if (g->graph_exec.empty()) return;
    (void)hipStreamSynchronize(g_stream);
    for (auto &kv : g->graph_exec) (void)hipGraphExecDestroy(kv.second.first);
    g->graph_exec.clear();
}

int workspace(pg_graph *g, const char *name, size_t bytes, void **p) {
    auto it = g->ws.find(name);
    if (it != g->ws.end() && it->second.second >= bytes) {
        *p = it->second.first;
        return PG_OK;
    }
    if (it != g->ws.end()) {
        drop_captures(g);
        PG_CHECK_HIP(hipStreamSynchronize(g_stream));
        PG_CHECK_HIP(hipFree(it->second.first));
        g->ws.erase(it);
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 16);
    if (e != hipSuccess) {
        set_error("workspace %s: hipMalloc(%zu) failed: %s", name, bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? PG_ERR_OOM : PG_ERR_HIP;
    }
    g->ws[name] = {q, bytes};
    *p = q;
    return PG_OK;
}

// ---- small reads of device words by the host -------------------------------------
// A copy back into pageable memory with a stream synchronisation measured 22 us;
// here one block copies the words into pinned host memory and then writes a
// sequence number there, and the host spins on that number (a launch and the
// spin measured 6 us).  Everything queued before the copy has finished once the
// number is visible.  After 2 ms the host falls back to the synchronisation,
// which reports a kernel that failed.
constexpr int kFetchWords = 1024;  // 32-bit words per fetch
struct FetchBuf {
    uint32_t *cpu = nullptr;  // kFetchWords words, then the 64-bit sequence number
    uint32_t *dev = nullptr;  // the same memory as the kernels address it
    unsigned long long seq = 0;
};
static FetchBuf g_fetch;

__global__ void k_fetch(const uint32_t *src, int nw, uint32_t *dst, unsigned long long seq) {
    for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        *reinterpret_cast<volatile unsigned long long *>(dst + kFetchWords) = seq;
        __threadfence_system();
    }
}

int fetch_words(void *host, const void *dev, size_t bytes) {
    PG_REQUIRE(bytes % 4 == 0 && bytes <= 4 * (size_t)kFetchWords, PG_ERR_ARG,
               "fetch_words: %zu bytes", bytes);
    static std::mutex mu;  // one pinned buffer: calls from several host threads take turns
    std::lock_guard<std::mutex> lock(mu);
    if (!g_fetch.cpu) {
        PG_CHECK_HIP(hipHostMalloc((void **)&g_fetch.cpu, 4 * kFetchWords + 8, hipHostMallocDefault));
        PG_CHECK_HIP(hipHostGetDevicePointer((void **)&g_fetch.dev, g_fetch.cpu, 0));
        *reinterpret_cast<unsigned long long *>(g_fetch.cpu + kFetchWords) = 0ull;
    }
    const unsigned long long seq = ++g_fetch.seq;
    hipLaunchKernelGGL(k_fetch, dim3(1), dim3(256), 0, g_stream, static_cast<const uint32_t *>(dev),
                       (int)(bytes / 4), g_fetch.dev, seq);
    PG_CHECK_HIP(hipGetLastError());
    volatile unsigned long long *h =
        reinterpret_cast<volatile unsigned long long *>(g_fetch.cpu + kFetchWords);
    const auto t0 = std::chrono::steady_clock::now();
    while (*h != seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
            PG_CHECK_HIP(hipStreamSynchronize(g_stream));
            PG_REQUIRE(*h == seq, PG_ERR_HIP, "fetch_words: the copy did not arrive");
            break;
        }
    }
    std::memcpy(host, const_cast<const uint32_t *>(g_fetch.cpu), bytes);
    return PG_OK;
}

int Timer::start() {
    if (!a) PG_CHECK_HIP(hipEventCreate(&a));
    if (!b) PG_CHECK_HIP(hipEventCreate(&b));
    PG_CHECK_HIP(hipEventRecord(a, g_stream));
    return PG_OK;
}
int Timer::stop(double *ms) {
    PG_CHECK_HIP(hipEventRecord(b, g_stream));
    PG_CHECK_HIP(hipEventSynchronize(b));
    float f = 0.f;
    PG_CHECK_HIP(hipEventElapsedTime(&f, a, b));
    *ms = f;
    return PG_OK;
}
Timer::~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
}

int KernelTimer::begin() {
    if (!on) return PG_OK;
    hipEvent_t e;
    PG_CHECK_HIP(hipEventCreate(&e));
    ev.push_back(e);
    PG_CHECK_HIP(hipEventRecord(e, g_stream));
    return PG_OK;
}
int KernelTimer::end() { return begin(); }
int KernelTimer::total(double *ms, int *launches) {
    *ms = 0.0;
    *launches = 0;
    if (!on || ev.size() < 2) return PG_OK;
    PG_CHECK_HIP(hipEventSynchronize(ev.back()));
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
        float f = 0.f;
        PG_CHECK_HIP(hipEventElapsedTime(&f, ev[i], ev[i + 1]));
        *ms += f;
        (*launches)++;
    }
    return PG_OK;
}
KernelTimer::~KernelTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
}

// ---- hipcub wrappers ----------------------------------------------------------
static int sort_pairs(uint64_t *kin, uint64_t *kout, uint32_t *vin, uint32_t *vout, int64_t n,
                      int end_bit) {
    size_t tb = 0;
    PG_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, n, 0,
                                                    end_bit, g_stream));
    DevBuf tmp;
    PG_TRY(tmp.alloc(tb));
    PG_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kin, kout, vin, vout, n, 0,
                                                    end_bit, g_stream));
    return PG_OK;
}

static int sort_keys(uint64_t *kin, uint64_t *kout, int64_t n, int end_bit) {
    size_t tb = 0;
    PG_CHECK_HIP(
        hipcub::DeviceRadixSort::SortKeys(nullptr, tb, kin, kout, n, 0, end_bit, g_stream));
    DevBuf tmp;
    PG_TRY(tmp.alloc(tb));
    PG_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, kin, kout, n, 0, end_bit, g_stream));
    return PG_OK;
}

template <class TI, class TO>
static int exclusive_sum(const TI *in, TO *out, int64_t n) {
    size_t tb = 0;
    PG_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, g_stream));
    DevBuf tmp;
    PG_TRY(tmp.alloc(tb));
    PG_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, in, out, n, g_stream));
    return PG_OK;
}

template <class TI, class TO>
static int inclusive_sum(const TI *in, TO *out, int64_t n) {
    size_t tb = 0;
    PG_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out, n, g_stream));
    DevBuf tmp;
    PG_TRY(tmp.alloc(tb));
    PG_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, in, out, n, g_stream));
    return PG_OK;
}

static int select_flagged(const uint64_t *in, const uint8_t *flags, uint64_t *out,
                          unsigned long long *d_nsel, int64_t n) {
    size_t tb = 0;
    PG_CHECK_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, in, flags, out, d_nsel, n, g_stream));
    DevBuf tmp;
    PG_TRY(tmp.alloc(tb));
    PG_CHECK_HIP(hipcub::DeviceSelect::Flagged(tmp.p, tb, in, flags, out, d_nsel, n, g_stream));
    return PG_OK;
}

// ---- exclusive scans of 32-bit counts (frontier compaction, push degree scans) ----------
// Wave scans by shuffles, 16 waves of a 1024-thread workgroup joined in LDS.  A scan of
// n items is one workgroup when n fits a few tiles, otherwise three launches: tile sums,
// one workgroup over the tile sums, tiles rescanned with their offsets (2 reads and one
// write of the items; no temporary storage beyond one word per 4096 items).
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    return v;
}

// exclusive scan of one value per thread over the workgroup; *total = the sum.  sm holds
// kScanThreads / 64 + 1 words; the call ends with a barrier, so it can be repeated.
__device__ __forceinline__ uint32_t workgroup_exclusive_scan(uint32_t v, uint32_t *sm,
                                                             uint32_t *total) {
    constexpr int W = kScanThreads / 64;
    const uint32_t inc = wave_inclusive_scan(v);
    const int w = threadIdx.x >> 6, lane = lane_id();
    if (lane == 63) sm[w] = inc;
    __syncthreads();
    if (w == 0) {
        const uint32_t x = lane < W ? sm[lane] : 0u;
        const uint32_t xi = wave_inclusive_scan(x);
        if (lane < W) sm[lane] = xi - x;
        if (lane == W - 1) sm[W] = xi;
    }
    __syncthreads();
    const uint32_t r = sm[w] + inc - v;
    *total = sm[W];
    __syncthreads();
    return r;
}

// out[i] = in[0] + ... + in[i-1] for i < n, by one workgroup walking tiles with a carry
// (in and out may be the same array); out[n] = the total when with_total
__global__ __launch_bounds__(kScanThreads) void k_scan_one_group(const uint32_t *in, uint32_t *out,
                                                                 int64_t n, int with_total) {
    __shared__ uint32_t sm[kScanThreads / 64 + 1];
    uint32_t carry = 0;
    for (int64_t base = 0; base < n; base += kScanTile) {
        const int64_t i0 = base + (int64_t)threadIdx.x * kScanItems;
        uint32_t v[kScanItems], s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            v[k] = i0 + k < n ? in[i0 + k] : 0u;
            s += v[k];
        }
        uint32_t tot = 0;
        uint32_t run = carry + workgroup_exclusive_scan(s, sm, &tot);
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            if (i0 + k < n) out[i0 + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    if (with_total && threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const uint32_t *in, int64_t n,
                                                                 uint32_t *sums) {
    __shared__ uint32_t sm[kScanThreads / 64 + 1];
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) s += i0 + k < n ? in[i0 + k] : 0u;
    uint32_t tot = 0;
    (void)workgroup_exclusive_scan(s, sm, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint32_t *in, int64_t n,
                                                             const uint32_t *offs, uint32_t *out) {
    __shared__ uint32_t sm[kScanThreads / 64 + 1];
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = i0 + k < n ? in[i0 + k] : 0u;
        s += v[k];
    }
    uint32_t tot = 0;
    uint32_t run = offs[blockIdx.x] + workgroup_exclusive_scan(s, sm, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
}

// out[i] = in[0] + ... + in[i-1], i < n, on the library stream; tile sums (one word per
// 4096 items) in the caller's buffer tmp (scan_tmp_words(n) words), or in a graph
// workspace when tmp is null
int64_t scan_tmp_words(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

int exclusive_scan_u32_on(hipStream_t st, const uint32_t *in, uint32_t *out, int64_t n,
                          uint32_t *tmp) {
    if (n <= 0) return PG_OK;
    if (n <= 4 * (int64_t)kScanTile) {
        hipLaunchKernelGGL(k_scan_one_group, dim3(1), dim3(kScanThreads), 0, st, in, out, n, 0);
    } else {
        const int64_t nt = (n + kScanTile - 1) / kScanTile;
        PG_REQUIRE(nt < 0x7FFFFFFF, PG_ERR_UNSUPPORTED, "scan: too many items");
        hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nt), dim3(kScanThreads), 0, st, in, n,
                           tmp);
        hipLaunchKernelGGL(k_scan_one_group, dim3(1), dim3(kScanThreads), 0, st, tmp, tmp, nt, 0);
        hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(kScanThreads), 0, st, in, n, tmp,
                           out);
    }
    PG_CHECK_HIP(hipGetLastError());
    return PG_OK;
}

// Scans inside traversal rounds keep their tile sums in a graph workspace sized for the
// longest scan of the graph (n + 1 items): no allocation after the first call.
int exclusive_scan_u32(pg_graph *g, const uint32_t *in, uint32_t *out, int64_t n) {
    uint32_t *tmp = nullptr;
    PG_TRY(workspace_t(g, "scan_tmp", (size_t)scan_tmp_words(std::max<int64_t>(n, g->n + 1)),
                       &tmp));
    return exclusive_scan_u32_on(g_stream, in, out, n, tmp);
}

static int bits_for(int64_t n) {
    int b = 1;
    while ((int64_t(1) << b) < n) b++;
    return b;
}

#define PG_LAUNCH(kern, grid, ...)                                                     \
    do {                                                                               \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, g_stream, __VA_ARGS__); \
        PG_CHECK_HIP(hipGetLastError());                                               \
    } while (0)

// ---- build kernels -------------------------------------------------------------
__global__ void k_rmat_keys(uint64_t base, int scale, uint32_t t0, uint32_t t1, uint32_t t2,
                            int64_t count, uint64_t *keys, uint32_t *ctr) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (int64_t)gridDim.x * blockDim.x) {
        uint32_t s = 0, d = 0;
        for (int i = 0; i < scale; i++) {
            const uint32_t r = (uint32_t)(dmix64(base + (uint64_t)k * 32ull + (uint64_t)i) >> 32);
            const uint32_t q = r < t0 ? 0u : (r < t1 ? 1u : (r < t2 ? 2u : 3u));
            s = (s << 1) | (q >> 1);
            d = (d << 1) | (q & 1u);
        }
        keys[k] = (s == d) ? ~0ull : (((uint64_t)s << 32) | d);
        ctr[k] = (uint32_t)k;
    }
}

__global__ void k_first_by_ctr(const uint64_t *keys, const uint32_t *ctr, int64_t count,
                               uint32_t *mark) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        if (k != ~0ull && (i == 0 || keys[i - 1] != k)) mark[ctr[i]] = 1;
    }
}

__global__ void k_find_cutoff(const uint32_t *cnt, int64_t count, uint32_t target,
                              int64_t *cutoff) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (int64_t)gridDim.x * blockDim.x) {
        if (cnt[k] == target && (k == 0 || cnt[k - 1] + 1 == target)) *cutoff = k;
    }
}

__global__ void k_select_flags(const uint64_t *keys, const uint32_t *ctr, int64_t count,
                               const int64_t *cutoff, uint8_t *sel) {
    const int64_t cut = *cutoff;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        sel[i] = (k != ~0ull && (i == 0 || keys[i - 1] != k) && (int64_t)ctr[i] <= cut) ? 1 : 0;
    }
}

__global__ void k_unique_flags(const uint64_t *keys, int64_t count, uint8_t *sel) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        const uint32_t s = (uint32_t)(k >> 32), d = (uint32_t)k;
        sel[i] = (s != d && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
    }
}

__global__ void k_mirror(const uint64_t *in, int64_t m, uint64_t *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = in[i];
        out[i] = k;
        out[m + i] = (k << 32) | (k >> 32);
    }
}

__global__ void k_keys_from_csr(const int32_t *off, const int32_t *adj, int64_t n,
                                uint64_t *keys) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        for (int e = off[v]; e < off[v + 1]; e++) keys[e] = ((uint64_t)v << 32) | (uint32_t)adj[e];
    }
}

__global__ void k_degrees(const uint64_t *keys, int64_t m, uint32_t *outdeg, uint32_t *indeg) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        atomicAdd(&outdeg[k >> 32], 1u);
        atomicAdd(&indeg[(uint32_t)k], 1u);
    }
}

__global__ void k_vertex_keys(const uint32_t *deg, int64_t n, uint64_t *vk) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        vk[v] = ((uint64_t)(0xFFFFFFFFu - deg[v]) << 32) | (uint32_t)v;
}

__global__ void k_perm_from_sorted(const uint64_t *vk, int64_t n, int32_t *iperm, int32_t *perm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = (int32_t)(uint32_t)vk[i];
        iperm[i] = v;
        perm[v] = (int32_t)i;
    }
}

// relabelled edge keys; dir 0: (src,dst), dir 1: (dst,src)
__global__ void k_relabel_keys(const uint64_t *keys, int64_t m, const int32_t *perm, int dir,
                               uint64_t *out, uint32_t *eidx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        const uint32_t s = (uint32_t)perm[k >> 32], d = (uint32_t)perm[(uint32_t)k];
        out[i] = dir == 0 ? (((uint64_t)s << 32) | d) : (((uint64_t)d << 32) | s);
        eidx[i] = (uint32_t)i;
    }
}

__global__ void k_permute_counts(const uint32_t *deg, const int32_t *perm, int64_t n,
                                 uint32_t *out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n;
         v += (int64_t)gridDim.x * blockDim.x)
        out[v < n ? perm[v] : n] = v < n ? deg[v] : 0u;
}

__global__ void k_low32(const uint64_t *keys, int64_t m, int32_t *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (int32_t)(uint32_t)keys[i];
}

template <class T>
__global__ void k_gather(const T *src, const int32_t *idx, int64_t m, T *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

__global__ void k_pair_weights_i32(const uint64_t *keys, int64_t m, uint64_t seed, int32_t *w) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        w[i] = dweight_i32(seed, (int32_t)(k >> 32), (int32_t)(uint32_t)k);
    }
}

__global__ void k_pair_values_f32(const uint64_t *keys, int64_t m, uint64_t seed, float *w) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        w[i] = dvalue_f32(seed, (int32_t)(k >> 32), (int32_t)(uint32_t)k);
    }
}

__global__ void k_count_above(const int32_t *off, const int32_t *order, int64_t n,
                              const int64_t *thr, int nthr, int64_t *out) {
    // binary search per threshold over non-increasing degrees
    const int j = threadIdx.x;
    if (j >= nthr) return;
    const int64_t T = thr[j];
    int64_t lo = 0, hi = n;  // first position with deg <= T
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int v = order ? order[mid] : (int)mid;
        const int64_t d = off[v + 1] - off[v];
        if (d > T) lo = mid + 1; else hi = mid;
    }
    out[j] = lo;
}

__global__ void k_degree_of(const int32_t *off, const int32_t *order, int64_t count, int32_t *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int v = order ? order[i] : (int)i;
        out[i] = off[v + 1] - off[v];
    }
}

__global__ void k_out_order_keys(const int32_t *off, int64_t n, uint64_t *vk) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        vk[v] = ((uint64_t)(0xFFFFFFFFu - (uint32_t)(off[v + 1] - off[v])) << 32) | (uint32_t)v;
}

__global__ void k_bp_pot(const int32_t *in_off, const int32_t *in_adj, const int32_t *iperm,
                         int64_t n, uint64_t seed, float lo, float hi, float *pot) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t dv = iperm[v];
        for (int e = in_off[v]; e < in_off[v + 1]; e++) {
            const int32_t sv = iperm[in_adj[e]];
#pragma unroll
            for (int k = 0; k < 4; k++) pot[4ll * e + k] = dbp_pot(seed, sv, dv, k, lo, hi);
        }
    }
}

// input-order transfers
__global__ void k_to_input_i32(const int32_t *a, const int32_t *perm, const int32_t *iperm,
                               int64_t n, int map_ids, int32_t *out) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x) {
        int32_t x = a[perm[u]];
        if (map_ids && x >= 0) x = iperm[x];
        out[u] = x;
    }
}

__global__ void k_to_input_f32(const float *a, const int32_t *perm, int64_t n, int width,
                               float *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * width;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = i / width, c = i % width;
        out[i] = a[(int64_t)perm[u] * width + c];
    }
}

__global__ void k_from_input_f32(const float *in, const int32_t *perm, int64_t n, int width,
                                 float *a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * width;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = i / width, c = i % width;
        a[(int64_t)perm[u] * width + c] = in[i];
    }
}

// ---- frontier conversions (bitmap <-> ordered list) -------------------------------
// One 32-bit word per thread: popcounts scanned in LDS, block totals scanned
// on the device, then every thread writes its ids in order -- an ordered
// packIndex (polymer.h:552-567, utils.h:207-227).
__global__ void k_word_counts(const uint32_t *bits, int64_t nwords, uint32_t *block_cnt) {
    __shared__ uint32_t sm[kBlock / 64];
    const int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t c = w < nwords ? (uint32_t)__popc(bits[w]) : 0u;
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane_id() == 0) sm[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kBlock / 64; i++) t += sm[i];
        block_cnt[blockIdx.x] = t;
    }
}

__global__ void k_word_write(const uint32_t *bits, int64_t nwords, int64_t n,
                             const uint32_t *block_off, int32_t *list,
                             unsigned long long *count_out, int64_t nblocks) {
    __shared__ uint32_t s_scan[kBlock + 1];
    const int t = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * kBlock + t;
    const uint32_t word = w < nwords ? bits[w] : 0u;
    s_scan[t + 1] = (uint32_t)__popc(word);
    if (t == 0) s_scan[0] = 0;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        uint32_t x = (t + 1 - o >= 1) ? s_scan[t + 1 - o] : 0u;
        __syncthreads();
        s_scan[t + 1] += x;
        __syncthreads();
    }
    uint32_t pos = block_off[blockIdx.x] + s_scan[t];
    uint32_t x = word;
    while (x) {
        const int b = __ffs(x) - 1;
        x &= x - 1;
        const int64_t v = w * 32 + b;
        if (v < n) list[pos++] = (int32_t)v;
    }
    if (blockIdx.x == nblocks - 1 && t == kBlock - 1) *count_out = block_off[blockIdx.x] + s_scan[kBlock];
}

__global__ void k_list_to_bits(const int32_t *list, const unsigned long long *count, uint32_t *bits) {
    const int64_t c = (int64_t)*count;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int v = list[i];
        atomicOr(&bits[v >> 5], 1u << (v & 31));
    }
}

// out[stride * (block mod 64)] += the block's sum: a few thousand blocks adding to
// one word serialise (tens of microseconds on MI355X); spread == false adds to out[0]
__global__ void k_list_degree_sum(const int32_t *list, const unsigned long long *count,
                                  const int32_t *off, unsigned long long *out, int stride) {
    __shared__ unsigned long long sm[kBlock / 64];
    const int64_t c = (int64_t)*count;
    unsigned long long acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int v = list[i];
        acc += (unsigned long long)(off[v + 1] - off[v]);
    }
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane_id() == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kBlock / 64; i++) t += sm[i];
        if (t) atomicAdd(&out[(int64_t)stride * (blockIdx.x & 63)], t);
    }
}

__global__ void k_bits_from_bools_input(const uint8_t *flags, const int32_t *perm, int64_t n,
                                        uint32_t *bits) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x)
        if (flags[u]) {
            const int v = perm[u];
            atomicOr(&bits[v >> 5], 1u << (v & 31));
        }
}

__global__ void k_bits_from_ids_input(const int32_t *ids, int64_t k, const int32_t *perm,
                                      uint32_t *bits) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int v = perm[ids[i]];
        atomicOr(&bits[v >> 5], 1u << (v & 31));
    }
}

__global__ void k_bits_to_input_bits(const uint32_t *bits, const int32_t *perm, int64_t n,
                                     uint32_t *ibits) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x)
        if (bit_test(bits, perm[u])) atomicOr(&ibits[u >> 5], 1u << (u & 31));
}

__global__ void k_fill_bits_all(uint32_t *bits, int64_t n) {
    const int64_t nw = (n + 31) / 32;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
         w += (int64_t)gridDim.x * blockDim.x) {
        const int64_t rem = n - w * 32;
        bits[w] = rem >= 32 ? 0xFFFFFFFFu : ((1u << rem) - 1u);
    }
}

int bits_to_list(pg_graph *g, const uint32_t *bits, int32_t *list, unsigned long long *d_count) {
    const int64_t nwords = (g->n + 31) / 32;
    const int64_t nb = (nwords + kBlock - 1) / kBlock;
    uint32_t *bc = nullptr, *bo = nullptr;
    PG_TRY(workspace_t(g, "compact_bc", (size_t)nb + 1, &bc));
    PG_TRY(workspace_t(g, "compact_bo", (size_t)nb + 1, &bo));
    PG_LAUNCH(k_word_counts, (unsigned)nb, bits, nwords, bc);
    hipLaunchKernelGGL(k_scan_one_group, dim3(1), dim3(kScanThreads), 0, g_stream, bc, bo, nb, 0);
    PG_LAUNCH(k_word_write, (unsigned)nb, bits, nwords, g->n, bo, list, d_count, nb);
    return PG_OK;
}

int list_to_bits(pg_graph *g, const int32_t *list, const unsigned long long *d_count,
                 uint32_t *bits) {
    PG_LAUNCH(k_list_to_bits, grid_for(g->n, 8192), list, d_count, bits);
    return PG_OK;
}

int list_degree_sum(pg_graph *g, const int32_t *list, const unsigned long long *d_count,
                    const int32_t *off, unsigned long long *d_out, int stride) {
    PG_LAUNCH(k_list_degree_sum, grid_for(g->n, 4096), list, d_count, off, d_out, stride);
    return PG_OK;
}

int fill_bits_all(pg_graph *g, uint32_t *bits) {
    PG_LAUNCH(k_fill_bits_all, grid_for((g->n + 31) / 32), bits, g->n);
    return PG_OK;
}

// ---- input-order transfers ---------------------------------------------------------
int to_input_order_i32(pg_graph *g, const int32_t *dev_internal, int32_t *host, bool map_ids) {
    if (!host) return PG_OK;
    int32_t *tmp = nullptr;
    PG_TRY(workspace_t(g, "xfer", (size_t)g->n * 4, reinterpret_cast<float **>(&tmp)));
    PG_LAUNCH(k_to_input_i32, grid_for(g->n), dev_internal, g->perm, g->iperm, g->n,
              map_ids ? 1 : 0, tmp);
    PG_CHECK_HIP(hipMemcpyAsync(host, tmp, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

int to_input_order_f32(pg_graph *g, const float *dev_internal, float *host, int width) {
    if (!host) return PG_OK;
    float *tmp = nullptr;
    PG_TRY(workspace_t(g, "xfer", (size_t)g->n * width, &tmp));
    PG_LAUNCH(k_to_input_f32, grid_for(g->n * width), dev_internal, g->perm, g->n, width, tmp);
    PG_CHECK_HIP(hipMemcpyAsync(host, tmp, sizeof(float) * g->n * width, hipMemcpyDeviceToHost,
                                g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

int from_input_order_f32(pg_graph *g, const float *host, float *dev_internal, int width) {
    float *tmp = nullptr;
    PG_TRY(workspace_t(g, "xfer", (size_t)g->n * width, &tmp));
    PG_CHECK_HIP(hipMemcpyAsync(tmp, host, sizeof(float) * g->n * width, hipMemcpyHostToDevice,
                                g_stream));
    PG_LAUNCH(k_from_input_f32, grid_for(g->n * width), tmp, g->perm, g->n, width, dev_internal);
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

int edges_from_input_f32(pg_graph *g, const float *host, bool csc, float *dev) {
    float *tmp = nullptr;
    PG_TRY(workspace_t(g, "xfer_e", (size_t)g->m, &tmp));
    PG_CHECK_HIP(hipMemcpyAsync(tmp, host, sizeof(float) * g->m, hipMemcpyHostToDevice, g_stream));
    PG_LAUNCH(k_gather<float>, grid_for(g->m), tmp, csc ? g->in_eid : g->out_eid, g->m, dev);
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

__global__ void k_scatter_f32(const float *src, const int32_t *idx, int64_t m, float *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[idx[i]] = src[i];
}

int edges_to_input_f32(pg_graph *g, const float *dev, bool csc, float *host) {
    if (!host) return PG_OK;
    float *tmp = nullptr;
    PG_TRY(workspace_t(g, "xfer_e", (size_t)g->m, &tmp));
    PG_LAUNCH(k_scatter_f32, grid_for(g->m), dev, csc ? g->in_eid : g->out_eid, g->m, tmp);
    PG_CHECK_HIP(hipMemcpyAsync(host, tmp, sizeof(float) * g->m, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

// ---- schedule ---------------------------------------------------------------------
static int free_schedule(PullSchedule &s) {
    if (s.d_multi) (void)hipFree(s.d_multi);
    if (s.d_partial) (void)hipFree(s.d_partial);
    if (s.d_tickets) (void)hipFree(s.d_tickets);
    s.d_multi = nullptr;
    s.d_partial = nullptr;
    s.d_tickets = nullptr;
    s.built = false;
    s.nblocks = 0;
    s.nblocks_body = 0;
    s.body_end = 0;
    return PG_OK;
}

int build_schedule(pg_graph *g, bool in_direction) {
    PullSchedule &s = in_direction ? g->in_sched : g->out_sched;
    drop_captures(g);
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    free_schedule(s);
    if (in_direction) g->bp_light_built = false;
    const int32_t *off = in_direction ? g->in_off : g->out_off;
    const int32_t *adj = in_direction ? g->in_adj : g->out_adj;
    const int32_t *order = in_direction ? nullptr : g->out_order;
    const PullParams &P = s.params;
    const int64_t n = g->n;
    const int ipl = std::max(1, P.items_per_lane);
    // class upper bounds, descending: multi | block | 64 | 32 | 16 | 8 | 4 | 2 | 1
    const int gs[9] = {kModeMulti, kModeBlock, 64, 32, 16, 8, 4, 2, 1};
    int64_t thr[10];
    thr[0] = std::max<int64_t>(P.multi_threshold, P.block_threshold);
    thr[1] = std::max<int64_t>(P.block_threshold, 32LL * ipl);
    for (int j = 2; j < 8; j++) thr[j] = (int64_t)ipl << (7 - j);  // 32ipl,16,8,4,2,ipl
    // thr[2] = 32*ipl ... thr[7] = ipl; thr[8] = 0 counts the vertices with edges;
    // thr[9]: the in-degree above which a vertex is not put in a tile
    thr[8] = 0;
    thr[9] = std::max<int64_t>(0, g->tile_max_degree);
    DevBuf dthr, dcnt;
    PG_TRY(dthr.alloc(sizeof(thr)));
    PG_TRY(dcnt.alloc(sizeof(thr)));
    PG_CHECK_HIP(hipMemcpyAsync(dthr.p, thr, sizeof(thr), hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(k_count_above, dim3(1), dim3(64), 0, g_stream, off, order, n,
                       dthr.as<int64_t>(), 10, dcnt.as<int64_t>());
    PG_CHECK_HIP(hipGetLastError());
    int64_t cnt[10];
    PG_CHECK_HIP(hipMemcpyAsync(cnt, dcnt.p, sizeof(cnt), hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    int64_t bounds[10];
    bounds[0] = 0;
    for (int j = 0; j < 8; j++) bounds[j + 1] = std::max(cnt[j], bounds[j]);
    bounds[9] = n;
    PullArgs &a = s.args;
    a = PullArgs{};
    a.off = off;
    a.adj = adj;
    a.vlist = order;
    a.chunk = (int32_t)std::max<int64_t>(kBlock, P.chunk);
    int nseg = 0;
    int64_t blocks = 0;
    std::vector<MultiItem> items;
    for (int c = 0; c < 9; c++) {
        const int64_t vb = bounds[c], ve = bounds[c + 1];
        if (ve <= vb) continue;
        PullSeg seg{(int32_t)vb, (int32_t)ve, gs[c], (int32_t)blocks};
        if (gs[c] == kModeMulti) {
            std::vector<int32_t> deg((size_t)(ve - vb));
            DevBuf dd;
            PG_TRY(dd.alloc(sizeof(int32_t) * deg.size()));
            hipLaunchKernelGGL(k_degree_of, dim3(grid_for(ve)), dim3(kBlock), 0, g_stream, off,
                               order, ve, dd.as<int32_t>());
            PG_CHECK_HIP(hipGetLastError());
            PG_CHECK_HIP(hipMemcpyAsync(deg.data(), dd.p, sizeof(int32_t) * deg.size(),
                                        hipMemcpyDeviceToHost, g_stream));
            PG_CHECK_HIP(hipStreamSynchronize(g_stream));
            for (int64_t p = vb; p < ve; p++) {
                const int32_t nch = (int32_t)((deg[p - vb] + a.chunk - 1) / a.chunk);
                const int32_t first = (int32_t)items.size();
                for (int32_t ch = 0; ch < nch; ch++)
                    items.push_back(MultiItem{(int32_t)p, ch, nch, first, (int32_t)(p - vb), {0, 0, 0}});
            }
            blocks += (int64_t)items.size();
        } else if (gs[c] == kModeBlock) {
            blocks += ve - vb;
        } else {
            blocks += ((ve - vb) * gs[c] + kBlock - 1) / kBlock;
        }
        if (nseg >= kMaxSegs) {
            set_error("schedule: too many segments");
            return PG_ERR_STATE;
        }
        a.seg[nseg++] = seg;
    }
    a.nseg = nseg;
    if (!items.empty()) {
        PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&s.d_multi), sizeof(MultiItem) * items.size()));
        PG_CHECK_HIP(hipMemcpy(s.d_multi, items.data(), sizeof(MultiItem) * items.size(),
                               hipMemcpyHostToDevice));
        PG_CHECK_HIP(hipMalloc(&s.d_partial, 32 * items.size() + 64));
        const int64_t nmv = bounds[1] - bounds[0];
        PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&s.d_tickets), sizeof(unsigned) * (nmv + 1)));
        // on the library stream, ahead of the pulls that use them
        PG_CHECK_HIP(hipMemsetAsync(s.d_tickets, 0, sizeof(unsigned) * (nmv + 1), g_stream));
    }
    a.multi = s.d_multi;
    a.partial = s.d_partial;
    a.tickets = s.d_tickets;
    if (blocks > 0x7FFFFFFF) {
        set_error("schedule: grid too large");
        return PG_ERR_UNSUPPORTED;
    }
    s.nblocks = (int)blocks;
    // the zero-degree tail ends the last segment: a pull whose vertices without
    // edges have nothing to do launches only the blocks before it
    {
        const int64_t body = std::max<int64_t>(cnt[8], 0);
        int64_t bb = 0;
        for (int i = 0; i < a.nseg; i++) {
            const PullSeg &sg = a.seg[i];
            if (sg.v_begin >= body) break;
            if (sg.v_end <= body || sg.g == kModeMulti || sg.g == kModeBlock) {
                bb = i + 1 < a.nseg ? a.seg[i + 1].block_begin : blocks;
            } else {
                bb = sg.block_begin + ((body - sg.v_begin) * sg.g + kBlock - 1) / kBlock;
                break;
            }
        }
        s.nblocks_body = (int)bb;
        s.body_end = body;
    }
    // hot schedule for the tiled pull: the segments above clipped to the positions
    // whose in-degree exceeds the tile limit (a clipped multi-block segment keeps
    // its first items, which are in position order), then the vertices without
    // pulled edges
    s.light_begin = std::min(cnt[9], std::max(cnt[8], (int64_t)0));
    s.light_end = std::max(cnt[8], s.light_begin);
    PullArgs &h = s.args_hot;
    h = a;
    h.nseg = 0;
    int64_t hblocks = 0;
    for (int i = 0; i < a.nseg && a.seg[i].v_begin < s.light_begin; i++) {
        PullSeg sg = a.seg[i];
        sg.v_end = (int32_t)std::min<int64_t>(sg.v_end, s.light_begin);
        sg.block_begin = (int32_t)hblocks;
        if (sg.g == kModeMulti) {
            int64_t k = 0;
            while (k < (int64_t)items.size() && items[k].pos < sg.v_end) k++;
            hblocks += k;
        } else if (sg.g == kModeBlock) {
            hblocks += sg.v_end - sg.v_begin;
        } else {
            hblocks += ((int64_t)(sg.v_end - sg.v_begin) * sg.g + kBlock - 1) / kBlock;
        }
        h.seg[h.nseg++] = sg;
    }
    s.nblocks_hot_body = (int)hblocks;
    if (s.light_end < n) {
        h.seg[h.nseg++] = PullSeg{(int32_t)s.light_end, (int32_t)n, 1, (int32_t)hblocks};
        hblocks += (n - s.light_end + kBlock - 1) / kBlock;
    }
    s.nblocks_hot = (int)hblocks;
    s.built = true;
    return PG_OK;
}

// ---- source-sorted tiles ----------------------------------------------------------------
static void free_tile_set(TileSet &t) {
    if (t.d_vbeg) (void)hipFree(t.d_vbeg);
    if (t.d_parts) (void)hipFree(t.d_parts);
    if (t.d_fin) (void)hipFree(t.d_fin);
    if (t.d_src) (void)hipFree(t.d_src);
    if (t.d_dst) (void)hipFree(t.d_dst);
    if (t.d_partial) (void)hipFree(t.d_partial);
    if (t.d_packed) (void)hipFree(t.d_packed);
    if (t.d_gbase) (void)hipFree(t.d_gbase);
    if (t.d_esc) (void)hipFree(t.d_esc);
    if (t.d_cold_parts) (void)hipFree(t.d_cold_parts);
    if (t.d_cold_pos) (void)hipFree(t.d_cold_pos);
    if (t.d_cold_srcl) (void)hipFree(t.d_cold_srcl);
    if (t.d_cold_val) (void)hipFree(t.d_cold_val);
    if (t.d_wt) (void)hipFree(t.d_wt);
    if (t.ev_tiles) (void)hipEventDestroy(t.ev_tiles);
    if (t.ev_cold) (void)hipEventDestroy(t.ev_cold);
    t = TileSet{};
}

void free_tiles(pg_graph *g) {
    free_tile_set(g->tiles);
    free_tile_set(g->tiles_out);
}

// key = tile << 32 | source, value = the edge's index relative to e0, for the
// edges of the positions [p0, p1) (a position's list: adj[off[p], off[p+1])); dloc[index]
// = destination offset within the tile.  A wave per position (a hub's list is not
// one thread's loop).
__global__ void k_tile_keys(const int32_t *off, const int32_t *adj, int64_t p0, int64_t p1,
                            int64_t e0, uint64_t *keys, uint32_t *vals, uint16_t *dloc) {
    const int lane = lane_id();
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t p = p0 + (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6); p < p1; p += nw) {
        const int t = (int)((p - p0) / kTileMaxV);
        const uint16_t local = (uint16_t)((p - p0) % kTileMaxV);
        for (int64_t e = off[p] + lane; e < off[p + 1]; e += 64) {
            keys[e - e0] = ((uint64_t)(uint32_t)t << 32) | (uint32_t)adj[e];
            vals[e - e0] = (uint32_t)(e - e0);
            dloc[e - e0] = local;
        }
    }
}

// tile-order sources and destinations, and the weights (as fp32) of a weighted
// graph: the sorted values are CSC edge indices, so every edge keeps its own
// weight whatever its source or duplicates
__global__ void k_tile_split(const uint64_t *keys, const uint32_t *vals, int64_t m,
                             const uint16_t *dloc, int64_t e0, const float *wf,
                             const int32_t *wi, int32_t *src, uint16_t *dst, float *wt) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t ei = vals[i];
        src[i] = (int32_t)(uint32_t)keys[i];
        dst[i] = dloc[ei];
        if (wt) wt[i] = wf ? wf[e0 + ei] : (float)wi[e0 + ei];
    }
}

// first position in [tb[t], tb[t+1]) of src whose value is >= cold (src sorted there)
__global__ void k_tile_cold_start(const int32_t *src, const int64_t *tb, int nt, int32_t cold,
                                  int64_t *out) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        int64_t lo = tb[t], hi = tb[t + 1];
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (src[mid] < cold) lo = mid + 1; else hi = mid;
        }
        out[t] = lo;
    }
}

// key = chunk << 32 | cold value index, value = source - chunk's first source,
// for every cold tile edge (tile found by binary search over tile edge starts)
__global__ void k_cold_keys(const int32_t *src, int64_t mt, const int64_t *tb, int nt,
                            const int64_t *cstart, const int64_t *cbase, int32_t cold,
                            uint64_t *keys, uint32_t *vals) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mt;
         i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nt - 1;  // last tile with tb[t] <= i
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tb[mid] <= i) lo = mid; else hi = mid - 1;
        }
        if (i < cstart[lo]) continue;
        const int64_t k = cbase[lo] + (i - cstart[lo]);
        const uint32_t rel = (uint32_t)(src[i] - cold);
        keys[k] = ((uint64_t)(rel / kColdChunk) << 32) | (uint64_t)k;
        vals[k] = rel % kColdChunk;
    }
}

__global__ void k_cold_split(const uint64_t *keys, const uint32_t *vals, int64_t m, uint32_t *pos,
                             uint16_t *srcl, uint32_t *chunk) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        pos[i] = (uint32_t)keys[i];
        srcl[i] = (uint16_t)vals[i];
        chunk[i] = (uint32_t)(keys[i] >> 32);
    }
}

// off[c] = first i with chunk[i] >= c, c in [0, nc]
__global__ void k_chunk_offsets(const uint32_t *chunk, int64_t m, int64_t nc, int64_t *off) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t hi = i < m ? (int64_t)chunk[i] : nc;
        const int64_t lo = i > 0 ? (int64_t)chunk[i - 1] + 1 : 0;
        for (int64_t c = lo; c <= hi; c++) off[c] = i;
    }
}

// per 64-edge group: min and max of the warm sources (src < cold); empty -> min = max = 0
__global__ void k_group_span(const int32_t *src, int64_t mt, int32_t cold, int32_t *gmin,
                             uint8_t *wide) {
    const int64_t ng = (mt + kPackGroup - 1) / kPackGroup;
    for (int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < ng;
         gi += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0x7FFFFFFF, hi = -1;
        const int64_t e1 = min(mt, (gi + 1) * kPackGroup);
        for (int64_t i = gi * kPackGroup; i < e1; i++) {
            const int32_t s = src[i];
            if (s < cold) {
                lo = min(lo, s);
                hi = max(hi, s);
            }
        }
        if (hi < 0) lo = hi = 0;
        gmin[gi] = lo;
        wide[gi] = (int64_t)hi - lo >= ((int64_t)1 << (32 - kPackDstBits)) ? 1 : 0;
    }
}

__global__ void k_tile_pack(const int32_t *src, const uint16_t *dst, int64_t mt, int32_t cold,
                            const int32_t *gbase, uint32_t *packed, int32_t *esc) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mt;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = src[i], b = gbase[i / kPackGroup];
        uint32_t w = dst[i];
        if (s < cold) {
            if (b >= 0)
                w |= (uint32_t)(s - b) << kPackDstBits;
            else
                esc[(int64_t)(-b - 1) * kPackGroup + (i % kPackGroup)] = s;
        }
        packed[i] = w;
    }
}

// Packs the tile edges (source offset + destination in one word, escape rows for
// wide groups) and frees the separate source / destination arrays.
static int pack_tiles(TileSet &t, int64_t mt, int32_t cold) {
    const int64_t ng = (mt + kPackGroup - 1) / kPackGroup;
    DevBuf dmin, dwide;
    PG_TRY(dmin.alloc(sizeof(int32_t) * (ng + 1)));
    PG_TRY(dwide.alloc(ng + 1));
    PG_LAUNCH(k_group_span, grid_for(ng), t.d_src, mt, cold, dmin.as<int32_t>(), dwide.as<uint8_t>());
    std::vector<int32_t> base((size_t)ng);
    std::vector<uint8_t> wide((size_t)ng);
    PG_CHECK_HIP(hipMemcpyAsync(base.data(), dmin.p, sizeof(int32_t) * ng, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipMemcpyAsync(wide.data(), dwide.p, ng, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    int64_t nesc = 0;
    for (int64_t gi = 0; gi < ng; gi++)
        if (wide[gi]) base[gi] = -(int32_t)(++nesc);
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_gbase), sizeof(int32_t) * (ng + 1)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_esc), sizeof(int32_t) * (nesc * kPackGroup + 1)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_packed), sizeof(uint32_t) * (mt + 16)));
    PG_CHECK_HIP(hipMemcpyAsync(t.d_gbase, base.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice, g_stream));
    PG_LAUNCH(k_tile_pack, grid_for(mt), t.d_src, t.d_dst, mt, cold, t.d_gbase, t.d_packed, t.d_esc);
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    (void)hipFree(t.d_src);
    (void)hipFree(t.d_dst);
    t.d_src = nullptr;
    t.d_dst = nullptr;
    t.n_esc = nesc;
    return PG_OK;
}

// Builds the cold pre-pass of the tiles: for each tile the first edge whose source
// is cold (tile lists are sorted by source), the cold value array in tile order,
// and the pre-pass lists grouped by source chunk.  Fills the cold fields of parts.
static int build_cold(TileSet &t, int64_t n, int32_t cold, const std::vector<int32_t> &vbeg,
                      const std::vector<int32_t> &off_h, std::vector<TilePart> &parts) {
    const int nt = (int)vbeg.size() - 1;
    const int64_t p0 = vbeg.front(), mt = (int64_t)off_h.back() - off_h[0];
    std::vector<int64_t> tb((size_t)nt + 1);
    for (int k = 0; k <= nt; k++) tb[k] = (int64_t)off_h[vbeg[k] - p0] - off_h[0];
    DevBuf dtb, dcs, dcb;
    PG_TRY(dtb.alloc(sizeof(int64_t) * (nt + 1)));
    PG_TRY(dcs.alloc(sizeof(int64_t) * (nt + 1)));
    PG_TRY(dcb.alloc(sizeof(int64_t) * (nt + 1)));
    PG_CHECK_HIP(hipMemcpy(dtb.p, tb.data(), sizeof(int64_t) * tb.size(), hipMemcpyHostToDevice));
    PG_LAUNCH(k_tile_cold_start, grid_for(nt), t.d_src, dtb.as<int64_t>(), nt, cold, dcs.as<int64_t>());
    std::vector<int64_t> cs((size_t)nt), cb((size_t)nt + 1);
    // the library stream does not synchronise with the null stream: copy on it
    PG_CHECK_HIP(hipMemcpyAsync(cs.data(), dcs.p, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    cb[0] = 0;
    for (int k = 0; k < nt; k++) cb[k + 1] = cb[k] + (tb[k + 1] - cs[k]);
    const int64_t mc = cb[nt];
    PG_REQUIRE(mc < ((int64_t)1 << 32), PG_ERR_UNSUPPORTED, "tiles: too many cold edges");
    for (auto &pt : parts) {
        const int64_t c0 = cs[pt.tile];
        pt.cold_e = std::min(std::max(c0, pt.e0), pt.e1);
        pt.cold_v = cb[pt.tile] + (pt.cold_e - c0);
    }
    t.m_cold = mc;
    if (mc == 0) return PG_OK;
    PG_CHECK_HIP(hipMemcpy(dcb.p, cb.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice));
    DevBuf k0, k1, v0, v1, ch;
    PG_TRY(k0.alloc(sizeof(uint64_t) * mc));
    PG_TRY(k1.alloc(sizeof(uint64_t) * mc));
    PG_TRY(v0.alloc(sizeof(uint32_t) * mc));
    PG_TRY(v1.alloc(sizeof(uint32_t) * mc));
    PG_TRY(ch.alloc(sizeof(uint32_t) * mc));
    PG_LAUNCH(k_cold_keys, grid_for(mt), t.d_src, mt, dtb.as<int64_t>(), nt, dcs.as<int64_t>(),
              dcb.as<int64_t>(), cold, k0.as<uint64_t>(), v0.as<uint32_t>());
    const int64_t nchunks = (n - cold + kColdChunk - 1) / kColdChunk;
    PG_TRY(sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(), v1.as<uint32_t>(), mc,
                      32 + bits_for(nchunks + 1)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_cold_pos), sizeof(uint32_t) * (mc + 16)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_cold_srcl), sizeof(uint16_t) * (mc + 16)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_cold_val), sizeof(float) * (mc + 16)));
    PG_LAUNCH(k_cold_split, grid_for(mc), k1.as<uint64_t>(), v1.as<uint32_t>(), mc, t.d_cold_pos,
              t.d_cold_srcl, ch.as<uint32_t>());
    DevBuf dco;
    PG_TRY(dco.alloc(sizeof(int64_t) * (nchunks + 1)));
    PG_LAUNCH(k_chunk_offsets, grid_for(mc + 1), ch.as<uint32_t>(), mc, nchunks, dco.as<int64_t>());
    std::vector<int64_t> co((size_t)nchunks + 1);
    PG_CHECK_HIP(hipMemcpyAsync(co.data(), dco.p, sizeof(int64_t) * co.size(), hipMemcpyDeviceToHost,
                                g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    std::vector<ColdPart> cparts;
    const int64_t per = 65536;  // pre-pass edges per workgroup
    for (int64_t c = 0; c < nchunks; c++)
        for (int64_t a = co[c]; a < co[c + 1]; a += per)
            cparts.push_back(ColdPart{(int32_t)c, 0, a, std::min(co[c + 1], a + per)});
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_cold_parts),
                           sizeof(ColdPart) * (cparts.size() + 1)));
    if (!cparts.empty())
        PG_CHECK_HIP(hipMemcpy(t.d_cold_parts, cparts.data(), sizeof(ColdPart) * cparts.size(),
                               hipMemcpyHostToDevice));
    t.cold = ColdArgs{t.d_cold_parts, t.d_cold_pos, t.d_cold_srcl, t.d_cold_val,
                      (int32_t)cparts.size(), cold, n};
    return PG_OK;
}

// The tiles of positions [p0, p1) of a pull over the lists adj[off[p], off[p+1])
// (weights wf or wi in list order, or none): kTileMaxV consecutive positions a tile,
// parts of about tile_edges edges, sources from g->tile_cold on cold.
static int build_tile_set(pg_graph *g, TileSet &t, const int32_t *off, const int32_t *adj,
                          const float *wf, const int32_t *wi, int64_t p0, int64_t p1) {
    drop_captures(g);
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    free_tile_set(t);
    PG_REQUIRE(p1 - p0 < ((int64_t)1 << 31), PG_ERR_UNSUPPORTED, "tiles: light range too large");
    // tiles: kTileMaxV consecutive light positions; parts: about tile_edges edges
    std::vector<int32_t> off_h((size_t)(p1 - p0 + 1));
    if (p1 > p0)
        PG_CHECK_HIP(hipMemcpy(off_h.data(), off + p0, sizeof(int32_t) * off_h.size(),
                               hipMemcpyDeviceToHost));
    std::vector<int32_t> vbeg;
    std::vector<TilePart> parts;
    std::vector<TileFin> fin;
    int64_t slots = 0;  // partial elements
    for (int64_t p = p0; p < p1; p += kTileMaxV) {
        const int64_t q = std::min<int64_t>(p1, p + kTileMaxV);
        const int64_t nv = q - p;
        const int64_t e0 = (int64_t)off_h[p - p0] - off_h[0], e1 = (int64_t)off_h[q - p0] - off_h[0];
        const int64_t pe = std::max<int64_t>(g->tile_edges, nv);
        const int64_t np = std::max<int64_t>(1, (e1 - e0 + pe - 1) / pe);
        const int32_t tile = (int32_t)vbeg.size();
        for (int64_t k = 0; k < np; k++)
            parts.push_back(TilePart{tile, 0, np > 1 ? slots + k * nv : -1,
                                     e0 + (e1 - e0) * k / np, e0 + (e1 - e0) * (k + 1) / np});
        if (np > 1) {
            for (int64_t v0 = 0; v0 < nv; v0 += kBlock)
                fin.push_back(TileFin{tile, (int32_t)np, (int32_t)v0, 0, slots});
            slots += np * nv;
        }
        vbeg.push_back((int32_t)p);
    }
    const int64_t mt = p1 > p0 ? (int64_t)off_h[p1 - p0] - off_h[0] : 0;
    vbeg.push_back((int32_t)p1);
    const int nt = (int)vbeg.size() - 1;
    PG_REQUIRE(parts.size() < 0x7FFFFFFF && fin.size() < 0x7FFFFFFF, PG_ERR_UNSUPPORTED,
               "tiles: too many parts");
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_vbeg), sizeof(int32_t) * vbeg.size()));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_parts), sizeof(TilePart) * (parts.size() + 1)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_fin), sizeof(TileFin) * (fin.size() + 1)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_src), sizeof(int32_t) * (mt + 16)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_dst), sizeof(uint16_t) * (mt + 16)));
    PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_partial),
                           sizeof(unsigned long long) * (slots + 1)));
    PG_CHECK_HIP(hipMemcpy(t.d_vbeg, vbeg.data(), sizeof(int32_t) * vbeg.size(), hipMemcpyHostToDevice));
    if (!fin.empty())
        PG_CHECK_HIP(hipMemcpy(t.d_fin, fin.data(), sizeof(TileFin) * fin.size(), hipMemcpyHostToDevice));
    const bool weighted = wf || wi;
    if (weighted && mt > 0)
        PG_CHECK_HIP(hipMalloc(reinterpret_cast<void **>(&t.d_wt), sizeof(float) * (mt + 16)));
    if (mt > 0) {
        PG_REQUIRE(mt < ((int64_t)1 << 32), PG_ERR_UNSUPPORTED, "tiles: too many light edges");
        DevBuf k0, k1, v0, v1, dl;
        PG_TRY(k0.alloc(sizeof(uint64_t) * mt));
        PG_TRY(k1.alloc(sizeof(uint64_t) * mt));
        PG_TRY(v0.alloc(sizeof(uint32_t) * mt));
        PG_TRY(v1.alloc(sizeof(uint32_t) * mt));
        PG_TRY(dl.alloc(sizeof(uint16_t) * mt));
        PG_LAUNCH(k_tile_keys, grid_for((p1 - p0) * 64), off, adj, p0, p1, (int64_t)off_h[0],
                  k0.as<uint64_t>(), v0.as<uint32_t>(), dl.as<uint16_t>());
        PG_TRY(sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(), v1.as<uint32_t>(),
                          mt, 32 + bits_for(nt)));
        PG_LAUNCH(k_tile_split, grid_for(mt), k1.as<uint64_t>(), v1.as<uint32_t>(), mt,
                  dl.as<uint16_t>(), (int64_t)off_h[0], wf, wi, t.d_src, t.d_dst, t.d_wt);
        PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    }
    // cold split: every part's first cold edge; the pre-pass lists by source chunk
    for (auto &pt : parts) {
        pt.cold_e = pt.e1;
        pt.cold_v = 0;
    }
    t.cold = ColdArgs{};
    t.cold_begin = g->tile_cold;
    const bool cold_split = mt > 0 && g->tile_cold >= 0 && g->tile_cold < g->n;
    if (cold_split) PG_TRY(build_cold(t, g->n, (int32_t)g->tile_cold, vbeg, off_h, parts));
    if (mt > 0) PG_TRY(pack_tiles(t, mt, cold_split ? (int32_t)g->tile_cold : (int32_t)g->n));
    // the next iteration's cold pre-pass may overlap this one's finish pass when no
    // vertex the finish pass writes is a cold source
    t.cold_overlap = cold_split && t.cold.nparts > 0;
    for (const auto &pt : parts)
        if (pt.slot >= 0 && vbeg[(size_t)pt.tile + 1] > g->tile_cold) t.cold_overlap = false;
    if (t.cold_overlap) {
        PG_CHECK_HIP(hipEventCreateWithFlags(&t.ev_tiles, hipEventDisableTiming));
        PG_CHECK_HIP(hipEventCreateWithFlags(&t.ev_cold, hipEventDisableTiming));
    }
    if (!parts.empty())
        PG_CHECK_HIP(hipMemcpy(t.d_parts, parts.data(), sizeof(TilePart) * parts.size(),
                               hipMemcpyHostToDevice));
    t.args = TileArgs{t.d_vbeg,    t.d_parts,   t.d_fin,      t.d_packed,
                      t.d_gbase,   t.d_esc,     t.d_partial,  t.d_cold_val,
                      (int32_t)parts.size(),    (int32_t)fin.size(), t.d_wt};
    t.ntiles = nt;
    t.m_tiled = mt;
    t.tile_edges = g->tile_edges;
    t.light_begin = p0;
    t.light_end = p1;
    t.built = true;
    return PG_OK;
}

int ensure_tiles(pg_graph *g) {
    PullSchedule &s = g->in_sched;
    PG_REQUIRE(s.built, PG_ERR_STATE, "tiles: the pull schedule is not built");
    TileSet &t = g->tiles;
    if (t.built && t.light_begin == s.light_begin && t.light_end == s.light_end &&
        t.tile_edges == g->tile_edges && t.cold_begin == g->tile_cold)
        return PG_OK;
    return build_tile_set(g, t, g->in_off, g->in_adj,
                          g->w_dtype == PG_W_F32 ? g->in_w_f32 : (const float *)nullptr,
                          g->w_dtype == PG_W_I32 ? g->in_w_i32 : (const int32_t *)nullptr,
                          s.light_begin, s.light_end);
}

int ensure_tiles_out(pg_graph *g) {
    TileSet &t = g->tiles_out;
    if (t.built && t.tile_edges == g->tile_edges && t.cold_begin == g->tile_cold) return PG_OK;
    return build_tile_set(g, t, g->out_off, g->out_adj,
                          g->w_dtype == PG_W_F32 ? g->out_w_f32 : (const float *)nullptr,
                          g->w_dtype == PG_W_I32 ? g->out_w_i32 : (const int32_t *)nullptr, 0,
                          g->n);
}

int ensure_out_schedule(pg_graph *g) {
    if (g->out_sched.built) return PG_OK;
    if (!g->out_order) {
        DevBuf vk, vk2;
        PG_TRY(vk.alloc(sizeof(uint64_t) * g->n));
        PG_TRY(vk2.alloc(sizeof(uint64_t) * g->n));
        PG_LAUNCH(k_out_order_keys, grid_for(g->n), g->out_off, g->n, vk.as<uint64_t>());
        PG_TRY(sort_keys(vk.as<uint64_t>(), vk2.as<uint64_t>(), g->n, 64));
        PG_TRY(galloc_t(g, &g->out_order, g->n));
        PG_LAUNCH(k_low32, grid_for(g->n), vk2.as<uint64_t>(), g->n, g->out_order);
    }
    return build_schedule(g, false);
}

// ---- graph construction ---------------------------------------------------------------
// keys: input-id (src<<32|dst) pairs in input CSR order (sorted by src, and for
// generated graphs by (src,dst)); m entries; w_input: device weights in the
// same order or nullptr.
static int build_graph(pg_graph *g, uint64_t *keys, int64_t m, const void *w_input, int w_dtype) {
    const int64_t n = g->n;
    g->m = m;
    const int vb = bits_for(n + 1);
    DevBuf d_out, d_in, d_cnt, d_vk, d_vk2, d_k2, d_e1, d_e2;
    PG_TRY(d_out.alloc(sizeof(uint32_t) * (n + 1)));
    PG_TRY(d_in.alloc(sizeof(uint32_t) * (n + 1)));
    PG_CHECK_HIP(hipMemsetAsync(d_out.p, 0, sizeof(uint32_t) * (n + 1), g_stream));
    PG_CHECK_HIP(hipMemsetAsync(d_in.p, 0, sizeof(uint32_t) * (n + 1), g_stream));
    PG_LAUNCH(k_degrees, grid_for(m), keys, m, d_out.as<uint32_t>(), d_in.as<uint32_t>());
    // relabel by in-degree, descending, ties by input id
    PG_TRY(d_vk.alloc(sizeof(uint64_t) * n));
    PG_TRY(d_vk2.alloc(sizeof(uint64_t) * n));
    PG_LAUNCH(k_vertex_keys, grid_for(n), d_in.as<uint32_t>(), n, d_vk.as<uint64_t>());
    PG_TRY(sort_keys(d_vk.as<uint64_t>(), d_vk2.as<uint64_t>(), n, 64));
    PG_TRY(galloc_t(g, &g->perm, n));
    PG_TRY(galloc_t(g, &g->iperm, n));
    PG_LAUNCH(k_perm_from_sorted, grid_for(n), d_vk2.as<uint64_t>(), n, g->iperm, g->perm);
    d_vk.alloc(8);
    d_vk2.alloc(8);
    // CSR in internal ids
    PG_TRY(d_k2.alloc(sizeof(uint64_t) * m));
    uint64_t *k_sorted = nullptr;
    PG_TRY(d_e1.alloc(sizeof(uint32_t) * m));
    PG_TRY(d_e2.alloc(sizeof(uint32_t) * m));
    PG_TRY(d_cnt.alloc(sizeof(uint32_t) * (n + 1)));
    DevBuf d_k3;
    PG_TRY(d_k3.alloc(sizeof(uint64_t) * m));
    k_sorted = d_k3.as<uint64_t>();
    for (int dir = 0; dir < (g->sym ? 1 : 2); dir++) {
        PG_LAUNCH(k_relabel_keys, grid_for(m), keys, m, g->perm, dir, d_k2.as<uint64_t>(),
                  d_e1.as<uint32_t>());
        PG_TRY(sort_pairs(d_k2.as<uint64_t>(), k_sorted, d_e1.as<uint32_t>(), d_e2.as<uint32_t>(), m,
                          32 + vb));
        int32_t *off = nullptr, *adj = nullptr, *eid = nullptr;
        PG_TRY(galloc_t(g, &off, n + 1));
        PG_TRY(galloc_t(g, &adj, m));
        PG_TRY(galloc_t(g, &eid, m));
        PG_LAUNCH(k_permute_counts, grid_for(n + 1),
                  dir == 0 ? d_out.as<uint32_t>() : d_in.as<uint32_t>(), g->perm, n,
                  d_cnt.as<uint32_t>());
        PG_TRY(exclusive_sum(d_cnt.as<uint32_t>(), reinterpret_cast<uint32_t *>(off), n + 1));
        PG_LAUNCH(k_low32, grid_for(m), k_sorted, m, adj);
        PG_CHECK_HIP(hipMemcpyAsync(eid, d_e2.p, sizeof(int32_t) * m, hipMemcpyDeviceToDevice, g_stream));
        if (dir == 0) {
            g->out_off = off; g->out_adj = adj; g->out_eid = eid;
        } else {
            g->in_off = off; g->in_adj = adj; g->in_eid = eid;
        }
    }
    if (g->sym) {
        g->in_off = g->out_off; g->in_adj = g->out_adj; g->in_eid = g->out_eid;
    }
    // weights
    g->w_dtype = w_dtype;
    if (w_dtype == PG_W_I32) {
        PG_TRY(galloc_t(g, &g->out_w_i32, m));
        PG_LAUNCH(k_gather<int32_t>, grid_for(m), (const int32_t *)w_input, g->out_eid, m, g->out_w_i32);
        if (g->sym) g->in_w_i32 = g->out_w_i32;
        else {
            PG_TRY(galloc_t(g, &g->in_w_i32, m));
            PG_LAUNCH(k_gather<int32_t>, grid_for(m), (const int32_t *)w_input, g->in_eid, m, g->in_w_i32);
        }
    } else if (w_dtype == PG_W_F32) {
        PG_TRY(galloc_t(g, &g->out_w_f32, m));
        PG_LAUNCH(k_gather<float>, grid_for(m), (const float *)w_input, g->out_eid, m, g->out_w_f32);
        if (g->sym) g->in_w_f32 = g->out_w_f32;
        else {
            PG_TRY(galloc_t(g, &g->in_w_f32, m));
            PG_LAUNCH(k_gather<float>, grid_for(m), (const float *)w_input, g->in_eid, m, g->in_w_f32);
        }
    }
    // max degrees
    int32_t o0[2] = {0, 0};
    if (n > 0) {
        PG_CHECK_HIP(hipMemcpyAsync(o0, g->in_off, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, g_stream));
    }
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    g->max_in_deg = o0[1] - o0[0];
    PG_TRY(build_schedule(g, true));
    return PG_OK;
}

static int new_graph(int64_t n, pg_graph **out) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(out, PG_ERR_ARG, "null output handle");
    PG_REQUIRE(n > 0 && n < 0x7FFFFFFF, PG_ERR_ARG, "n out of range: %lld", (long long)n);
    *out = new pg_graph();
    (*out)->n = n;
    return PG_OK;
}

static void delete_graph(pg_graph *g) {
    if (!g) return;
    drop_captures(g);
    (void)hipStreamSynchronize(g_stream);
    free_schedule(g->in_sched);
    free_schedule(g->out_sched);
    free_tiles(g);
    for (void *p : g->owned) (void)hipFree(p);
    for (auto &kv : g->ws) (void)hipFree(kv.second.first);
    delete g;
}

// ---- a graph from a CSR in device memory ------------------------------------------
// Checks (offsets monotone from 0 to m, neighbours in range; the first offending
// index is reported), then forms the (source, neighbour) keys of build_graph: an
// edge's source is the number of vertices v >= 1 whose list starts at or before
// it (an inclusive scan of start marks), so hubs cost no more than other
// vertices.  PG_GRAPH_SKIP_SELF_LOOPS drops v -> v edges (IO-numa.h:225-268) with a
// compaction scan, their weights with them.
__global__ void k_csr_check(const int64_t *off, int64_t n, const int32_t *adj, int64_t m,
                            unsigned long long *err) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        if (off[v] > off[v + 1]) atomicMin(&err[0], (unsigned long long)v);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride)
        if (adj[e] < 0 || adj[e] >= n) atomicMin(&err[1], (unsigned long long)e);
}

__global__ void k_mark_starts(const int64_t *off, int64_t n, int64_t m, uint32_t *mark) {
    for (int64_t v = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        if (off[v] < m) atomicAdd(&mark[off[v]], 1u);
}

// keep[e] = 1 unless e is a self-loop that is dropped; keys when nothing is dropped
__global__ void k_edge_keys(const uint32_t *src, const int32_t *adj, int64_t m, bool skip_loops,
                            uint32_t *keep, uint64_t *keys) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t s = src[e];
        const uint32_t d = (uint32_t)adj[e];
        if (skip_loops) keep[e] = s != d ? 1u : 0u;
        else keys[e] = ((uint64_t)s << 32) | d;
    }
}

__global__ void k_edge_keys_compact(const uint32_t *src, const int32_t *adj, const uint32_t *keep,
                                    const uint32_t *pos, const int32_t *w, int64_t m, uint64_t *keys,
                                    int32_t *wout) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[e]) continue;
        keys[pos[e]] = ((uint64_t)src[e] << 32) | (uint32_t)adj[e];
        if (w) wout[pos[e]] = w[e];
    }
}

int create_from_device(int64_t n, int64_t m, const int64_t *d_off, const int32_t *d_adj,
                       const void *d_w, int w_dtype, uint32_t flags, pg_graph **out) {
    *out = nullptr;
    PG_REQUIRE(n > 0 && n < 0x7FFFFFFF, PG_ERR_ARG, "n out of range: %lld", (long long)n);
    PG_REQUIRE(m >= 0 && m < 0x7FFFFFFF, PG_ERR_ARG, "m out of range: %lld", (long long)m);
    PG_REQUIRE(w_dtype == PG_W_NONE || w_dtype == PG_W_I32 || w_dtype == PG_W_F32, PG_ERR_ARG,
               "bad weight dtype %d", w_dtype);
    PG_REQUIRE(w_dtype == PG_W_NONE || d_w || m == 0, PG_ERR_ARG, "weights requested but w is NULL");
    int64_t ends[2];
    PG_CHECK_HIP(hipMemcpyAsync(&ends[0], d_off, 8, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipMemcpyAsync(&ends[1], d_off + n, 8, hipMemcpyDeviceToHost, g_stream));
    DevBuf derr;
    PG_TRY(derr.alloc(16));
    PG_LAUNCH(k_fill<unsigned long long>, 1, derr.as<unsigned long long>(), ~0ull, 2);
    PG_LAUNCH(k_csr_check, grid_for(std::max(n, m)), d_off, n, d_adj, m, derr.as<unsigned long long>());
    unsigned long long err[2];
    PG_CHECK_HIP(hipMemcpyAsync(err, derr.p, 16, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    PG_REQUIRE(ends[0] == 0 && ends[1] == m, PG_ERR_ARG, "offsets must span [0, m]");
    PG_REQUIRE(err[0] == ~0ull, PG_ERR_ARG, "offsets not monotone at %lld", (long long)err[0]);
    if (err[1] != ~0ull) {
        int32_t d = 0;
        PG_CHECK_HIP(hipMemcpy(&d, d_adj + err[1], 4, hipMemcpyDeviceToHost));
        set_error("edge %lld: neighbour %d out of range", (long long)err[1], d);
        return PG_ERR_ARG;
    }
    const bool skip = (flags & PG_GRAPH_SKIP_SELF_LOOPS) != 0;
    DevBuf mark, src, keys, keep, pos, wc;
    PG_TRY(mark.alloc(4 * (size_t)m + 16));
    PG_TRY(src.alloc(4 * (size_t)m + 16));
    PG_TRY(keys.alloc(8 * (size_t)m + 16));
    int64_t mm = m;
    const void *wk = d_w;
    if (m > 0) {
        PG_LAUNCH(k_fill<uint32_t>, grid_for(m), mark.as<uint32_t>(), 0u, m);
        PG_LAUNCH(k_mark_starts, grid_for(n), d_off, n, m, mark.as<uint32_t>());
        PG_TRY(inclusive_sum(mark.as<uint32_t>(), src.as<uint32_t>(), m));
        if (!skip) {
            PG_LAUNCH(k_edge_keys, grid_for(m), src.as<uint32_t>(), d_adj, m, false, nullptr,
                      keys.as<uint64_t>());
        } else {
            PG_TRY(keep.alloc(4 * (size_t)m + 16));
            PG_TRY(pos.alloc(4 * (size_t)m + 16));
            PG_LAUNCH(k_edge_keys, grid_for(m), src.as<uint32_t>(), d_adj, m, true, keep.as<uint32_t>(),
                      nullptr);
            PG_TRY(exclusive_sum(keep.as<uint32_t>(), pos.as<uint32_t>(), m));
            uint32_t last[2];
            PG_CHECK_HIP(hipMemcpyAsync(&last[0], pos.as<uint32_t>() + m - 1, 4, hipMemcpyDeviceToHost,
                                        g_stream));
            PG_CHECK_HIP(hipMemcpyAsync(&last[1], keep.as<uint32_t>() + m - 1, 4,
                                        hipMemcpyDeviceToHost, g_stream));
            PG_CHECK_HIP(hipStreamSynchronize(g_stream));
            mm = (int64_t)last[0] + last[1];
            if (d_w) {
                PG_TRY(wc.alloc(4 * (size_t)mm + 16));
                wk = wc.p;
            }
            PG_LAUNCH(k_edge_keys_compact, grid_for(m), src.as<uint32_t>(), d_adj, keep.as<uint32_t>(),
                      pos.as<uint32_t>(), static_cast<const int32_t *>(d_w), m, keys.as<uint64_t>(),
                      wc.as<int32_t>());
        }
    }
    PG_TRY(new_graph(n, out));
    pg_graph *g = *out;
    g->sym = (flags & PG_GRAPH_SYMMETRIC) != 0;
    const int r = build_graph(g, keys.as<uint64_t>(), mm, w_dtype == PG_W_NONE ? nullptr : wk, w_dtype);
    if (r != PG_OK) {
        delete_graph(g);
        *out = nullptr;
    }
    return r;
}

// a host CSR: uploaded, then built as a device CSR
int create_from_host(int64_t n, int64_t m, const int64_t *off, const int32_t *adj, const void *w,
                     int w_dtype, uint32_t flags, pg_graph **out) {
    *out = nullptr;
    PG_REQUIRE(m >= 0 && m < 0x7FFFFFFF, PG_ERR_ARG, "m out of range: %lld", (long long)m);
    PG_REQUIRE(off && (adj || m == 0), PG_ERR_ARG, "null CSR arrays");
    PG_REQUIRE(w_dtype == PG_W_NONE || w || m == 0, PG_ERR_ARG, "weights requested but w is NULL");
    DevBuf doff, dadj, dw;
    PG_TRY(doff.alloc(8 * ((size_t)n + 1)));
    PG_TRY(dadj.alloc(4 * (size_t)m + 16));
    PG_CHECK_HIP(hipMemcpyAsync(doff.p, off, 8 * ((size_t)n + 1), hipMemcpyHostToDevice, g_stream));
    if (m) PG_CHECK_HIP(hipMemcpyAsync(dadj.p, adj, 4 * (size_t)m, hipMemcpyHostToDevice, g_stream));
    if (w_dtype != PG_W_NONE) {
        PG_TRY(dw.alloc(4 * (size_t)m + 16));
        if (m) PG_CHECK_HIP(hipMemcpyAsync(dw.p, w, 4 * (size_t)m, hipMemcpyHostToDevice, g_stream));
    }
    return create_from_device(n, m, doff.as<int64_t>(), dadj.as<int32_t>(),
                              w_dtype == PG_W_NONE ? nullptr : dw.p, w_dtype, flags, out);
}

// device rMat: the first m distinct non-loop pairs of the counter sequence
static int rmat_keys(int scale, int64_t m, double a, double b, double c, uint64_t seed,
                     DevBuf &out_keys) {
    uint32_t thr[3];
    host_rmat_thresholds(a, b, c, thr);
    const uint64_t base = host_mix64(seed ^ 0x5851F42D4C957F2Dull);
    int64_t C = m + m / 8 + 1024;
    for (;;) {
        PG_REQUIRE(C < 0xFFFFFFFFll, PG_ERR_ARG, "rmat: too many duplicates for m=%lld", (long long)m);
        DevBuf k1, k2, c1, c2, mark, cnt, cut, sel, nsel;
        PG_TRY(k1.alloc(8 * C)); PG_TRY(k2.alloc(8 * C));
        PG_TRY(c1.alloc(4 * C)); PG_TRY(c2.alloc(4 * C));
        PG_LAUNCH(k_rmat_keys, grid_for(C), base, scale, thr[0], thr[1], thr[2], C,
                  k1.as<uint64_t>(), c1.as<uint32_t>());
        PG_TRY(sort_pairs(k1.as<uint64_t>(), k2.as<uint64_t>(), c1.as<uint32_t>(), c2.as<uint32_t>(),
                          C, 64));
        PG_TRY(mark.alloc(4 * C));
        PG_CHECK_HIP(hipMemsetAsync(mark.p, 0, 4 * C, g_stream));
        PG_LAUNCH(k_first_by_ctr, grid_for(C), k2.as<uint64_t>(), c2.as<uint32_t>(), C, mark.as<uint32_t>());
        PG_TRY(cnt.alloc(4 * C));
        PG_TRY(inclusive_sum(mark.as<uint32_t>(), cnt.as<uint32_t>(), C));
        uint32_t total = 0;
        PG_CHECK_HIP(hipMemcpyAsync(&total, cnt.as<uint32_t>() + (C - 1), 4, hipMemcpyDeviceToHost, g_stream));
        PG_CHECK_HIP(hipStreamSynchronize(g_stream));
        if ((int64_t)total >= m) {
            PG_TRY(cut.alloc(8));
            int64_t minus1 = -1;
            PG_CHECK_HIP(hipMemcpyAsync(cut.p, &minus1, 8, hipMemcpyHostToDevice, g_stream));
            if (m > 0) PG_LAUNCH(k_find_cutoff, grid_for(C), cnt.as<uint32_t>(), C, (uint32_t)m, cut.as<int64_t>());
            PG_TRY(sel.alloc(C));
            PG_LAUNCH(k_select_flags, grid_for(C), k2.as<uint64_t>(), c2.as<uint32_t>(), C,
                      cut.as<int64_t>(), sel.as<uint8_t>());
            PG_TRY(out_keys.alloc(8 * (m + 1)));
            PG_TRY(nsel.alloc(8));
            PG_TRY(select_flagged(k2.as<uint64_t>(), sel.as<uint8_t>(), out_keys.as<uint64_t>(),
                                  nsel.as<unsigned long long>(), C));
            unsigned long long got = 0;
            PG_CHECK_HIP(hipMemcpyAsync(&got, nsel.p, 8, hipMemcpyDeviceToHost, g_stream));
            PG_CHECK_HIP(hipStreamSynchronize(g_stream));
            PG_REQUIRE((int64_t)got == m, PG_ERR_STATE, "rmat: selected %llu of %lld", got, (long long)m);
            return PG_OK;
        }
        C = C + C / 2 + 1024;
    }
}

static int symmetrize_keys(DevBuf &keys, int64_t m, int64_t *m_out) {
    DevBuf k2, k3, sel, nsel;
    PG_TRY(k2.alloc(8 * (2 * m + 1)));
    PG_TRY(k3.alloc(8 * (2 * m + 1)));
    PG_LAUNCH(k_mirror, grid_for(m), keys.as<uint64_t>(), m, k2.as<uint64_t>());
    PG_TRY(sort_keys(k2.as<uint64_t>(), k3.as<uint64_t>(), 2 * m, 64));
    PG_TRY(sel.alloc(2 * m + 1));
    PG_LAUNCH(k_unique_flags, grid_for(2 * m), k3.as<uint64_t>(), 2 * m, sel.as<uint8_t>());
    PG_TRY(keys.alloc(8 * (2 * m + 1)));
    PG_TRY(nsel.alloc(8));
    PG_TRY(select_flagged(k3.as<uint64_t>(), sel.as<uint8_t>(), keys.as<uint64_t>(),
                          nsel.as<unsigned long long>(), 2 * m));
    unsigned long long got = 0;
    PG_CHECK_HIP(hipMemcpyAsync(&got, nsel.p, 8, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    *m_out = (int64_t)got;
    return PG_OK;
}

int create_rmat(int scale, int64_t m, double a, double b, double c, uint64_t seed, uint32_t flags,
                uint64_t wseed, pg_graph **out) {
    PG_REQUIRE(scale >= 1 && scale <= 30, PG_ERR_ARG, "scale out of range: %d", scale);
    PG_REQUIRE(m >= 0 && m < 0x40000000ll, PG_ERR_ARG, "m out of range");
    PG_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, PG_ERR_ARG, "bad rmat probabilities");
    PG_REQUIRE(!((flags & PG_GRAPH_WEIGHT_I32) && (flags & PG_GRAPH_VALUE_F32)), PG_ERR_ARG,
               "choose one weight kind");
    PG_TRY(new_graph(int64_t(1) << scale, out));
    pg_graph *g = *out;
    DevBuf keys, w;
    int64_t mm = m;
    int r = rmat_keys(scale, m, a, b, c, seed, keys);
    if (r == PG_OK && (flags & PG_GRAPH_SYMMETRIZE)) {
        r = symmetrize_keys(keys, m, &mm);
        g->sym = true;
    }
    int wd = PG_W_NONE;
    if (r == PG_OK && (flags & PG_GRAPH_WEIGHT_I32)) {
        wd = PG_W_I32;
        r = w.alloc(4 * (mm + 1));
        if (r == PG_OK && mm) {
            hipLaunchKernelGGL(k_pair_weights_i32, dim3(grid_for(mm)), dim3(kBlock), 0, g_stream,
                               keys.as<uint64_t>(), mm, wseed, w.as<int32_t>());
            if (hipGetLastError() != hipSuccess) { set_error("weights launch failed"); r = PG_ERR_HIP; }
        }
    } else if (r == PG_OK && (flags & PG_GRAPH_VALUE_F32)) {
        wd = PG_W_F32;
        r = w.alloc(4 * (mm + 1));
        if (r == PG_OK && mm) {
            hipLaunchKernelGGL(k_pair_values_f32, dim3(grid_for(mm)), dim3(kBlock), 0, g_stream,
                               keys.as<uint64_t>(), mm, wseed, w.as<float>());
            if (hipGetLastError() != hipSuccess) { set_error("values launch failed"); r = PG_ERR_HIP; }
        }
    }
    if (r == PG_OK) r = build_graph(g, keys.as<uint64_t>(), mm, w.p, wd);
    if (r != PG_OK) {
        delete_graph(g);
        *out = nullptr;
    }
    return r;
}

}  // namespace pg

using namespace pg;

// ====================================================================================
// C ABI
// ====================================================================================
extern "C" int pg_init(int device) {
    if (g_init) {
        PG_REQUIRE(device == g_device, PG_ERR_STATE, "already initialised on device %d", g_device);
        return PG_OK;
    }
    int count = 0;
    PG_CHECK_HIP(hipGetDeviceCount(&count));
    PG_REQUIRE(device >= 0 && device < count, PG_ERR_ARG, "device %d not present (%d devices)",
               device, count);
    PG_CHECK_HIP(hipSetDevice(device));
    if (device < kMaxDevices && !g_scratch[device]) {
        void *p = nullptr;
        PG_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[device] = static_cast<uint32_t *>(p);
    }
    PG_CHECK_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    PG_CHECK_HIP(hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking));
    g_device = device;
    g_init = true;
    return PG_OK;
}

extern "C" int pg_finalize(void) {
    if (!g_init) return PG_OK;
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_side));
    PG_CHECK_HIP(hipStreamDestroy(g_stream));
    PG_CHECK_HIP(hipStreamDestroy(g_side));
    g_stream = nullptr;
    g_side = nullptr;
    g_init = false;
    g_device = -1;
    return PG_OK;
}

extern "C" int pg_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return PG_ERR_ARG;
    snprintf(buf, len, "%s", g_err);
    return PG_OK;
}

extern "C" int pg_device_name(char *buf, size_t len) {
    PG_REQUIRE(g_init, PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(buf && len > 0, PG_ERR_ARG, "null buffer");
    hipDeviceProp_t p;
    PG_CHECK_HIP(hipGetDeviceProperties(&p, g_device));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return PG_OK;
}

extern "C" int pg_graph_create(int64_t n, int64_t m, const int64_t *off, const int32_t *adj,
                               const void *w, int w_dtype, uint32_t flags, pg_graph **out) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(out, PG_ERR_ARG, "null output handle");
    *out = nullptr;
    PG_REQUIRE(n > 0 && n < 0x7FFFFFFF, PG_ERR_ARG, "n out of range: %lld", (long long)n);
    return create_from_host(n, m, off, adj, w, w_dtype, flags, out);
}

extern "C" int pg_graph_rmat(int scale, int64_t m, double a, double b, double c, uint64_t seed,
                             uint32_t flags, uint64_t wseed, pg_graph **out) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(out, PG_ERR_ARG, "null output handle");
    *out = nullptr;
    return create_rmat(scale, m, a, b, c, seed, flags, wseed, out);
}

// ---- PBBS text on the GPU (readGraphFromFile IO-numa.h:112-190, IO.h:186-272) ------
// The file is read into pinned host memory by parallel preads and copied to the
// device once.  There, each thread takes 16 bytes and marks the bytes that start a
// word (not a separator, after a separator: IO-numa.h:90-109's words); per-block
// counts and an exclusive scan number the words, and each word is parsed by the
// thread that holds its first byte (an optional sign, then digits, as strtoll) into
// its slot: words 3 .. 3+n-1 the offsets, then m neighbour ids, then m weights.
inline bool txt_sep_host(char c) {
    return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == 0;
}

constexpr int kTxtThreadBytes = 16;
constexpr int kTxtBlockBytes = kBlock * kTxtThreadBytes;

__device__ __forceinline__ bool txt_sep(unsigned c) {
    return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == 0;
}

// bit k: byte p0 + k starts a word
__device__ __forceinline__ unsigned txt_starts(const uint8_t *__restrict__ txt, int64_t size,
                                               int64_t p0) {
    if (p0 >= size) return 0u;
    uint8_t b[kTxtThreadBytes];
    if (p0 + kTxtThreadBytes <= size) {
        const uint4 v = *reinterpret_cast<const uint4 *>(txt + p0);
        __builtin_memcpy(b, &v, 16);
    } else {
        for (int k = 0; k < kTxtThreadBytes; k++) b[k] = p0 + k < size ? txt[p0 + k] : ' ';
    }
    unsigned prev = p0 == 0 ? ' ' : txt[p0 - 1];
    unsigned mask = 0;
#pragma unroll
    for (int k = 0; k < kTxtThreadBytes; k++) {
        if (!txt_sep(b[k]) && txt_sep(prev)) mask |= 1u << k;
        prev = b[k];
    }
    return mask;
}

__global__ __launch_bounds__(kBlock) void k_txt_count(const uint8_t *txt, int64_t size,
                                                     uint32_t *bcount) {
    const int64_t p0 = (int64_t)blockIdx.x * kTxtBlockBytes + (int64_t)threadIdx.x * kTxtThreadBytes;
    unsigned c = __popc(txt_starts(txt, size, p0));
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    __shared__ unsigned s[kBlock / 64];
    if (lane_id() == 0) s[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned tot = 0;
        for (int i = 0; i < kBlock / 64; i++) tot += s[i];
        bcount[blockIdx.x] = tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_txt_parse(const uint8_t *txt, int64_t size,
                                                     const uint64_t *bbase, int64_t n, int64_t m,
                                                     bool weighted, int64_t *off, int32_t *adj,
                                                     int32_t *w) {
    const int64_t p0 = (int64_t)blockIdx.x * kTxtBlockBytes + (int64_t)threadIdx.x * kTxtThreadBytes;
    unsigned mask = txt_starts(txt, size, p0);
    // the block's words before this thread's: a wave scan, then the waves' totals
    const int lane = lane_id();
    unsigned x = __popc(mask);
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    __shared__ unsigned s[kBlock / 64];
    if (lane == 63) s[threadIdx.x >> 6] = x;
    __syncthreads();
    unsigned before = x - __popc(mask);
    for (int i = 0; i < (int)(threadIdx.x >> 6); i++) before += s[i];
    uint64_t j = bbase[blockIdx.x] + before;
    const uint64_t nn = (uint64_t)n, mm = (uint64_t)m;
    while (mask) {
        const int k = __ffs(mask) - 1;
        mask &= mask - 1;
        int64_t p = p0 + k;
        bool neg = false;
        if (txt[p] == '-' || txt[p] == '+') {
            neg = txt[p] == '-';
            p++;
        }
        long long val = 0;
        while (p < size) {
            const unsigned c = txt[p];
            if (c < '0' || c > '9') break;
            val = val * 10 + (long long)(c - '0');
            p++;
        }
        if (neg) val = -val;
        if (j >= 3) {
            const uint64_t i = j - 3;
            if (i < nn) off[i] = val;
            else if (i < nn + mm) adj[i - nn] = (int32_t)val;
            else if (weighted && i < nn + 2 * mm) w[i - nn - mm] = (int32_t)val;
        }
        j++;
    }
}

// the whole file into pinned host memory, read in parallel
static int read_file_pinned(const char *path, uint8_t **buf, int64_t *size) {
    *buf = nullptr;
    const int fd = open(path, O_RDONLY);
    PG_REQUIRE(fd >= 0, PG_ERR_ARG, "Unable to open file: %s", path);  // IO-numa.h:75-78
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        set_error("Unable to open file: %s", path);
        return PG_ERR_ARG;
    }
    const int64_t sz = (int64_t)sb.st_size;
    void *p = nullptr;
    if (hipHostMalloc(&p, (size_t)sz + 16, hipHostMallocDefault) != hipSuccess) {
        close(fd);
        set_error("hipHostMalloc(%lld) failed", (long long)sz + 16);
        return PG_ERR_OOM;
    }
    uint8_t *b = static_cast<uint8_t *>(p);
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(16, sz >> 26));
    std::vector<int> ok(nt, 1);
    std::vector<std::thread> th;
    const int64_t chunk = (sz + nt - 1) / nt;
    for (int i = 0; i < nt; i++) {
        th.emplace_back([=, &ok]() {
            int64_t pos = (int64_t)i * chunk;
            const int64_t endp = std::min(sz, pos + chunk);
            while (pos < endp) {
                const ssize_t r = pread(fd, b + pos, (size_t)std::min<int64_t>(endp - pos, 1 << 30), pos);
                if (r <= 0) {
                    ok[i] = 0;
                    return;
                }
                pos += r;
            }
        });
    }
    for (auto &x : th) x.join();
    close(fd);
    for (int i = 0; i < nt; i++)
        if (!ok[i]) {
            (void)hipHostFree(p);
            set_error("read failed: %s", path);
            return PG_ERR_ARG;
        }
    std::memset(b + sz, 0, 16);
    *buf = b;
    *size = sz;
    return PG_OK;
}

extern "C" int pg_graph_load_adj(const char *path, uint32_t flags, pg_graph **out) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(path && out, PG_ERR_ARG, "null argument");
    *out = nullptr;
    uint8_t *hb = nullptr;
    int64_t size = 0;
    PG_TRY(read_file_pinned(path, &hb, &size));
    // the header words (IO-numa.h:112-122): the format's name, n, m
    const char *p = reinterpret_cast<const char *>(hb), *endp = p + size;
    auto next_word = [&](std::string *wd) -> bool {
        while (p < endp && txt_sep_host(*p)) p++;
        if (p >= endp) return false;
        const char *b = p;
        while (p < endp && !txt_sep_host(*p)) p++;
        wd->assign(b, p);
        return true;
    };
    std::string head, sn, sm;
    const bool got = next_word(&head) && next_word(&sn) && next_word(&sm);
    const bool weighted = got && head == "WeightedAdjacencyGraph";
    const long long n = got ? strtoll(sn.c_str(), nullptr, 10) : 0;
    const long long m = got ? strtoll(sm.c_str(), nullptr, 10) : 0;
    if (!got || !(weighted || head == "AdjacencyGraph") || n <= 0 || m < 0 || n >= 0x7FFFFFFF ||
        m >= 0x7FFFFFFF) {
        (void)hipHostFree(hb);
        set_error("Bad input file: %s", path);
        return PG_ERR_ARG;
    }
    DevBuf dtxt, dcnt, dbase, doff, dadj, dw;
    int r = dtxt.alloc((size_t)size + 16);
    if (r == PG_OK) {
        const hipError_t e = hipMemcpyAsync(dtxt.p, hb, (size_t)size + 16, hipMemcpyHostToDevice, g_stream);
        if (e == hipSuccess) (void)hipStreamSynchronize(g_stream);
        if (e != hipSuccess) {
            set_error("upload failed: %s", hipGetErrorString(e));
            r = PG_ERR_HIP;
        }
    }
    (void)hipHostFree(hb);
    PG_TRY(r);
    const int64_t nb = (size + kTxtBlockBytes - 1) / kTxtBlockBytes;
    PG_REQUIRE(nb < 0x7FFFFFFF, PG_ERR_UNSUPPORTED, "file too large: %s", path);
    PG_TRY(dcnt.alloc(4 * (size_t)nb + 16));
    PG_TRY(dbase.alloc(8 * (size_t)nb + 16));
    PG_LAUNCH(k_txt_count, nb, dtxt.as<uint8_t>(), size, dcnt.as<uint32_t>());
    PG_TRY((exclusive_sum<uint32_t, uint64_t>(dcnt.as<uint32_t>(), dbase.as<uint64_t>(), nb)));
    uint64_t lastb = 0;
    uint32_t lastc = 0;
    PG_CHECK_HIP(hipMemcpyAsync(&lastb, dbase.as<uint64_t>() + nb - 1, 8, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipMemcpyAsync(&lastc, dcnt.as<uint32_t>() + nb - 1, 4, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    const uint64_t words = lastb + lastc;
    const uint64_t need_off = 3 + (uint64_t)n, need_adj = need_off + (uint64_t)m;
    const uint64_t need = weighted ? need_adj + (uint64_t)m : need_adj;
    PG_REQUIRE(words >= need_off, PG_ERR_ARG, "Bad input file (offsets): %s", path);
    PG_REQUIRE(words >= need_adj, PG_ERR_ARG, "Bad input file (edges): %s", path);
    PG_REQUIRE(words >= need, PG_ERR_ARG, "Bad input file (weights): %s", path);
    PG_REQUIRE(words == need, PG_ERR_ARG, "Bad input file (length): %s", path);  // IO-numa.h:122
    PG_TRY(doff.alloc(8 * ((size_t)n + 1)));
    PG_TRY(dadj.alloc(4 * (size_t)m + 16));
    if (weighted) PG_TRY(dw.alloc(4 * (size_t)m + 16));
    PG_LAUNCH(k_txt_parse, nb, dtxt.as<uint8_t>(), size, dbase.as<uint64_t>(), (int64_t)n, (int64_t)m,
              weighted, doff.as<int64_t>(), dadj.as<int32_t>(), dw.as<int32_t>());
    const int64_t mm = m;
    PG_CHECK_HIP(hipMemcpyAsync(doff.as<int64_t>() + n, &mm, 8, hipMemcpyHostToDevice, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return create_from_device(n, m, doff.as<int64_t>(), dadj.as<int32_t>(), weighted ? dw.p : nullptr,
                              weighted ? PG_W_I32 : PG_W_NONE, flags, out);
}

// The binary graph of readGraphFromBinary / readWghGraphFromBinary
// (IO-numa.h:543-635, 637-720): <base>.config holds n as text, <base>.adj the m
// neighbour ids as 32-bit words (m = its size / 4), <base>.idx the n offsets of the
// out-lists (no closing m).  The reference's two readers disagree on the width of
// an offset (4 bytes at IO-numa.h:579, 8 at :672); here the width follows the
// file's size (4n or 8n bytes) and any other size is "File size wrong", where the
// reference aborts.  PG_GRAPH_WEIGHT_I32 gives every edge weight 1, as the
// weighted reader does (IO-numa.h:683).
extern "C" int pg_graph_load_bin(const char *base, uint32_t flags, pg_graph **out) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(base && out, PG_ERR_ARG, "null argument");
    *out = nullptr;
    const std::string b(base);
    long long n = 0;
    {
        std::ifstream fc(b + ".config");
        PG_REQUIRE(fc.is_open(), PG_ERR_ARG, "Unable to open file: %s.config", base);
        PG_REQUIRE((fc >> n) && n > 0 && n < 0x7FFFFFFF, PG_ERR_ARG, "Bad config file: %s.config",
                   base);
    }
    auto slurp = [](const std::string &path, std::vector<char> *buf) -> bool {
        std::ifstream f(path, std::ios::in | std::ios::binary | std::ios::ate);
        if (!f.is_open()) return false;
        const std::streamsize size = f.tellg();
        f.seekg(0);
        buf->resize((size_t)size);
        return size == 0 || (bool)f.read(buf->data(), size);
    };
    std::vector<char> abuf, ibuf;
    PG_REQUIRE(slurp(b + ".adj", &abuf), PG_ERR_ARG, "Unable to read file: %s.adj", base);
    PG_REQUIRE(slurp(b + ".idx", &ibuf), PG_ERR_ARG, "Unable to read file: %s.idx", base);
    const int64_t m = (int64_t)(abuf.size() / 4);
    const bool wide = ibuf.size() == 8 * (size_t)n;
    PG_REQUIRE(wide || ibuf.size() == 4 * (size_t)n, PG_ERR_ARG,
               "File size wrong: %s.idx has %zu bytes for n = %lld", base, ibuf.size(), n);
    std::vector<int64_t> off((size_t)n + 1);
    for (long long i = 0; i < n; i++) {
        if (wide) {
            int64_t x;
            std::memcpy(&x, ibuf.data() + 8 * i, 8);
            off[i] = x;
        } else {
            int32_t x;
            std::memcpy(&x, ibuf.data() + 4 * i, 4);
            off[i] = x;
        }
    }
    off[n] = m;
    std::vector<int32_t> adj((size_t)m);
    if (m > 0) std::memcpy(adj.data(), abuf.data(), 4 * (size_t)m);
    const bool weighted = (flags & PG_GRAPH_WEIGHT_I32) != 0;
    std::vector<int32_t> w(weighted ? (size_t)m : 0, 1);
    return create_from_host(n, m, off.data(), adj.data(), weighted ? w.data() : nullptr,
                            weighted ? PG_W_I32 : PG_W_NONE, flags & ~(uint32_t)PG_GRAPH_WEIGHT_I32,
                            out);
}

// The whole-graph dump of dumpGraphToBin / loadGraphFromBin (IO-numa.h:758-900,
// read by numa-PageRank-bin.C): int64 total size, int32 n, int64 m (20 bytes, no
// padding), then per vertex int32 -1, int32 its id, int32 out-degree, int32
// in-degree, the out-neighbours and the in-neighbours.  The out-lists define the
// graph; the markers, ids, the totals and every vertex's recorded in-degree are
// checked against them (the reference aborts on a marker or id mismatch).
extern "C" int pg_graph_load_dump(const char *path, uint32_t flags, pg_graph **out) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(path && out, PG_ERR_ARG, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::in | std::ios::binary | std::ios::ate);
    PG_REQUIRE(f.is_open(), PG_ERR_ARG, "Unable to open file: %s", path);
    const int64_t size = (int64_t)f.tellg();
    f.seekg(0);
    std::vector<char> buf((size_t)size);
    PG_REQUIRE(size == 0 || (bool)f.read(buf.data(), size), PG_ERR_ARG, "Unable to read %s", path);
    int64_t pos = 0;
    auto get = [&](void *dst, int64_t bytes) -> bool {
        if (pos + bytes > size) return false;
        std::memcpy(dst, buf.data() + pos, (size_t)bytes);
        pos += bytes;
        return true;
    };
    int64_t total = 0, m = 0;
    int32_t n = 0;
    PG_REQUIRE(get(&total, 8) && get(&n, 4) && get(&m, 8), PG_ERR_ARG, "Bad dump file: %s", path);
    PG_REQUIRE(total == size && n > 0 && m >= 0 && m < 0x7FFFFFFF, PG_ERR_ARG,
               "Bad dump file %s: size %lld, recorded %lld, n %d, m %lld", path, (long long)size,
               (long long)total, n, (long long)m);
    std::vector<int64_t> off((size_t)n + 1, 0);
    std::vector<int32_t> adj((size_t)m), indeg((size_t)n);
    int64_t e = 0, in_total = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t hdr[4];
        PG_REQUIRE(get(hdr, 16), PG_ERR_ARG, "Bad dump file %s: vertex %d truncated", path, i);
        PG_REQUIRE(hdr[0] == -1 && hdr[1] == i, PG_ERR_ARG,
                   "Bad dump file %s: vertex %d has marker %d and id %d", path, i, hdr[0], hdr[1]);
        PG_REQUIRE(hdr[2] >= 0 && hdr[3] >= 0 && e + hdr[2] <= m, PG_ERR_ARG,
                   "Bad dump file %s: vertex %d degrees %d / %d", path, i, hdr[2], hdr[3]);
        PG_REQUIRE(get(adj.data() + e, 4 * (int64_t)hdr[2]), PG_ERR_ARG,
                   "Bad dump file %s: vertex %d truncated", path, i);
        e += hdr[2];
        off[i + 1] = e;
        indeg[i] = hdr[3];
        in_total += hdr[3];
        pos += 4 * (int64_t)hdr[3];  // the in-lists: the transpose, rebuilt on the device
        PG_REQUIRE(pos <= size, PG_ERR_ARG, "Bad dump file %s: vertex %d truncated", path, i);
    }
    PG_REQUIRE(e == m && in_total == m && pos == size, PG_ERR_ARG,
               "Bad dump file %s: %lld out-edges and %lld in-edges for m = %lld", path,
               (long long)e, (long long)in_total, (long long)m);
    std::vector<int32_t> count((size_t)n, 0);
    for (int64_t k = 0; k < m; k++) {
        PG_REQUIRE(adj[k] >= 0 && adj[k] < n, PG_ERR_ARG, "Bad dump file %s: neighbour %d out of range",
                   path, adj[k]);
        count[adj[k]]++;
    }
    for (int32_t i = 0; i < n; i++)
        PG_REQUIRE(count[i] == indeg[i], PG_ERR_ARG,
                   "Bad dump file %s: vertex %d records in-degree %d, its in-edges number %d", path,
                   i, indeg[i], count[i]);
    return create_from_host(n, m, off.data(), adj.data(), nullptr, PG_W_NONE, flags, out);
}

extern "C" int pg_graph_dump(pg_graph *g, const char *path) {
    PG_REQUIRE(g && path, PG_ERR_ARG, "null argument");
    const int64_t n = g->n, m = g->m;
    std::vector<int64_t> off((size_t)n + 1);
    std::vector<int32_t> adj((size_t)std::max<int64_t>(m, 1));
    PG_TRY(pg_graph_download(g, off.data(), adj.data(), nullptr));
    // in-lists by a counting sort of the out-edges on their destination
    std::vector<int64_t> ioff((size_t)n + 1, 0);
    for (int64_t k = 0; k < m; k++) ioff[(size_t)adj[k] + 1]++;
    for (int64_t i = 0; i < n; i++) ioff[i + 1] += ioff[i];
    std::vector<int32_t> iadj((size_t)std::max<int64_t>(m, 1));
    {
        std::vector<int64_t> fillp(ioff.begin(), ioff.end() - 1);
        for (int64_t v = 0; v < n; v++)
            for (int64_t k = off[v]; k < off[v + 1]; k++) iadj[(size_t)fillp[adj[k]]++] = (int32_t)v;
    }
    const int64_t total = 20 + 16 * n + 8 * m;
    std::vector<char> buf((size_t)total);
    char *p = buf.data();
    auto put = [&](const void *src, size_t bytes) {
        std::memcpy(p, src, bytes);
        p += bytes;
    };
    const int32_t n32 = (int32_t)n;
    put(&total, 8);
    put(&n32, 4);
    put(&m, 8);
    for (int64_t i = 0; i < n; i++) {
        const int32_t hdr[4] = {-1, (int32_t)i, (int32_t)(off[i + 1] - off[i]),
                                (int32_t)(ioff[i + 1] - ioff[i])};
        put(hdr, 16);
        put(adj.data() + off[i], 4 * (size_t)(off[i + 1] - off[i]));
        put(iadj.data() + ioff[i], 4 * (size_t)(ioff[i + 1] - ioff[i]));
    }
    std::ofstream o(path, std::ios::out | std::ios::binary | std::ios::trunc);
    PG_REQUIRE(o.is_open(), PG_ERR_ARG, "Unable to write %s", path);
    o.write(buf.data(), total);
    PG_REQUIRE((bool)o, PG_ERR_ARG, "Unable to write %s", path);
    return PG_OK;
}

extern "C" int pg_graph_destroy(pg_graph *g) {
    if (!g) return PG_OK;
    delete_graph(g);
    return PG_OK;
}

extern "C" int pg_graph_info(const pg_graph *g, int64_t *n, int64_t *m, int64_t *max_in_degree,
                             int32_t *w_dtype) {
    PG_REQUIRE(g, PG_ERR_ARG, "null graph");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (max_in_degree) *max_in_degree = g->max_in_deg;
    if (w_dtype) *w_dtype = g->w_dtype;
    return PG_OK;
}

__global__ void k_download_keys(const int32_t *off, const int32_t *adj, const int32_t *iperm,
                                int64_t n, uint64_t *keys, uint32_t *eidx) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t s = (uint32_t)iperm[v];
        for (int e = off[v]; e < off[v + 1]; e++) {
            keys[e] = (s << 32) | (uint32_t)iperm[adj[e]];
            eidx[e] = (uint32_t)e;
        }
    }
}

extern "C" int pg_graph_download(const pg_graph *gc, int64_t *off, int32_t *adj, void *w) {
    PG_REQUIRE(gc, PG_ERR_ARG, "null graph");
    pg_graph *g = const_cast<pg_graph *>(gc);
    const int64_t n = g->n, m = g->m;
    DevBuf k1, k2, e1, e2, low, wbuf;
    PG_TRY(k1.alloc(8 * (m + 1))); PG_TRY(k2.alloc(8 * (m + 1)));
    PG_TRY(e1.alloc(4 * (m + 1))); PG_TRY(e2.alloc(4 * (m + 1)));
    PG_LAUNCH(k_download_keys, grid_for(n), g->out_off, g->out_adj, g->iperm, n,
              k1.as<uint64_t>(), e1.as<uint32_t>());
    PG_TRY(sort_pairs(k1.as<uint64_t>(), k2.as<uint64_t>(), e1.as<uint32_t>(), e2.as<uint32_t>(), m,
                      32 + bits_for(n + 1)));
    std::vector<uint64_t> keys((size_t)m);
    PG_CHECK_HIP(hipMemcpyAsync(keys.data(), k2.p, 8 * m, hipMemcpyDeviceToHost, g_stream));
    if (w && g->w_dtype != PG_W_NONE) {
        PG_TRY(wbuf.alloc(4 * (m + 1)));
        if (g->w_dtype == PG_W_I32)
            PG_LAUNCH(k_gather<int32_t>, grid_for(m), g->out_w_i32, (const int32_t *)e2.p, m, wbuf.as<int32_t>());
        else
            PG_LAUNCH(k_gather<float>, grid_for(m), g->out_w_f32, (const int32_t *)e2.p, m, wbuf.as<float>());
        PG_CHECK_HIP(hipMemcpyAsync(w, wbuf.p, 4 * m, hipMemcpyDeviceToHost, g_stream));
    }
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    if (off) {
        for (int64_t i = 0; i <= n; i++) off[i] = 0;
        for (int64_t e = 0; e < m; e++) off[(keys[e] >> 32) + 1]++;
        for (int64_t i = 0; i < n; i++) off[i + 1] += off[i];
    }
    if (adj)
        for (int64_t e = 0; e < m; e++) adj[e] = (int32_t)(uint32_t)keys[e];
    return PG_OK;
}

// off[v] = first i with (keys[i] >> 32) >= v, for v in [0, n]; keys sorted
__global__ void k_offsets_from_sorted(const uint64_t *keys, int64_t m, int64_t n, int32_t *off) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t hi = i < m ? (int64_t)(keys[i] >> 32) : n;
        const int64_t lo = i > 0 ? (int64_t)(keys[i - 1] >> 32) + 1 : 0;
        for (int64_t v = lo; v <= hi; v++) off[v] = (int32_t)i;
    }
}

__global__ void k_swap_halves(const uint64_t *keys, int64_t m, uint64_t *out, uint32_t *idx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        out[i] = (k << 32) | (k >> 32);
        idx[i] = (uint32_t)i;
    }
}

extern "C" int pg_graph_view(pg_graph *g, pg_view *v) {
    PG_REQUIRE(g && v, PG_ERR_ARG, "null argument");
    const int64_t n = g->n, m = g->m;
    const size_t wsz = 4;
    if (!g->view_built) {
        DevBuf k1, k2, e1, e2;
        PG_TRY(k1.alloc(8 * (m + 1)));
        PG_TRY(k2.alloc(8 * (m + 1)));
        PG_TRY(e1.alloc(4 * (m + 1)));
        PG_TRY(e2.alloc(4 * (m + 1)));
        // CSR: (src, dst) in input ids, sorted; e2 = internal CSR edge of each entry
        PG_LAUNCH(k_download_keys, grid_for(n), g->out_off, g->out_adj, g->iperm, n,
                  k1.as<uint64_t>(), e1.as<uint32_t>());
        PG_TRY(sort_pairs(k1.as<uint64_t>(), k2.as<uint64_t>(), e1.as<uint32_t>(), e2.as<uint32_t>(),
                          m, 32 + bits_for(n + 1)));
        PG_TRY(galloc_t(g, &g->v_out_off, n + 1));
        PG_TRY(galloc_t(g, &g->v_out_adj, m + 1));
        PG_LAUNCH(k_offsets_from_sorted, grid_for(m + 1), k2.as<uint64_t>(), m, n, g->v_out_off);
        PG_LAUNCH(k_low32, grid_for(m), k2.as<uint64_t>(), m, g->v_out_adj);
        if (g->w_dtype != PG_W_NONE) {
            PG_TRY(galloc(g, &g->v_out_w, wsz * (m + 1)));
            if (g->w_dtype == PG_W_I32)
                PG_LAUNCH(k_gather<int32_t>, grid_for(m), g->out_w_i32, (const int32_t *)e2.p, m,
                          static_cast<int32_t *>(g->v_out_w));
            else
                PG_LAUNCH(k_gather<float>, grid_for(m), g->out_w_f32, (const int32_t *)e2.p, m,
                          static_cast<float *>(g->v_out_w));
        }
        if (g->sym) {
            g->v_in_off = g->v_out_off;
            g->v_in_adj = g->v_out_adj;
            g->v_in_w = g->v_out_w;
        } else {
            // CSC: (dst, src) sorted; e2 = position of the entry in the view's CSR
            PG_LAUNCH(k_swap_halves, grid_for(m), k2.as<uint64_t>(), m, k1.as<uint64_t>(),
                      e1.as<uint32_t>());
            PG_TRY(sort_pairs(k1.as<uint64_t>(), k2.as<uint64_t>(), e1.as<uint32_t>(),
                              e2.as<uint32_t>(), m, 32 + bits_for(n + 1)));
            PG_TRY(galloc_t(g, &g->v_in_off, n + 1));
            PG_TRY(galloc_t(g, &g->v_in_adj, m + 1));
            PG_LAUNCH(k_offsets_from_sorted, grid_for(m + 1), k2.as<uint64_t>(), m, n, g->v_in_off);
            PG_LAUNCH(k_low32, grid_for(m), k2.as<uint64_t>(), m, g->v_in_adj);
            if (g->w_dtype != PG_W_NONE) {
                PG_TRY(galloc(g, &g->v_in_w, wsz * (m + 1)));
                if (g->w_dtype == PG_W_I32)
                    PG_LAUNCH(k_gather<int32_t>, grid_for(m), static_cast<const int32_t *>(g->v_out_w),
                              (const int32_t *)e2.p, m, static_cast<int32_t *>(g->v_in_w));
                else
                    PG_LAUNCH(k_gather<float>, grid_for(m), static_cast<const float *>(g->v_out_w),
                              (const int32_t *)e2.p, m, static_cast<float *>(g->v_in_w));
            }
        }
        PG_CHECK_HIP(hipStreamSynchronize(g_stream));
        g->view_built = true;
    }
    v->n = n;
    v->m = m;
    v->symmetric = g->sym ? 1 : 0;
    v->w_dtype = g->w_dtype;
    v->out_off = g->v_out_off;
    v->out_adj = g->v_out_adj;
    v->in_off = g->v_in_off;
    v->in_adj = g->v_in_adj;
    v->out_w = g->v_out_w;
    v->in_w = g->v_in_w;
    v->stream = (void *)g_stream;
    return PG_OK;
}

extern "C" int pg_graph_set_schedule(pg_graph *g, int items_per_lane, int64_t block_threshold,
                                     int64_t multi_threshold, int64_t chunk) {
    PG_REQUIRE(g, PG_ERR_ARG, "null graph");
    for (PullSchedule *s : {&g->in_sched, &g->out_sched}) {
        if (items_per_lane > 0) s->params.items_per_lane = items_per_lane;
        if (block_threshold > 0) s->params.block_threshold = block_threshold;
        if (multi_threshold > 0) s->params.multi_threshold = multi_threshold;
        if (chunk > 0) s->params.chunk = chunk;
    }
    PG_TRY(build_schedule(g, true));
    if (g->out_sched.built) PG_TRY(build_schedule(g, false));
    return PG_OK;
}

extern "C" int pg_graph_set_tiles(pg_graph *g, int enable, int64_t max_degree, int64_t tile_edges,
                                  int64_t cold_from) {
    PG_REQUIRE(g, PG_ERR_ARG, "null graph");
    PG_REQUIRE(max_degree >= 0, PG_ERR_ARG, "max_degree out of range: %lld", (long long)max_degree);
    PG_REQUIRE(tile_edges >= 0 && tile_edges <= (int64_t)1 << 30, PG_ERR_ARG,
               "tile_edges out of range: %lld", (long long)tile_edges);
    PG_REQUIRE(cold_from >= -1, PG_ERR_ARG, "cold_from out of range: %lld", (long long)cold_from);
    drop_captures(g);
    if (enable >= 0) g->use_tiles = enable != 0;
    if (tile_edges > 0) g->tile_edges = tile_edges;
    if (cold_from != 0) g->tile_cold = cold_from;
    if (max_degree > 0 && max_degree != g->tile_max_degree) {
        g->tile_max_degree = max_degree;
        PG_TRY(build_schedule(g, true));  // the hot / light split moves
    }
    return PG_OK;
}

extern "C" int pg_graph_tile_stats(const pg_graph *g, int64_t *tiles, int64_t *parts,
                                   int64_t *tiled_edges, int64_t *cold_edges,
                                   int64_t *escape_groups, int32_t *cold_overlap) {
    PG_REQUIRE(g, PG_ERR_ARG, "null graph");
    const TileSet &t = g->tiles;
    if (tiles) *tiles = t.built ? t.ntiles : 0;
    if (parts) *parts = t.built ? t.args.nparts : 0;
    if (tiled_edges) *tiled_edges = t.built ? t.m_tiled : 0;
    if (cold_edges) *cold_edges = t.built ? t.m_cold : 0;
    if (escape_groups) *escape_groups = t.built ? t.n_esc : 0;
    if (cold_overlap) *cold_overlap = (t.built && t.cold_overlap && t.cold.nparts > 0) ? 1 : 0;
    return PG_OK;
}

extern "C" int pg_graph_get_schedule(const pg_graph *g, int32_t *items_per_lane,
                                     int64_t *block_threshold, int64_t *multi_threshold,
                                     int64_t *chunk, int32_t *tiles, int64_t *max_degree,
                                     int64_t *tile_edges, int64_t *cold_from) {
    PG_REQUIRE(g, PG_ERR_ARG, "null graph");
    const PullParams &p = g->in_sched.params;
    if (items_per_lane) *items_per_lane = p.items_per_lane;
    if (block_threshold) *block_threshold = p.block_threshold;
    if (multi_threshold) *multi_threshold = p.multi_threshold;
    if (chunk) *chunk = p.chunk;
    if (tiles) *tiles = g->use_tiles ? 1 : 0;
    if (max_degree) *max_degree = g->tile_max_degree;
    if (tile_edges) *tile_edges = g->tile_edges;
    if (cold_from) *cold_from = g->tile_cold;
    return PG_OK;
}

extern "C" int pg_graph_bp_potentials(pg_graph *g, uint64_t seed, float lo, float hi) {
    PG_REQUIRE(g, PG_ERR_ARG, "null graph");
    if (!g->bp_pot) PG_TRY(galloc_t(g, &g->bp_pot, 4 * (size_t)g->m));
    PG_LAUNCH(k_bp_pot, grid_for(g->n), g->in_off, g->in_adj, g->iperm, g->n, seed, lo, hi, g->bp_pot);
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

// ---- subsets -------------------------------------------------------------------------
extern "C" int pg_subset_create(pg_graph *g, pg_subset **out) {
    PG_REQUIRE(g && out, PG_ERR_ARG, "null argument");
    pg_subset *s = new pg_subset();
    s->g = g;
    const int64_t nw = (g->n + 31) / 32;
    hipError_t e1 = hipMalloc(reinterpret_cast<void **>(&s->bits), 4 * (nw + 1));
    hipError_t e2 = hipMalloc(reinterpret_cast<void **>(&s->list), 4 * (g->n + 1));
    hipError_t e3 = hipMalloc(reinterpret_cast<void **>(&s->counters), 64);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        if (s->bits) (void)hipFree(s->bits);
        if (s->list) (void)hipFree(s->list);
        if (s->counters) (void)hipFree(s->counters);
        delete s;
        set_error("subset allocation failed");
        return PG_ERR_OOM;
    }
    *out = s;
    return pg_subset_clear(s);
}

extern "C" int pg_subset_destroy(pg_subset *s) {
    if (!s) return PG_OK;
    (void)hipStreamSynchronize(g_stream);
    (void)hipFree(s->bits);
    (void)hipFree(s->list);
    (void)hipFree(s->counters);
    delete s;
    return PG_OK;
}

extern "C" int pg_subset_clear(pg_subset *s) {
    PG_REQUIRE(s, PG_ERR_ARG, "null subset");
    PG_CHECK_HIP(hipMemsetAsync(s->bits, 0, 4 * ((s->g->n + 31) / 32 + 1), g_stream));
    PG_CHECK_HIP(hipMemsetAsync(s->counters, 0, 64, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

static int subset_finish(pg_subset *s) {
    // list + |F| from the bitmap, then sum of out-degrees (vertexCounter)
    PG_CHECK_HIP(hipMemsetAsync(s->counters, 0, 64, g_stream));
    PG_TRY(bits_to_list(s->g, s->bits, s->list, &s->counters[0]));
    PG_TRY(list_degree_sum(s->g, s->list, &s->counters[0], s->g->out_off, &s->counters[1]));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    return PG_OK;
}

extern "C" int pg_subset_fill(pg_subset *s) {
    PG_REQUIRE(s, PG_ERR_ARG, "null subset");
    PG_TRY(fill_bits_all(s->g, s->bits));
    return subset_finish(s);
}

extern "C" int pg_subset_from_dense(pg_subset *s, const uint8_t *flags) {
    PG_REQUIRE(s && flags, PG_ERR_ARG, "null argument");
    pg_graph *g = s->g;
    uint8_t *tmp = nullptr;
    PG_TRY(workspace_t(g, "subset_in", (size_t)g->n, &tmp));
    PG_CHECK_HIP(hipMemcpyAsync(tmp, flags, g->n, hipMemcpyHostToDevice, g_stream));
    PG_CHECK_HIP(hipMemsetAsync(s->bits, 0, 4 * ((g->n + 31) / 32 + 1), g_stream));
    PG_LAUNCH(k_bits_from_bools_input, grid_for(g->n), tmp, g->perm, g->n, s->bits);
    return subset_finish(s);
}

extern "C" int pg_subset_from_sparse(pg_subset *s, const int32_t *ids, int64_t k) {
    PG_REQUIRE(s && (ids || k == 0) && k >= 0, PG_ERR_ARG, "bad argument");
    pg_graph *g = s->g;
    for (int64_t i = 0; i < k; i++)
        PG_REQUIRE(ids[i] >= 0 && ids[i] < g->n, PG_ERR_ARG, "id %d out of range", ids[i]);
    int32_t *tmp = nullptr;
    PG_TRY(workspace_t(g, "subset_ids", (size_t)k + 1, &tmp));
    if (k) PG_CHECK_HIP(hipMemcpyAsync(tmp, ids, 4 * k, hipMemcpyHostToDevice, g_stream));
    PG_CHECK_HIP(hipMemsetAsync(s->bits, 0, 4 * ((g->n + 31) / 32 + 1), g_stream));
    if (k) PG_LAUNCH(k_bits_from_ids_input, grid_for(k), tmp, k, g->perm, s->bits);
    return subset_finish(s);
}

extern "C" int pg_subset_to_dense(pg_subset *s, uint8_t *flags) {
    PG_REQUIRE(s && flags, PG_ERR_ARG, "null argument");
    pg_graph *g = s->g;
    const int64_t nw = (g->n + 31) / 32;
    uint32_t *ib = nullptr;
    PG_TRY(workspace_t(g, "subset_ibits", (size_t)nw + 1, &ib));
    PG_CHECK_HIP(hipMemsetAsync(ib, 0, 4 * (nw + 1), g_stream));
    PG_LAUNCH(k_bits_to_input_bits, grid_for(g->n), s->bits, g->perm, g->n, ib);
    std::vector<uint32_t> hb((size_t)nw);
    PG_CHECK_HIP(hipMemcpyAsync(hb.data(), ib, 4 * nw, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    for (int64_t u = 0; u < g->n; u++) flags[u] = (hb[u >> 5] >> (u & 31)) & 1u;
    return PG_OK;
}

// toSparse of a subset into caller-owned device memory, allocating nothing: the
// kernels of pg_subset_to_sparse (input-order bitmap, per-block popcounts, one-group
// scan, ordered write) on the file-scope scratch array.  Asynchronous on the given
// stream (null: the library stream).
extern "C" int pg_subset_to_sparse_noalloc(pg_subset *s, int32_t *d_ids, int64_t *d_count,
                                           void *stream) {
    PG_REQUIRE(g_init, PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(s && d_ids && d_count, PG_ERR_ARG, "null argument");
    pg_graph *g = s->g;
    PG_REQUIRE(g->n <= kNoAllocMaxN, PG_ERR_UNSUPPORTED,
               "pg_subset_to_sparse_noalloc: n = %lld exceeds its maximum of %lld vertices",
               (long long)g->n, (long long)kNoAllocMaxN);
    PG_REQUIRE(g_device >= 0 && g_device < kMaxDevices && g_scratch[g_device], PG_ERR_STATE,
               "no scratch address for device %d", g_device);
    uint32_t *ib = g_scratch[g_device];
    uint32_t *bc = ib + kNoAllocWords;
    uint32_t *bo = bc + kNoAllocBlocks;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_stream;
    const int64_t n = g->n, nw = (n + 31) / 32;
    const int64_t nb = (nw + kBlock - 1) / kBlock;
    // the bitmap keeps the last call's contents: cleared with a kernel (no memset node
    // in a captured graph)
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(nw + 1)), dim3(kBlock), 0, st, ib, 0u,
                       nw + 1);
    hipLaunchKernelGGL(k_bits_to_input_bits, dim3(grid_for(n)), dim3(kBlock), 0, st,
                       (const uint32_t *)s->bits, (const int32_t *)g->perm, n, ib);
    if (nb > 0) {
        hipLaunchKernelGGL(k_word_counts, dim3((unsigned)nb), dim3(kBlock), 0, st,
                           (const uint32_t *)ib, nw, bc);
        hipLaunchKernelGGL(k_scan_one_group, dim3(1), dim3(kScanThreads), 0, st,
                           (const uint32_t *)bc, bo, nb, 0);
        hipLaunchKernelGGL(k_word_write, dim3((unsigned)nb), dim3(kBlock), 0, st,
                           (const uint32_t *)ib, nw, n, (const uint32_t *)bo, d_ids,
                           reinterpret_cast<unsigned long long *>(d_count), nb);
    } else {
        hipLaunchKernelGGL(k_fill<int64_t>, dim3(1), dim3(kBlock), 0, st, d_count, (int64_t)0,
                           (int64_t)1);
    }
    PG_CHECK_HIP(hipGetLastError());
    return PG_OK;
}

extern "C" int pg_subset_to_sparse(pg_subset *s, int32_t *ids, int64_t cap, int64_t *k) {
    PG_REQUIRE(s && k, PG_ERR_ARG, "null argument");
    pg_graph *g = s->g;
    const int64_t nw = (g->n + 31) / 32;
    uint32_t *ib = nullptr;
    int32_t *il = nullptr;
    unsigned long long *cnt = nullptr;
    PG_TRY(workspace_t(g, "subset_ibits", (size_t)nw + 1, &ib));
    PG_TRY(workspace_t(g, "subset_ilist", (size_t)g->n + 1, &il));
    PG_TRY(workspace_t(g, "subset_icnt", 2, &cnt));
    PG_CHECK_HIP(hipMemsetAsync(ib, 0, 4 * (nw + 1), g_stream));
    PG_LAUNCH(k_bits_to_input_bits, grid_for(g->n), s->bits, g->perm, g->n, ib);
    PG_TRY(bits_to_list(g, ib, il, cnt));
    unsigned long long c = 0;
    PG_CHECK_HIP(hipMemcpyAsync(&c, cnt, 8, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    *k = (int64_t)c;
    PG_REQUIRE(ids || c == 0 || cap == 0, PG_ERR_ARG, "null ids");
    const int64_t ncopy = std::min<int64_t>((int64_t)c, cap);
    if (ncopy > 0) {
        PG_CHECK_HIP(hipMemcpy(ids, il, 4 * ncopy, hipMemcpyDeviceToHost));
    }
    return PG_OK;
}

extern "C" int pg_subset_size(pg_subset *s, int64_t *count, int64_t *out_edges) {
    PG_REQUIRE(s, PG_ERR_ARG, "null subset");
    unsigned long long c[2];
    PG_CHECK_HIP(hipMemcpyAsync(c, s->counters, 16, hipMemcpyDeviceToHost, g_stream));
    PG_CHECK_HIP(hipStreamSynchronize(g_stream));
    if (count) *count = (int64_t)c[0];
    if (out_edges) *out_edges = (int64_t)c[1];
    return PG_OK;
}
