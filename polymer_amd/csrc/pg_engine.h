// pg_engine.h -- the device side of the edgeMap engine for CDNA4 (wave64).
//
//  * k_pull<Op>: dense pull (edgeMapDense / edgeMapDenseReduce, polymer.h:
//    924-968, 1097-1168) over a degree-binned schedule: G lanes per vertex
//    (G = 1..64) for light vertices, a 256-thread block per heavy vertex and
//    several blocks (chunked, finalised by the last block to arrive, in chunk
//    order) for the heaviest.  Neighbour lists are read coalesced by the lanes
//    of a group; results are reduced with cross-lane shuffles.
//  * k_push<Op>: push over a frontier (edgeMapSparseV3 polymer.h:1928-2003,
//    edgeMapDenseForward polymer.h:970-1032): each block takes 256 sources,
//    scans their out-degrees in LDS and spreads the block's edges evenly over
//    its threads (one binary search in LDS per edge).
//  * frontier helpers: bitmap test/set, wave-aggregated list append.
//
// An Op for k_push provides
//     bool src_active(int s);     void push(int s, int d, int e);
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "pg_internal.h"

namespace pg {

// ---- lane-level helpers ------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ float shfl_xor_t(float v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ int shfl_xor_t(int v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ unsigned shfl_xor_t(unsigned v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ unsigned long long shfl_xor_t(unsigned long long v, int m) {
    return __shfl_xor(v, m, 64);
}
__device__ __forceinline__ double shfl_xor_t(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ float2 shfl_xor_t(float2 v, int m) {
    return make_float2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}
__device__ __forceinline__ float4 shfl_xor_t(float4 v, int m) {
    return make_float4(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64), __shfl_xor(v.z, m, 64),
                       __shfl_xor(v.w, m, 64));
}

template <int G, class Op, class Acc>
__device__ __forceinline__ Acc group_reduce(const Op &op, Acc a) {
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) a = op.combine(a, shfl_xor_t(a, o));
    return a;
}

template <int G>
__device__ __forceinline__ bool group_any(bool p) {
    unsigned long long b = __ballot(p);
    if constexpr (G == 64) {
        return b != 0ull;
    } else {
        const int base = lane_id() & ~(G - 1);
        const unsigned long long mask = (1ull << G) - 1ull;
        return ((b >> base) & mask) != 0ull;
    }
}

// atomics that bypass the non-coherent vector L1 (used for cross-block data)
template <class T>
__device__ __forceinline__ T load_agent(const T *p) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4/8-byte loads only");
    if constexpr (sizeof(T) == 4) {
        unsigned u = __hip_atomic_load(reinterpret_cast<const unsigned *>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(T, u);
    } else {
        unsigned long long u =
            __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(T, u);
    }
}

template <class T>
__device__ __forceinline__ void store_agent(T *p, T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4/8-byte stores only");
    if constexpr (sizeof(T) == 4) {
        __hip_atomic_store(reinterpret_cast<unsigned *>(p), __builtin_bit_cast(unsigned, v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(reinterpret_cast<unsigned long long *>(p),
                           __builtin_bit_cast(unsigned long long, v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- frontier helpers ------------------------------------------------------------
__device__ __forceinline__ bool bit_test(const uint32_t *bits, int v) {
    return (bits[v >> 5] >> (v & 31)) & 1u;
}
// returns true if this call set the bit
__device__ __forceinline__ bool bit_set_new(uint32_t *bits, int v) {
    const uint32_t m = 1u << (v & 31);
    return (atomicOr(&bits[v >> 5], m) & m) == 0u;
}

// Wave-aggregated append: the lanes with pred write val to list; one atomic
// per wave on *counter.  Works under divergence: the ballot only sees the
// lanes that reach it, and the base is read from the (active) leader lane.
__device__ __forceinline__ void agg_append(int32_t *list, unsigned long long *counter, bool pred,
                                           int32_t val) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const int lane = lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred) {
        const unsigned long long below = mask & ((1ull << lane) - 1ull);
        list[base + __popcll(below)] = val;
    }
}

// ---- pull engine ------------------------------------------------------------------
// An Op for k_pull provides
//     using Acc;  using Ctx;  using Stat;
//     static constexpr bool kEarlyExit, kStats;
//     bool active(int v);               Ctx ctx(int v);
//     Acc identity();                   Acc combine(Acc, Acc);
//     Acc edge(int s, int e, const Ctx&);  bool found(Acc, const Ctx&);
//     Stat finish(int v, Acc, const Ctx&);  void skip(int v);
// finish() and skip() run once per vertex (one lane).  When kStats, the Stat
// values are combined per block with Op::stat_reduce and published once per
// block with Op::stat_commit(stat_out, value) (e.g. an atomic add or max).
template <class T>
__device__ __forceinline__ void store_partial(T *p, const T &v) {
    if constexpr (sizeof(T) <= 8 && (sizeof(T) == 4 || sizeof(T) == 8)) {
        store_agent(p, v);
    } else {
        static_assert(sizeof(T) % 4 == 0, "partials are whole words");
        const uint32_t *src = reinterpret_cast<const uint32_t *>(&v);
        uint32_t *dst = reinterpret_cast<uint32_t *>(p);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 4); i++) store_agent(&dst[i], src[i]);
    }
}

template <class T>
__device__ __forceinline__ T load_partial(const T *p) {
    if constexpr (sizeof(T) == 4 || sizeof(T) == 8) {
        return load_agent(p);
    } else {
        T v;
        uint32_t *dst = reinterpret_cast<uint32_t *>(&v);
        const uint32_t *src = reinterpret_cast<const uint32_t *>(p);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 4); i++) dst[i] = load_agent(&src[i]);
        return v;
    }
}

// Neighbour ids are read once per launch: non-temporal loads keep them from
// evicting the gathered per-vertex values from L1/L2.
__device__ __forceinline__ int ld_stream(const int32_t *p) { return __builtin_nontemporal_load(p); }

// An Op may replace the neighbour id with a per-edge token of its own: with
// `static constexpr bool kEdgeToken = true` and `int edge_token(int e)`, the
// engine passes edge_token(e) to edge() in place of adj[e] (and never reads adj).
template <class Op, class = void>
struct has_edge_token : std::false_type {};
template <class Op>
struct has_edge_token<Op, std::enable_if_t<Op::kEdgeToken>> : std::true_type {};

// An Op may gate whole launches: with `static constexpr bool kGated = true` and
// `bool gated()` (the same for every thread, read from device memory), a k_pull
// launch whose gate is set returns at once (iterations enqueued past a
// convergence found on the device).
template <class Op, class = void>
struct has_gate : std::false_type {};
template <class Op>
struct has_gate<Op, std::enable_if_t<Op::kGated>> : std::true_type {};

template <class Op>
__device__ __forceinline__ int edge_src(const Op &op, const int32_t *adj, int e) {
    if constexpr (has_edge_token<Op>::value) return op.edge_token(e);
    else return ld_stream(adj + e);
}

// acc combined with op.edge over e, e+S, e+2S, ... < end; eight loads in flight
template <int S, class Op>
__device__ __forceinline__ typename Op::Acc edge_range(const Op &op, const int32_t *__restrict__ adj,
                                                       int e, int end, const typename Op::Ctx &c,
                                                       typename Op::Acc acc) {
    using Acc = typename Op::Acc;
    for (; e + 7 * S < end; e += 8 * S) {
        int s[8];
#pragma unroll
        for (int k = 0; k < 8; k++) s[k] = edge_src(op, adj, e + k * S);
        Acc a[8];
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = op.edge(s[k], e + k * S, c);
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] = op.combine(a[k], a[k + 4]);
        a[0] = op.combine(a[0], a[2]);
        a[1] = op.combine(a[1], a[3]);
        acc = op.combine(acc, op.combine(a[0], a[1]));
    }
    if (e + 3 * S < end) {
        const int s0 = edge_src(op, adj, e), s1 = edge_src(op, adj, e + S),
                  s2 = edge_src(op, adj, e + 2 * S), s3 = edge_src(op, adj, e + 3 * S);
        const Acc a0 = op.edge(s0, e, c), a1 = op.edge(s1, e + S, c), a2 = op.edge(s2, e + 2 * S, c),
                  a3 = op.edge(s3, e + 3 * S, c);
        acc = op.combine(acc, op.combine(op.combine(a0, a1), op.combine(a2, a3)));
        e += 4 * S;
    }
    for (; e < end; e += S) acc = op.combine(acc, op.edge(edge_src(op, adj, e), e, c));
    return acc;
}

template <int G, class Op>
__device__ __forceinline__ typename Op::Stat pull_group(const Op &op, const PullArgs &pa,
                                                        const PullSeg &seg, int rb) {
    using Acc = typename Op::Acc;
    using Stat = typename Op::Stat;
    const int t = threadIdx.x;
    const int64_t pos = (int64_t)seg.v_begin + ((int64_t)rb * kBlock + t) / G;
    if (pos >= seg.v_end) return Stat(0);  // uniform within the group
    const int lane = t & (G - 1);
    const int v = pa.vlist ? pa.vlist[pos] : (int)pos;
    if (!op.active(v)) {
        if (lane == 0) op.skip(v);
        return Stat(0);
    }
    const typename Op::Ctx c = op.ctx(v);
    const int beg = pa.off[v], end = pa.off[v + 1];
    Acc acc = op.identity();
    if constexpr (Op::kEarlyExit) {
        // four chunks of G edges are loaded before each exit test, so four
        // independent (neighbour id, neighbour state) load pairs are in flight;
        // a hit in an earlier chunk still wins the min over edge positions (eight
        // or sixteen chunks per test measured slower: every lane then issues all
        // its loads, and the loads of a G = 1 group each touch 64 cache lines)
        constexpr int U = 4;
        for (int base = beg; base < end; base += U * G) {
            int s[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int e = base + k * G + lane;
                s[k] = e < end ? edge_src(op, pa.adj, e) : -1;
            }
#pragma unroll
            for (int k = 0; k < U; k++)
                if (s[k] >= 0) acc = op.combine(acc, op.edge(s[k], base + k * G + lane, c));
            if (group_any<G>(op.found(acc, c))) break;
        }
    } else {
        acc = edge_range<G>(op, pa.adj, beg + lane, end, c, acc);
    }
    acc = group_reduce<G>(op, acc);
    if (lane == 0) return op.finish(v, acc, c);
    return Stat(0);
}

template <class Op>
__device__ __forceinline__ typename Op::Acc block_reduce(const Op &op, typename Op::Acc a,
                                                         typename Op::Acc *sm) {
    a = group_reduce<64>(op, a);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) sm[w] = a;
    __syncthreads();
    typename Op::Acc r = sm[0];
#pragma unroll
    for (int i = 1; i < kBlock / 64; i++) r = op.combine(r, sm[i]);
    __syncthreads();
    return r;
}

template <class Op>
__device__ __forceinline__ typename Op::Stat pull_block(const Op &op, const PullArgs &pa,
                                                        const PullSeg &seg, int rb) {
    using Acc = typename Op::Acc;
    using Stat = typename Op::Stat;
    __shared__ Acc sm[kBlock / 64];
    __shared__ int sflag;
    const int64_t pos = (int64_t)seg.v_begin + rb;
    const int v = pa.vlist ? pa.vlist[pos] : (int)pos;
    if (!op.active(v)) {  // uniform over the block
        if (threadIdx.x == 0) op.skip(v);
        return Stat(0);
    }
    const typename Op::Ctx c = op.ctx(v);
    const int beg = pa.off[v], end = pa.off[v + 1];
    Acc acc = op.identity();
    if constexpr (Op::kEarlyExit) {
        if (threadIdx.x == 0) sflag = 0;
        __syncthreads();
        for (int base = beg; base < end; base += kBlock) {
            const int e = base + (int)threadIdx.x;
            if (e < end) acc = op.combine(acc, op.edge(edge_src(op, pa.adj, e), e, c));
            if (group_any<64>(op.found(acc, c)) && lane_id() == 0) sflag = 1;
            __syncthreads();
            const int f = sflag;
            __syncthreads();
            if (f) break;
        }
    } else {
        acc = edge_range<kBlock>(op, pa.adj, beg + (int)threadIdx.x, end, c, acc);
    }
    acc = block_reduce(op, acc, sm);
    if (threadIdx.x == 0) return op.finish(v, acc, c);
    return Stat(0);
}

template <class Op>
__device__ __forceinline__ typename Op::Stat pull_multi(const Op &op, const PullArgs &pa,
                                                        const PullSeg &seg, int rb) {
    using Acc = typename Op::Acc;
    using Stat = typename Op::Stat;
    static_assert(sizeof(Acc) <= 32, "multi-block partials hold 32 bytes");
    __shared__ Acc sm[kBlock / 64];
    __shared__ int slast;
    const MultiItem it = pa.multi[rb];
    const int v = pa.vlist ? pa.vlist[it.pos] : it.pos;
    if (!op.active(v)) {
        if (threadIdx.x == 0 && it.chunk == 0) op.skip(v);
        return Stat(0);
    }
    const typename Op::Ctx c = op.ctx(v);
    const int vbeg = pa.off[v], vend = pa.off[v + 1];
    const int beg = vbeg + it.chunk * pa.chunk;
    const int end = min(vend, beg + pa.chunk);
    Acc acc = op.identity();
    acc = edge_range<kBlock>(op, pa.adj, beg + (int)threadIdx.x, end, c, acc);
    acc = block_reduce(op, acc, sm);
    Acc *part = reinterpret_cast<Acc *>(pa.partial);
    if (threadIdx.x == 0) {
        store_partial(&part[rb], acc);
        __threadfence();
        const unsigned prev = atomicAdd(&pa.tickets[it.slot], 1u);
        slast = (prev == (unsigned)(it.nchunks - 1));
    }
    __syncthreads();
    Stat st = Stat(0);
    if (slast && threadIdx.x == 0) {
        __threadfence();
        Acc tot = op.identity();
        for (int k = 0; k < it.nchunks; k++) tot = op.combine(tot, load_partial(&part[it.first + k]));
        pa.tickets[it.slot] = 0u;  // ready for the next launch
        st = op.finish(v, tot, c);
    }
    return st;
}

// the work of block b of a k_pull launch (uniform over the block)
template <class Op>
__device__ __forceinline__ typename Op::Stat pull_dispatch(const Op &op, const PullArgs &pa, int b) {
    using Stat = typename Op::Stat;
    int si = 0;
    while (si + 1 < pa.nseg && b >= pa.seg[si + 1].block_begin) ++si;
    const PullSeg seg = pa.seg[si];
    const int rb = b - seg.block_begin;
    Stat st;
    switch (seg.g) {
        case kModeMulti: st = pull_multi(op, pa, seg, rb); break;
        case kModeBlock: st = pull_block(op, pa, seg, rb); break;
        case 64: st = pull_group<64>(op, pa, seg, rb); break;
        case 32: st = pull_group<32>(op, pa, seg, rb); break;
        case 16: st = pull_group<16>(op, pa, seg, rb); break;
        case 8: st = pull_group<8>(op, pa, seg, rb); break;
        case 4: st = pull_group<4>(op, pa, seg, rb); break;
        case 2: st = pull_group<2>(op, pa, seg, rb); break;
        default: st = pull_group<1>(op, pa, seg, rb); break;
    }
    return st;
}

// combines every thread's statistic over the block and commits it once
template <class Op>
__device__ __forceinline__ void commit_block_stat(typename Op::Stat st, typename Op::Stat *stat_out) {
    using Stat = typename Op::Stat;
    if constexpr (Op::kStats) {
        __shared__ Stat ssm[kBlock / 64];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) st = Op::stat_reduce(st, shfl_xor_t(st, o));
        if (lane_id() == 0) ssm[threadIdx.x >> 6] = st;
        __syncthreads();
        if (threadIdx.x == 0) {
            Stat tot = ssm[0];
#pragma unroll
            for (int i = 1; i < kBlock / 64; i++) tot = Op::stat_reduce(tot, ssm[i]);
            if (tot != Stat(0)) Op::stat_commit(stat_out, tot);
        }
    }
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_pull(Op op, PullArgs pa,
                                                 typename Op::Stat *stat_out) {
    if constexpr (has_gate<Op>::value) {
        if (op.gated()) return;  // uniform: every thread reads the same word
    }
    commit_block_stat<Op>(pull_dispatch(op, pa, (int)blockIdx.x), stat_out);
}

// body_only: skip the vertices without edges (for an Op whose finish() and skip()
// do nothing to a vertex with no active in-neighbour)
template <class Op>
inline int launch_pull(const Op &op, const PullSchedule &s, typename Op::Stat *stat_out = nullptr,
                       bool body_only = false) {
    const int nb = body_only ? s.nblocks_body : s.nblocks;
    if (nb == 0) return PG_OK;
    hipLaunchKernelGGL(k_pull<Op>, dim3(nb), dim3(kBlock), 0, stream(), op, s.args, stat_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("k_pull launch failed: %s", hipGetErrorString(e));
        return PG_ERR_HIP;
    }
    return PG_OK;
}

// ---- tiled pull ---------------------------------------------------------------------
// The light destinations are pulled tile by tile (TileArgs): a 1024-thread
// workgroup streams its tile's (source, destination) pairs, sorted by source, and
// adds the gathered values into per-destination accumulators in LDS.  An Op used
// here also provides
//     V tile_raw(int s);                          // the value gathered from source s
//     T tile_fix(V x);                            // an edge's term
//     Acc tile_sum(T q);                          // an accumulated sum
//     static constexpr bool kTileWeighted;        // term = tile_raw(s) * w(edge)
// (weights from TileArgs::wt, the graph's edge weights in tile order) with V = float
// and T = unsigned long long (fixed point: integer addition is associative, so the
// result does not depend on the order in which the LDS atomics land), or V = T =
// double (fp64 LDS adds, rounded in arrival order like any parallel fp64 sum; the
// cold values then travel as doubles, TileArgs::cold_val64), or V = T = unsigned
// long long with Op::kTileReduce = kReduceOr (64-bit masks ORed; the cold values as
// 8-byte words), or V = uint32_t and T = unsigned long long with kReduceMin.
// The per-vertex protocol (active / finish / skip, statistics) is the one k_pull
// follows.
template <class Op>
using tile_val_t = decltype(static_cast<const Op *>(nullptr)->tile_raw(0));
template <class Op>
using tile_acc_t = decltype(static_cast<const Op *>(nullptr)->tile_fix(tile_val_t<Op>{}));

__device__ __forceinline__ void lds_add(unsigned long long *p, unsigned long long v) { atomicAdd(p, v); }
__device__ __forceinline__ void lds_add(double *p, double v) { unsafeAtomicAdd(p, v); }

// Op::kTileReduce (optional, 64-bit integer accumulators): kReduceOr ORs the
// values (masks; a zero value is not applied), kReduceMin keeps the smallest (an
// all-ones value is not applied); the partial slots of a tile are combined the same
// way.  Without it the values are added.
constexpr int kReduceAdd = 0, kReduceOr = 1, kReduceMin = 2;
template <class Op, class = void>
struct tile_reduce_trait : std::integral_constant<int, kReduceAdd> {};
template <class Op>
struct tile_reduce_trait<Op, std::void_t<decltype(Op::kTileReduce)>>
    : std::integral_constant<int, Op::kTileReduce> {};
template <class Op>
constexpr int kTileReduceOf = tile_reduce_trait<Op>::value;

template <class Op, class T>
__device__ __forceinline__ void lds_acc(T *p, T v) {
    if constexpr (kTileReduceOf<Op> == kReduceOr) {
        if (v) atomicOr(p, v);
    } else if constexpr (kTileReduceOf<Op> == kReduceMin) {
        if (v != ~T(0)) atomicMin(p, v);
    } else {
        lds_add(p, v);
    }
}
template <class Op, class T>
__device__ __forceinline__ T tile_combine(T a, T b) {
    if constexpr (kTileReduceOf<Op> == kReduceOr) return a | b;
    else if constexpr (kTileReduceOf<Op> == kReduceMin) return a < b ? a : b;
    else return a + b;
}
template <class Op, class T>
__device__ __forceinline__ T tile_identity() {
    if constexpr (kTileReduceOf<Op> == kReduceMin) return ~T(0);
    else return T(0);
}

template <class T>
__device__ __forceinline__ const T *cold_values(const TileArgs &ta) {
    if constexpr (sizeof(T) == 8) return static_cast<const T *>(ta.cold_val64);
    else return reinterpret_cast<const T *>(ta.cold_val);
}
__device__ __forceinline__ unsigned ld_stream_u16(const uint16_t *p) {
    return __builtin_nontemporal_load(p);
}

// x >= 0 (finite, at most 2^(63-F)) as an unsigned fixed-point number with F
// fraction bits, truncated: the mantissa shifted by the exponent, no float->int64
// conversion sequence.  (A branch-free form that forms both shifts measured 3%
// slower in the tile loop: the lanes of a wave nearly always take the same side.)
template <int F>
__host__ __device__ __forceinline__ unsigned long long to_fixed(float x) {
    const unsigned u = __builtin_bit_cast(unsigned, x);
    const int ex = (int)(u >> 23);  // biased exponent (sign bit is 0)
    const unsigned long long mant =
        ex ? (unsigned long long)((u & 0x7FFFFFu) | 0x800000u) : (unsigned long long)(u & 0x7FFFFFu);
    const int sh = (ex ? ex : 1) - 150 + F;  // x = mant * 2^(ex-150)
    if (sh >= 0) return mant << sh;
    return sh > -64 ? mant >> (-sh) : 0ull;
}

// the weight of tile edge i (Op::kTileWeighted: weights streamed in tile order;
// otherwise nothing is loaded) and the fixed-point term of an edge whose source
// value is x: x itself, or x times the edge's weight
template <class Op>
__device__ __forceinline__ float tile_weight(const Op &, const float *wt, int64_t i) {
    if constexpr (Op::kTileWeighted) return __builtin_nontemporal_load(wt + i);
    else return 1.f;
}
// (a weighted Op that reduces by min adds the weight instead: its tile_fix_w(x, w)
// makes the term, as a min-plus relaxation does)
template <class Op>
__device__ __forceinline__ tile_acc_t<Op> tile_term(const Op &op, tile_val_t<Op> x, float wv) {
    if constexpr (Op::kTileWeighted && kTileReduceOf<Op> == kReduceMin) return op.tile_fix_w(x, wv);
    else if constexpr (Op::kTileWeighted) return op.tile_fix(x * wv);
    else return op.tile_fix(x);
}

template <class Op>
__global__ __launch_bounds__(kTileThreads) void k_pull_tiles(Op op, TileArgs ta,
                                                             typename Op::Stat *stat_out) {
    using Stat = typename Op::Stat;
    using Val = tile_val_t<Op>;
    using T = tile_acc_t<Op>;
    static_assert(sizeof(T) == 8, "8-byte accumulators (the partial slots)");
    __shared__ T acc[kTileMaxV];
    __shared__ Stat ssm[kTileThreads / 64];
    const TilePart part = ta.parts[blockIdx.x];
    const int t = part.tile;
    const int vb = ta.vbeg[t];
    const int nv = ta.vbeg[t + 1] - vb;  // <= kTileMaxV by construction
    for (int i = threadIdx.x; i < nv; i += kTileThreads) acc[i] = tile_identity<Op, T>();
    __syncthreads();
    constexpr int U = 4;  // measured on MI355X: 4 edges in flight per lane beat 2, 3, 8, 16
    constexpr unsigned kDstMask = (1u << kPackDstBits) - 1u;
    static_assert(kTileMaxV == (1 << kPackDstBits), "destination offsets fill the low bits");
    {  // warm sources: gathered
        const int64_t e0 = part.e0;
        const int cnt = (int)(part.cold_e - part.e0);
        int e = threadIdx.x;
        // A step of U edges per lane in stages, each issued for all U edges before
        // the next: source ids from the packed words and group bases (an escaped
        // group's ids in a fix-up that is almost never taken, so no per-edge branch
        // separates the gathers); the U gathers, in flight together; the next
        // step's packed words, bases and weights (index clamped into the part), in
        // flight while this step waits for its gathers; terms; LDS adds.  Measured
        // on MI355X (S24 PageRank iteration): 613 us with one gather in flight per
        // lane and no prefetch, 590 with the gathers together, 552 with the
        // prefetch; a prefetch two steps ahead, or U = 3, 6, 8, no better.
        if (e + (U - 1) * kTileThreads < cnt) {
            unsigned w[U];
            int b[U];
            float wv[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int64_t i = e0 + e + k * kTileThreads;
                w[k] = __builtin_nontemporal_load(ta.packed + i);
                b[k] = ta.gbase[i / kPackGroup];
                wv[k] = tile_weight(op, ta.wt, i);
            }
            while (true) {
                int s[U];
                bool esc = false;
#pragma unroll
                for (int k = 0; k < U; k++) {
                    s[k] = b[k] + (int)(w[k] >> kPackDstBits);
                    esc = esc || b[k] < 0;
                }
                if (esc) {
#pragma unroll
                    for (int k = 0; k < U; k++) {
                        const int64_t i = e0 + e + k * kTileThreads;
                        if (b[k] < 0)
                            s[k] = ta.esc[(int64_t)(-b[k] - 1) * kPackGroup + (i % kPackGroup)];
                    }
                }
                Val x[U];
#pragma unroll
                for (int k = 0; k < U; k++) x[k] = op.tile_raw(s[k]);
                const int en = e + U * kTileThreads;
                unsigned wn[U];
                int bn[U];
                float wvn[U];
#pragma unroll
                for (int k = 0; k < U; k++) {
                    const int64_t i = e0 + min(en + k * kTileThreads, cnt - 1);
                    wn[k] = __builtin_nontemporal_load(ta.packed + i);
                    bn[k] = ta.gbase[i / kPackGroup];
                    wvn[k] = tile_weight(op, ta.wt, i);
                }
                T q[U];
#pragma unroll
                for (int k = 0; k < U; k++) q[k] = tile_term(op, x[k], wv[k]);
#pragma unroll
                for (int k = 0; k < U; k++) lds_acc<Op>(&acc[w[k] & kDstMask], q[k]);
                e = en;
                if (!(e + (U - 1) * kTileThreads < cnt)) break;
#pragma unroll
                for (int k = 0; k < U; k++) {
                    w[k] = wn[k];
                    b[k] = bn[k];
                    wv[k] = wvn[k];
                }
            }
        }
        for (; e < cnt; e += kTileThreads) {
            const int64_t i = e0 + e;
            const unsigned w = __builtin_nontemporal_load(ta.packed + i);
            const int b = ta.gbase[i / kPackGroup];
            const int s = b >= 0 ? b + (int)(w >> kPackDstBits)
                                 : ta.esc[(int64_t)(-b - 1) * kPackGroup + (i % kPackGroup)];
            lds_acc<Op>(&acc[w & kDstMask], tile_term(op, op.tile_raw(s), tile_weight(op, ta.wt, i)));
        }
    }
    {  // cold sources: values streamed from the pre-pass
        const Val *__restrict__ val = cold_values<Val>(ta) + part.cold_v;
        const uint32_t *__restrict__ pk = ta.packed + part.cold_e;
        const int cnt = (int)(part.e1 - part.cold_e);
        int e = threadIdx.x;
        for (; e + (U - 1) * kTileThreads < cnt; e += U * kTileThreads) {
            Val x[U];
            float wv[U];
            unsigned d[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                x[k] = __builtin_nontemporal_load(val + e + k * kTileThreads);
                d[k] = __builtin_nontemporal_load(pk + e + k * kTileThreads) & kDstMask;
                wv[k] = tile_weight(op, ta.wt, part.cold_e + e + k * kTileThreads);
            }
#pragma unroll
            for (int k = 0; k < U; k++) lds_acc<Op>(&acc[d[k]], tile_term(op, x[k], wv[k]));
        }
        for (; e < cnt; e += kTileThreads)
            lds_acc<Op>(&acc[__builtin_nontemporal_load(pk + e) & kDstMask],
                    tile_term(op, __builtin_nontemporal_load(val + e),
                              tile_weight(op, ta.wt, part.cold_e + e)));
    }
    __syncthreads();
    if (part.slot >= 0) {
        // one of several parts: store the partial sums; k_tile_finish adds them up
        T *p = reinterpret_cast<T *>(ta.partial) + part.slot;
        for (int i = threadIdx.x; i < nv; i += kTileThreads) p[i] = acc[i];
        return;  // uniform over the block
    }
    Stat st = Stat(0);
    for (int i = threadIdx.x; i < nv; i += kTileThreads) {
        const int v = vb + i;
        if (!op.active(v)) {
            op.skip(v);
            continue;
        }
        const typename Op::Ctx c = op.ctx(v);
        st = Op::stat_reduce(st, op.finish(v, op.tile_sum(acc[i]), c));
    }
    if constexpr (Op::kStats) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) st = Op::stat_reduce(st, shfl_xor_t(st, o));
        if (lane_id() == 0) ssm[threadIdx.x >> 6] = st;
        __syncthreads();
        if (threadIdx.x == 0) {
            Stat tot = ssm[0];
#pragma unroll
            for (int i = 1; i < kTileThreads / 64; i++) tot = Op::stat_reduce(tot, ssm[i]);
            if (tot != Stat(0)) Op::stat_commit(stat_out, tot);
        }
    }
}

// Cold pre-pass: one workgroup per slice of a source chunk's list stages the
// chunk's kColdChunk source values in LDS, then writes each listed edge's value
// to its place in the cold value array (runs of consecutive places per tile).
// An 8-byte value type stages the chunk in two halves of kColdChunk / 2 sources (64
// KiB of LDS each), every half scanning the part's list for its own sources.
template <class Op>
__global__ __launch_bounds__(kTileThreads) void k_cold_gather(Op op, ColdArgs ca) {
    using Val = tile_val_t<Op>;
    constexpr int kHalves = sizeof(Val) == 8 ? 2 : 1;
    constexpr int kStage = kColdChunk / kHalves;
    __shared__ Val sv[kStage];
    Val *__restrict__ out = sizeof(Val) == 8 ? static_cast<Val *>(ca.val64)
                                             : reinterpret_cast<Val *>(ca.val);
    const ColdPart cp = ca.parts[blockIdx.x];
    const int64_t base = (int64_t)ca.cold_begin + (int64_t)cp.chunk * kColdChunk;
    const uint32_t *__restrict__ pos = ca.pos + cp.a;
    const uint16_t *__restrict__ srcl = ca.srcl + cp.a;
    const int cnt = (int)(cp.b - cp.a);
    for (int h = 0; h < kHalves; h++) {
        const int64_t hb = base + (int64_t)h * kStage;
        const int len = (int)max((int64_t)0, min((int64_t)kStage, ca.n - hb));
        if (h > 0) __syncthreads();  // the previous half's readers are done
        for (int j = threadIdx.x; j < len; j += kTileThreads) sv[j] = op.tile_raw((int)(hb + j));
        __syncthreads();
        const unsigned lo = (unsigned)(h * kStage);
        constexpr int U = 8;
        int e = threadIdx.x;
        for (; e + (U - 1) * kTileThreads < cnt; e += U * kTileThreads) {
            uint32_t p[U];
            unsigned l[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                p[k] = __builtin_nontemporal_load(pos + e + k * kTileThreads);
                l[k] = ld_stream_u16(srcl + e + k * kTileThreads);
            }
#pragma unroll
            for (int k = 0; k < U; k++)
                if (kHalves == 1 || l[k] - lo < (unsigned)kStage) out[p[k]] = sv[l[k] - lo];
        }
        for (; e < cnt; e += kTileThreads) {
            const unsigned l = ld_stream_u16(srcl + e);
            if (kHalves == 1 || l - lo < (unsigned)kStage)
                out[__builtin_nontemporal_load(pos + e)] = sv[l - lo];
        }
    }
}

// Vertices of the tiles with several parts: one thread per vertex sums the
// vertex's partial slots (coalesced across the block) and finalises.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_tile_finish(Op op, TileArgs ta,
                                                        typename Op::Stat *stat_out) {
    using Stat = typename Op::Stat;
    __shared__ Stat ssm[kBlock / 64];
    const TileFin f = ta.fin[blockIdx.x];
    const int vb = ta.vbeg[f.tile];
    const int nv = ta.vbeg[f.tile + 1] - vb;
    const int i = f.v0 + (int)threadIdx.x;
    Stat st = Stat(0);
    if (i < nv) {
        const int v = vb + i;
        if (!op.active(v)) {
            op.skip(v);
        } else {
            using T = tile_acc_t<Op>;
            const T *p = reinterpret_cast<const T *>(ta.partial) + f.slot0 + i;
            // eight slots in flight (a hub tile has dozens of parts, and its finishing
            // blocks are the tail of the launch)
            T q[8];
#pragma unroll
            for (int j = 0; j < 8; j++) q[j] = tile_identity<Op, T>();
            int k = 0;
            for (; k + 7 < f.nparts; k += 8) {
                T x[8];
#pragma unroll
                for (int j = 0; j < 8; j++) x[j] = __builtin_nontemporal_load(p + (int64_t)(k + j) * nv);
#pragma unroll
                for (int j = 0; j < 8; j++) q[j] = tile_combine<Op>(q[j], x[j]);
            }
            for (; k < f.nparts; k++)
                q[0] = tile_combine<Op>(q[0], __builtin_nontemporal_load(p + (int64_t)k * nv));
            // integers: any order; doubles: this fixed order for a given part count
            const T qa = tile_combine<Op>(tile_combine<Op>(q[0], q[1]), tile_combine<Op>(q[2], q[3]));
            const T qb = tile_combine<Op>(tile_combine<Op>(q[4], q[5]), tile_combine<Op>(q[6], q[7]));
            const T q8 = tile_combine<Op>(qa, qb);
            st = op.finish(v, op.tile_sum(q8), op.ctx(v));
        }
    }
    if constexpr (Op::kStats) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) st = Op::stat_reduce(st, shfl_xor_t(st, o));
        if (lane_id() == 0) ssm[threadIdx.x >> 6] = st;
        __syncthreads();
        if (threadIdx.x == 0) {
            Stat tot = ssm[0];
#pragma unroll
            for (int w = 1; w < kBlock / 64; w++) tot = Op::stat_reduce(tot, ssm[w]);
            if (tot != Stat(0)) Op::stat_commit(stat_out, tot);
        }
    }
}

// Pull over the CSC with the light range taken from the tiles: the cold pre-pass,
// the hot schedule (multi-block and block-per-vertex segments, zero-degree tail),
// the tile parts, then the vertices of the tiles with several parts.
// skip_tail: leave out the zero-in-degree vertices, for an Op whose results
// there are already in place (PageRank from its third iteration: those vertices
// left the frontier after the first, and both double buffers hold their zeros).
// ahead: the next iteration's op.  When given and the tile set allows it (every
// multi-part tile lies below the first cold source, so the finishing pass writes
// no value the cold pre-pass reads), the next iteration's cold pre-pass is launched
// on the side stream as soon as this iteration's tile parts are done, beside the
// finishing pass; the next call, with cold_ready, waits for it instead of running
// its own.  (The cold value array is read only by the tile parts, which are done.)
// finish_aside: with ahead, the roles of the streams swap: the finishing pass runs
// on the side stream and the next cold pre-pass follows the tile parts on this
// stream, which saves a cross-stream hand-off on the critical path (the cold
// pre-pass is the longer of the two).  Only for a caller that enqueues nothing
// reading this pull's results (values or stat_out) before its next call, whose
// cold_ready wait then covers the finishing pass.
template <class Op>
inline int launch_pull_tiled(const Op &op, const PullSchedule &s, const TileSet &t,
                             typename Op::Stat *stat_out = nullptr, bool skip_tail = false,
                             const Op *ahead = nullptr, bool cold_ready = false,
                             bool finish_aside = false) {
    if (cold_ready) {
        hipError_t e = hipStreamWaitEvent(stream(), t.ev_cold, 0);
        if (e != hipSuccess) {
            set_error("tiled pull: wait for the cold pre-pass failed: %s", hipGetErrorString(e));
            return PG_ERR_HIP;
        }
    } else if (t.cold.nparts > 0) {
        hipLaunchKernelGGL(k_cold_gather<Op>, dim3(t.cold.nparts), dim3(kTileThreads), 0, stream(),
                           op, t.cold);
    }
    const int hot_blocks = skip_tail ? s.nblocks_hot_body : s.nblocks_hot;
    if (hot_blocks > 0)
        hipLaunchKernelGGL(k_pull<Op>, dim3(hot_blocks), dim3(kBlock), 0, stream(), op,
                           s.args_hot, stat_out);
    if (t.args.nparts > 0)
        hipLaunchKernelGGL(k_pull_tiles<Op>, dim3(t.args.nparts), dim3(kTileThreads), 0, stream(),
                           op, t.args, stat_out);
    if (ahead && t.cold_overlap && t.cold.nparts > 0 && finish_aside) {
        hipError_t e = hipEventRecord(t.ev_tiles, stream());
        if (e == hipSuccess) e = hipStreamWaitEvent(side_stream(), t.ev_tiles, 0);
        if (e == hipSuccess && t.args.nfin > 0)
            hipLaunchKernelGGL(k_tile_finish<Op>, dim3(t.args.nfin), dim3(kBlock), 0,
                               side_stream(), op, t.args, stat_out);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(t.ev_cold, side_stream());
        if (e == hipSuccess)
            hipLaunchKernelGGL(k_cold_gather<Op>, dim3(t.cold.nparts), dim3(kTileThreads), 0,
                               stream(), *ahead, t.cold);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("tiled pull: finishing pass aside failed: %s", hipGetErrorString(e));
            return PG_ERR_HIP;
        }
        return PG_OK;
    }
    if (ahead && t.cold_overlap && t.cold.nparts > 0) {
        hipError_t e = hipEventRecord(t.ev_tiles, stream());
        if (e == hipSuccess) e = hipStreamWaitEvent(side_stream(), t.ev_tiles, 0);
        if (e == hipSuccess)
            hipLaunchKernelGGL(k_cold_gather<Op>, dim3(t.cold.nparts), dim3(kTileThreads), 0,
                               side_stream(), *ahead, t.cold);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(t.ev_cold, side_stream());
        if (e != hipSuccess) {
            set_error("tiled pull: cold pre-pass ahead failed: %s", hipGetErrorString(e));
            return PG_ERR_HIP;
        }
    }
    if (t.args.nfin > 0)
        hipLaunchKernelGGL(k_tile_finish<Op>, dim3(t.args.nfin), dim3(kBlock), 0, stream(), op,
                           t.args, stat_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("tiled pull launch failed: %s", hipGetErrorString(e));
        return PG_ERR_HIP;
    }
    return PG_OK;
}

// A pull whose every destination lies in the tiles (a tile set over all positions,
// no hot schedule): the cold pre-pass, the tile parts, the finishing pass.
template <class Op>
inline int launch_tiles(const Op &op, const TileSet &t, typename Op::Stat *stat_out = nullptr) {
    if (t.cold.nparts > 0)
        hipLaunchKernelGGL(k_cold_gather<Op>, dim3(t.cold.nparts), dim3(kTileThreads), 0, stream(),
                           op, t.cold);
    if (t.args.nparts > 0)
        hipLaunchKernelGGL(k_pull_tiles<Op>, dim3(t.args.nparts), dim3(kTileThreads), 0, stream(),
                           op, t.args, stat_out);
    if (t.args.nfin > 0)
        hipLaunchKernelGGL(k_tile_finish<Op>, dim3(t.args.nfin), dim3(kBlock), 0, stream(), op,
                           t.args, stat_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("tile pull launch failed: %s", hipGetErrorString(e));
        return PG_ERR_HIP;
    }
    return PG_OK;
}

// ---- push engine ------------------------------------------------------------------
// kDense: sources are the vertex ids [0, count) filtered by op.src_active;
// otherwise the sources are list[0, count).
template <class Op, bool kDense>
__global__ __launch_bounds__(kBlock) void k_push(Op op, const int32_t *__restrict__ off,
                                                 const int32_t *__restrict__ adj,
                                                 const int32_t *__restrict__ list, int64_t count) {
    __shared__ int32_t s_src[kBlock];
    __shared__ int32_t s_beg[kBlock];
    __shared__ uint32_t s_scan[kBlock + 1];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + t;
    int32_t s = -1;
    uint32_t deg = 0;
    int32_t beg = 0;
    if (i < count) {
        s = kDense ? (int32_t)i : list[i];
        if (op.src_active(s)) {
            beg = off[s];
            deg = (uint32_t)(off[s + 1] - beg);
        }
    }
    s_src[t] = s;
    s_beg[t] = beg;
    // block-wide inclusive scan of deg (Hillis-Steele over LDS)
    s_scan[t + 1] = deg;
    if (t == 0) s_scan[0] = 0;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        uint32_t x = (t + 1 - o >= 1) ? s_scan[t + 1 - o] : 0u;
        __syncthreads();
        s_scan[t + 1] += x;
        __syncthreads();
    }
    const uint32_t total = s_scan[kBlock];
    for (uint32_t k = t; k < total; k += kBlock) {
        // find j with s_scan[j] <= k < s_scan[j+1]
        int lo = 0, hi = kBlock - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_scan[mid] <= k) lo = mid; else hi = mid - 1;
        }
        const int src = s_src[lo];
        const int e = s_beg[lo] + (int)(k - s_scan[lo]);
        op.push(src, adj[e], e);
    }
}

template <class Op, bool kDense>
inline int launch_push(const Op &op, const int32_t *off, const int32_t *adj, const int32_t *list,
                       int64_t count) {
    if (count <= 0) return PG_OK;
    const int64_t nb = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((k_push<Op, kDense>), dim3((unsigned)nb), dim3(kBlock), 0, stream(), op,
                       off, adj, list, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("k_push launch failed: %s", hipGetErrorString(e));
        return PG_ERR_HIP;
    }
    return PG_OK;
}

// Device-wide edge-balanced push: scan[j] is the exclusive prefix sum of the
// out-degrees of list[0..count) (scan[count] = total).  Block b takes the edges
// [b*epb, (b+1)*epb) of the frontier, whatever vertices they belong to, so a hub's
// edges are spread over as many blocks as they fill.  epb is kPushEdges for large
// frontiers and shrinks (to one edge per lane) until the grid has kPushMinBlocks
// blocks: a push is a chain of dependent accesses per edge, and a few hundred
// blocks of eight-edge chains leave most of the chip idle.  The block's window of
// list entries is staged in LDS; the rare window larger than LDS falls back to
// binary searches in global memory.
constexpr int kPushEdges = 2048;
constexpr int kPushWin = 2048;
constexpr int64_t kPushMinBlocks = 2048;

__device__ __forceinline__ int64_t upper_bound_u32(const uint32_t *a, int64_t n, uint32_t x) {
    int64_t lo = 0, hi = n;  // first index with a[i] > x
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_push_global(Op op, const int32_t *__restrict__ off,
                                                        const int32_t *__restrict__ adj,
                                                        const int32_t *__restrict__ list,
                                                        const uint32_t *__restrict__ scan,
                                                        int64_t count, uint32_t total,
                                                        uint32_t epb) {
    __shared__ uint32_t s_scan[kPushWin + 1];
    __shared__ int32_t s_beg[kPushWin];
    __shared__ int64_t s_j0, s_j1;
    const uint32_t E0 = (uint32_t)blockIdx.x * epb;
    const uint32_t E1 = min(total, E0 + epb);
    if (threadIdx.x == 0) {
        s_j0 = upper_bound_u32(scan, count + 1, E0) - 1;
        s_j1 = upper_bound_u32(scan, count + 1, E1 - 1) - 1;
    }
    __syncthreads();
    const int64_t j0 = s_j0, nwin = s_j1 - s_j0 + 1;
    if (nwin <= kPushWin) {
        for (int64_t i = threadIdx.x; i < nwin; i += kBlock) {
            s_scan[i] = scan[j0 + i];
            s_beg[i] = off[list[j0 + i]];
        }
        __syncthreads();
        for (uint32_t k = E0 + threadIdx.x; k < E1; k += kBlock) {
            int lo = 0, hi = (int)nwin - 1;  // last i with s_scan[i] <= k
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_scan[mid] <= k) lo = mid; else hi = mid - 1;
            }
            const int e = s_beg[lo] + (int)(k - s_scan[lo]);
            op.push(list[j0 + lo], adj[e], e);
        }
    } else {
        for (uint32_t k = E0 + threadIdx.x; k < E1; k += kBlock) {
            const int64_t j = upper_bound_u32(scan, count + 1, k) - 1;
            const int s = list[j];
            const int e = off[s] + (int)(k - scan[j]);
            op.push(s, adj[e], e);
        }
    }
}

// deg[i] = out-degree of list[i] for i < count, deg[count] = 0
static __global__ void k_list_degrees(const int32_t *list, int64_t count, const int32_t *off,
                                      uint32_t *deg) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= count;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t d = 0;
        if (i < count) {
            const int v = list[i];
            d = (uint32_t)(off[v + 1] - off[v]);
        }
        deg[i] = d;
    }
}

// Push over list[0..count) whose out-degrees sum to total (host values).
template <class Op>
inline int launch_push_global(const Op &op, pg_graph *g, const int32_t *off, const int32_t *adj,
                              const int32_t *list, int64_t count, int64_t total) {
    if (count <= 0 || total <= 0) return PG_OK;
    if (total >= 0xFFFFFFFFll) {
        set_error("push: frontier edge count exceeds 32 bits");
        return PG_ERR_UNSUPPORTED;
    }
    // sized for the largest list once: growing a workspace frees and allocates
    uint32_t *deg = nullptr, *scan = nullptr;
    PG_TRY(workspace_t(g, "push_deg", (size_t)g->n + 1, &deg));
    PG_TRY(workspace_t(g, "push_scan", (size_t)g->n + 1, &scan));
    const int64_t dgrid = std::min<int64_t>((count + kBlock) / kBlock, 65536);
    hipLaunchKernelGGL(k_list_degrees, dim3((unsigned)dgrid), dim3(kBlock), 0, stream(), list,
                       count, off, deg);
    PG_TRY(exclusive_scan_u32(g, deg, scan, count + 1));
    const int64_t epb = std::min<int64_t>(
        kPushEdges, std::max<int64_t>(kBlock, (total + kPushMinBlocks - 1) / kPushMinBlocks));
    const int64_t nb = (total + epb - 1) / epb;
    hipLaunchKernelGGL(k_push_global<Op>, dim3((unsigned)nb), dim3(kBlock), 0, stream(), op, off,
                       adj, list, scan, count, (uint32_t)total, (uint32_t)epb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("k_push_global launch failed: %s", hipGetErrorString(e));
        return PG_ERR_HIP;
    }
    return PG_OK;
}

// ---- small element-wise kernels ------------------------------------------------------
template <class T>
__global__ void k_fill(T *p, T v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <class T>
inline int fill(T *p, T v, int64_t n) {
    if (n <= 0) return PG_OK;
    int64_t nb = (n + kBlock - 1) / kBlock;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)nb), dim3(kBlock), 0, stream(), p, v, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("fill launch failed: %s", hipGetErrorString(e));
        return PG_ERR_HIP;
    }
    return PG_OK;
}

inline unsigned grid_for(int64_t n, int64_t cap = 65536) {
    int64_t nb = (n + kBlock - 1) / kBlock;
    if (nb < 1) nb = 1;
    if (nb > cap) nb = cap;
    return (unsigned)nb;
}

}  // namespace pg
