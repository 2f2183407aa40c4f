// pg_internal.h -- host-side structures of libpolymer_gpu.so shared by the
// translation units (graph build, applications, batched forms, adjoint).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "polymer_gpu.h"

namespace pg {

// ---- errors ---------------------------------------------------------------
void set_error(const char *fmt, ...);

#define PG_CHECK_HIP(call)                                                       \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess) {                                                  \
            ::pg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                            __FILE__, __LINE__);                                 \
            return _e == hipErrorOutOfMemory ? PG_ERR_OOM : PG_ERR_HIP;          \
        }                                                                        \
    } while (0)

#define PG_TRY(call)                      \
    do {                                  \
        int _r = (call);                  \
        if (_r != PG_OK) return _r;       \
    } while (0)

#define PG_REQUIRE(cond, code, ...)              \
    do {                                         \
        if (!(cond)) {                           \
            ::pg::set_error(__VA_ARGS__);        \
            return (code);                       \
        }                                        \
    } while (0)

hipStream_t stream();
hipStream_t side_stream();  // a second stream (default priority) for overlapped work
bool initialised();

// ---- pull schedule ----------------------------------------------------------
// Vertices are visited in an order where the pulled degree is non-increasing
// (the in-degree relabelling makes this the identity for the CSC; a vertex
// list gives the order for the CSR).  Contiguous runs with the same
// lanes-per-vertex form segments; blocks of 256 threads are laid out segment
// after segment.
constexpr int kBlock = 256;
constexpr int kMaxSegs = 12;
constexpr int kStatSlots = 64;  // per-launch statistic counters (spread atomics)
constexpr int kModeMulti = 0;    // several blocks per vertex
constexpr int kModeBlock = 256;  // one block per vertex

struct PullSeg {
    int32_t v_begin, v_end;  // positions in the visiting order
    int32_t g;               // lanes per vertex: 1..64, kModeBlock, kModeMulti
    int32_t block_begin;
};

struct MultiItem {  // one block of a multi-block vertex
    int32_t pos;         // position in the visiting order
    int32_t chunk;       // chunk index within the vertex
    int32_t nchunks;
    int32_t first;       // first multi-block index of this vertex
    int32_t slot;        // ticket slot of this vertex
    int32_t pad[3];
};

struct PullArgs {
    const int32_t *off;    // offsets of the pulled direction
    const int32_t *adj;    // neighbour ids of the pulled direction
    const int32_t *vlist;  // visiting order (nullptr = identity)
    const MultiItem *multi;
    void *partial;         // 32 bytes per multi block
    unsigned *tickets;     // one per multi vertex, zero between launches
    int32_t chunk;
    int32_t nseg;
    PullSeg seg[kMaxSegs];
};

struct PullParams {
    int items_per_lane = 16;
    int64_t block_threshold = 16384;  // in-degree above which one block per vertex
    int64_t multi_threshold = 65536;  // above which several blocks per vertex
    int64_t chunk = 16384;            // edges per block for multi-block vertices
};

// ---- source-sorted tiles --------------------------------------------------------
// The light destinations of the CSC (in-degree in [1, block_threshold)) are cut
// into tiles of kTileMaxV consecutive positions.  A tile stores its in-edges
// sorted by source id, with the destination as a 16-bit offset from the tile's
// first position, so the gathers of a wavefront walking the list hit few cache
// lines.  A tile's list is split into parts of about tile_edges edges (at least
// as many edges as the tile has vertices); one workgroup per part adds the
// gathered values into per-destination accumulators in LDS.  A tile with one
// part finalises its vertices in place; the parts of a larger tile store their
// accumulators to a partial slot each, and a second launch (one thread per
// vertex) sums the slots and finalises.
constexpr int kTileMaxV = 8192;      // destinations per tile (64 KiB of LDS accumulators)
constexpr int kTileThreads = 1024;   // threads per tile workgroup

struct TilePart {
    int32_t tile;
    int32_t pad;
    int64_t slot;    // element offset of this part's partial slot, -1 = single part
    int64_t e0, e1;  // edge range in src / dst
    int64_t cold_e;  // first edge in [e0, e1] whose source is cold (e1 if none)
    int64_t cold_v;  // index in the cold value array of edge cold_e
};

// Cold sources (internal id >= cold_begin: low out-degree, little reuse) are not
// gathered by the tile kernel.  A pre-pass walks the cold edges grouped by source
// chunk (kColdChunk consecutive sources staged in LDS) and writes each edge's
// value into a cold value array laid out in tile order; the tile kernel streams
// it.  A random line fetch per cold edge becomes about 10 streamed bytes.
constexpr int kColdChunk = 16384;  // sources per chunk (64 KiB of LDS)

struct ColdPart {
    int32_t chunk;
    int32_t pad;
    int64_t a, b;  // range in the pre-pass lists
};

struct ColdArgs {
    const ColdPart *parts;
    const uint32_t *pos;    // index in the cold value array
    const uint16_t *srcl;   // source - chunk's first source
    float *val;             // cold values, tile order
    int32_t nparts;
    int32_t cold_begin;
    int64_t n;
    void *val64;            // the same for an Op with 8-byte values (or null)
};

struct TileFin {     // kBlock vertices of a tile with several parts
    int32_t tile;
    int32_t nparts;
    int32_t v0;      // first vertex offset within the tile
    int32_t pad;
    int64_t slot0;   // offset of the tile's first partial slot (slots are nv apart)
};

// A tile edge is one packed word: destination offset in the low 13 bits and, for a
// warm edge, the source as an offset from its 64-edge group's base in the high 19
// bits.  A group whose warm sources span 2^19 or more (at a tile boundary, or in a
// sparse stretch) has base -(k+1) and keeps its sources in escape row k.
constexpr int kPackDstBits = 13;  // kTileMaxV = 2^13
constexpr int kPackGroup = 64;

struct TileArgs {
    const int32_t *vbeg;         // ntiles + 1 positions (visiting order)
    const TilePart *parts;       // one per workgroup
    const TileFin *fin;          // one per finishing block
    const uint32_t *packed;      // one word per tile edge
    const int32_t *gbase;        // per 64-edge group: source base, or -(escape row + 1)
    const int32_t *esc;          // escape rows of 64 source ids
    unsigned long long *partial; // fixed-point partial sums
    const float *cold_val;       // values of the cold edges (written by the pre-pass)
    int32_t nparts;
    int32_t nfin;
    const float *wt;             // edge weights in tile order (weighted graphs), or null
    const void *cold_val64;      // cold values of an Op with 8-byte values (or null)
};

struct TileSet {
    bool built = false;
    int64_t light_begin = -1, light_end = -1;  // light range the tiles were cut for
    int64_t tile_edges = 0;
    TileArgs args{};
    int32_t *d_vbeg = nullptr;
    TilePart *d_parts = nullptr;
    TileFin *d_fin = nullptr;
    int32_t *d_src = nullptr;    // build only (freed once packed)
    uint16_t *d_dst = nullptr;   // build only
    uint32_t *d_packed = nullptr;
    int32_t *d_gbase = nullptr;
    int32_t *d_esc = nullptr;
    int64_t n_esc = 0;
    float *d_wt = nullptr;  // edge weights in tile order (weighted graphs only)
    unsigned long long *d_partial = nullptr;
    int32_t ntiles = 0;
    int64_t m_tiled = 0;
    // cold split
    int64_t cold_begin = -1;  // the setting the tiles were built for
    ColdArgs cold{};
    ColdPart *d_cold_parts = nullptr;
    uint32_t *d_cold_pos = nullptr;
    uint16_t *d_cold_srcl = nullptr;
    float *d_cold_val = nullptr;
    int64_t m_cold = 0;
    // every multi-part tile lies below the first cold source: the next iteration's
    // cold pre-pass can run beside this iteration's finish pass (two events order it)
    bool cold_overlap = false;
    hipEvent_t ev_tiles = nullptr, ev_cold = nullptr;
};

struct PullSchedule {
    PullArgs args{};
    int nblocks = 0;
    int nblocks_body = 0;  // the blocks that cover the positions with edges
    int64_t body_end = 0;  // the positions with edges are [0, body_end)
    PullParams params;
    MultiItem *d_multi = nullptr;
    void *d_partial = nullptr;
    unsigned *d_tickets = nullptr;
    bool built = false;
    // the same schedule without the light range [light_begin, light_end), which
    // the tiles cover (multi and block segments, then the zero-degree tail)
    PullArgs args_hot{};
    int nblocks_hot = 0;
    int nblocks_hot_body = 0;  // the same without the zero-degree tail (its last segment)
    int64_t light_begin = 0, light_end = 0;
};

// ---- graph ---------------------------------------------------------------------
struct pg_graph_impl;

}  // namespace pg

struct pg_graph {
    int64_t n = 0, m = 0;
    bool sym = false;
    int w_dtype = PG_W_NONE;
    int64_t max_in_deg = 0, max_out_deg = 0;
    // device arrays, internal ids (relabelled by in-degree, descending)
    int32_t *out_off = nullptr, *out_adj = nullptr;
    int32_t *in_off = nullptr, *in_adj = nullptr;  // alias out_* when sym
    int32_t *perm = nullptr;   // input id -> internal id
    int32_t *iperm = nullptr;  // internal id -> input id
    int32_t *out_eid = nullptr, *in_eid = nullptr;  // -> input CSR edge index
    int32_t *out_w_i32 = nullptr, *in_w_i32 = nullptr;
    float *out_w_f32 = nullptr, *in_w_f32 = nullptr;
    int32_t *out_order = nullptr;  // internal ids by out-degree, descending
    float *bp_pot = nullptr;       // 4 per edge, CSC order
    // many-graph form: per internal vertex (1-damping)/n_g factor 1/n_g
    float *inv_n = nullptr;
    int num_graphs = 1;
    // schedules
    pg::PullSchedule in_sched;   // pull over CSC (in-edges)
    pg::PullSchedule out_sched;  // pull over CSR (out-edges)
    pg::TileSet tiles;           // source-sorted tiles of the CSC light range
    pg::TileSet tiles_out;       // the same over the CSR, every vertex (the SpMV adjoint)
    bool use_tiles = true;       // PageRank-family and SpMV solves use the tiles
    int64_t tile_max_degree = (int64_t)1 << 30;  // heavier vertices stay in the hot pull
    int64_t tile_edges = 524288; // target edges per tile part
    int64_t tile_cold = 2 << 20; // first cold source id of the tiles (<0: no cold split)
    // input-id CSR / CSC for the header-level edgeMap API (built on request)
    bool view_built = false;
    int32_t *v_out_off = nullptr, *v_out_adj = nullptr, *v_in_off = nullptr, *v_in_adj = nullptr;
    void *v_out_w = nullptr, *v_in_w = nullptr;
    // cached device workspace
    std::map<std::string, std::pair<void *, size_t>> ws;
    // captured solves (hipGraph executables) and the buffer holding their result
    std::map<std::string, std::pair<hipGraphExec_t, float *>> graph_exec;
    // saved state of the last weighted PageRank forward (for the adjoint)
    // rounds of the first captured batch of the next BFS (from the last search)
    int bfs_rounds_hint = 0;
    bool wpr_valid = false;
    // iterations of the last pg_spmv run with PG_SPMV_SAVE (its iterates and exponent
    // shifts are kept for pg_spmv_adjoint), 0 = none kept
    int spmv_saved_iters = 0;
    // BP: the iteration at which each vertex leaves the frontier, as a byte per CSC
    // edge (workspace bp_de): 0 not built, 1 built, -1 not representable (the
    // frontier shrinks for more than 254 iterations)
    int bp_death_state = 0;
    // BP: the light range [lb, be) of the CSC and its edge chunks (workspaces
    // bp_chunks, bp_heads), valid while the in-edge schedule is unchanged
    bool bp_light_built = false;
    int64_t bp_light_lb = 0, bp_light_be = 0;
    int bp_light_nch = 0;
    double wpr_damping = 0.85;
    // allocations owned by the graph
    std::vector<void *> owned;
};

struct pg_subset {
    pg_graph *g = nullptr;
    uint32_t *bits = nullptr;  // n/32 words
    int32_t *list = nullptr;   // n entries
    unsigned long long *counters = nullptr;  // [0]=count [1]=out-edges
};

namespace pg {

// device allocation helpers (owned by the graph)
int galloc(pg_graph *g, void **p, size_t bytes);
template <class T>
int galloc_t(pg_graph *g, T **p, size_t count) {
    return galloc(g, reinterpret_cast<void **>(p), count * sizeof(T) + 16);
}
// named workspace, grown on demand, kept until the graph is destroyed
int workspace(pg_graph *g, const char *name, size_t bytes, void **p);
template <class T>
int workspace_t(pg_graph *g, const char *name, size_t count, T **p) {
    return workspace(g, name, count * sizeof(T) + 16, reinterpret_cast<void **>(p));
}

int build_schedule(pg_graph *g, bool in_direction);
int ensure_out_schedule(pg_graph *g);
// builds (once per block threshold) the source-sorted tiles of the CSC light range
int ensure_tiles(pg_graph *g);
// builds (once per tile setting) the tiles of the pull over the CSR: every vertex a
// destination of its out-lists, sources = the out-neighbours
int ensure_tiles_out(pg_graph *g);
void free_tiles(pg_graph *g);

// device-wide helpers implemented in pg_build.hip (hipcub)
int exclusive_scan_u32(pg_graph *g, const uint32_t *in, uint32_t *out, int64_t n);

// event timing on the library stream
// bytes (a multiple of 4, at most 4 KiB) of device memory into host memory once
// everything queued on the library stream before the call has finished
int fetch_words(void *host, const void *dev, size_t bytes);

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    int start();
    int stop(double *ms);
    ~Timer();
};

// per-kernel timing used with PG_PROFILE_KERNELS
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;
    int begin();
    int end();
    int total(double *ms, int *launches);
    ~KernelTimer();
};

// frontier conversions (pg_build.hip)
int bits_to_list(pg_graph *g, const uint32_t *bits, int32_t *list, unsigned long long *d_count);
int list_to_bits(pg_graph *g, const int32_t *list, const unsigned long long *d_count,
                 uint32_t *bits);
// adds the out-degrees of list[0..*d_count) to d_out[stride * k], k < 64 (stride 0:
// all to d_out[0])
int list_degree_sum(pg_graph *g, const int32_t *list, const unsigned long long *d_count,
                    const int32_t *off, unsigned long long *d_out, int stride = 0);
int fill_bits_all(pg_graph *g, uint32_t *bits);

// edge arrays between the caller's CSR order and the internal CSR/CSC order
int edges_from_input_f32(pg_graph *g, const float *host, bool csc, float *dev);
int edges_to_input_f32(pg_graph *g, const float *dev, bool csc, float *host);

int create_from_host(int64_t n, int64_t m, const int64_t *off, const int32_t *adj, const void *w,
                     int w_dtype, uint32_t flags, pg_graph **out);

// copy an internal-order array to a host array in input order
int to_input_order_i32(pg_graph *g, const int32_t *dev_internal, int32_t *host, bool map_ids);
int to_input_order_f32(pg_graph *g, const float *dev_internal, float *host, int width);
int from_input_order_f32(pg_graph *g, const float *host, float *dev_internal, int width);

}  // namespace pg
