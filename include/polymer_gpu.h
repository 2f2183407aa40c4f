/*
 * polymer_gpu.h -- C ABI of libpolymer_gpu.so, the MI355X (gfx950) engine
 * behind Polymer's edgeMap/vertexMap path.
 *
 * The reference has no binary interface: its "plugin API" is the header-only
 * template engine polymer.h / polymer-wgh.h instantiated by the six app
 * drivers.  This header is the binary boundary that engine's callers bind
 * instead: plain pointers and sizes, int status codes, opaque handles.
 *
 * Conventions
 *   - Every function returns PG_OK (0) or a negative PG_ERR_*; the message
 *     of the last failure on the calling thread is pg_last_error().
 *   - Vertex ids seen by the caller are the ids of the input graph.  The
 *     library stores the graph relabelled by in-degree (the reference also
 *     relabels, numa-PageRank.C:523-546 PR_Hash_F) and maps ids back on
 *     every output.
 *   - Pointers marked "host" are host memory owned by the caller.  Results
 *     stay resident in device memory owned by the graph handle; passing NULL
 *     for a host output skips the copy (the timed benchmark path).
 *   - A graph handle is used from one host thread at a time; all of its work
 *     runs on the library's stream, in order.
 *   - Offsets are int64 (n+1 entries), ids int32; m < 2^31.
 */
#ifndef POLYMER_GPU_H
#define POLYMER_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PG_OK 0
#define PG_ERR_ARG (-1)     /* bad argument (null handle, id out of range, ...) */
#define PG_ERR_HIP (-2)     /* a HIP runtime call failed */
#define PG_ERR_OOM (-3)     /* device allocation failed */
#define PG_ERR_STATE (-4)   /* library not initialised / wrong graph kind */
#define PG_ERR_UNSUPPORTED (-5)

/* ---- graph flags (pg_graph_create / pg_graph_rmat) --------------------- */
#define PG_GRAPH_SYMMETRIC 0x1u       /* adj is symmetric: in-lists = out-lists (readGraph -s) */
#define PG_GRAPH_SKIP_SELF_LOOPS 0x2u /* drop (v,v) edges as readGraphFromFileSkipRing, IO-numa.h:225-268 */
#define PG_GRAPH_SYMMETRIZE 0x4u      /* pg_graph_rmat only: union with the transpose, deduplicated */
#define PG_GRAPH_WEIGHT_I32 0x8u      /* pg_graph_rmat only: int32 weights U[1,255] (pair hash, symmetric) */
#define PG_GRAPH_VALUE_F32 0x10u      /* pg_graph_rmat only: fp32 edge values U(0,1] (directed pair hash) */

/* weight element types for pg_graph_create */
#define PG_W_NONE 0
#define PG_W_I32 1 /* wghGraph intE weights (IO.h:209-210) */
#define PG_W_F32 2

/* edgeMap directions (polymer.h:2241-2309): AUTO applies the reference's
 * rule "dense iff |F| + sum_{v in F} outdeg(v) >= threshold" (P9, >=). */
#define PG_DIR_AUTO 0
#define PG_DIR_SPARSE 1 /* push from a vertex list (edgeMapSparseV3) */
#define PG_DIR_DENSE 2  /* dense round (pull for BFS/CC, forward push for SSSP) */

typedef struct pg_graph pg_graph;
typedef struct pg_subset pg_subset;

/* Per-call statistics.  ms = device time of the solve between events on the
 * library stream (host copies excluded); kernel_ms/kernel_launches = time of
 * the dominant per-iteration kernel only (events around each launch when
 * PG_PROFILE_KERNELS is set in flags of the call, else 0). */
typedef struct pg_stats {
    int32_t rounds;          /* iterations / BFS levels processed */
    int32_t dense_rounds;    /* rounds executed in the dense direction */
    int64_t edges;           /* sum over rounds of out-edges of the active set */
    int64_t active;          /* sum over rounds of |frontier| */
    double ms;
    double kernel_ms;
    int32_t kernel_launches;
    int32_t pad;
    double residual;         /* last fp64 residual (fixed-point solvers) */
} pg_stats;

#define PG_PROFILE_KERNELS 0x1u
#define PG_BP_FP32 0x2u /* pg_bp: fp32 product state and messages (default fp64) */
/* pg_pagerank / pg_spmv on tiles: run every iteration's cold pre-pass in order on
 * the library stream instead of beside the previous iteration's finishing pass
 * (same sums; a check of the overlapped schedule) */
#define PG_SERIAL_COLD 0x4u
/* pg_spmv: keep every iterate y_0..y_T on the device for pg_spmv_adjoint */
#define PG_SPMV_SAVE 0x8u

/* ---- library ------------------------------------------------------------ */
int pg_init(int device);
int pg_finalize(void);
int pg_last_error(char *buf, size_t len);
int pg_device_name(char *buf, size_t len);

/* ---- graphs (graph.h:11-164; ingest IO-numa.h:112-315, IO.h:186-272) ----
 * pg_graph_create uploads a host CSR (off n+1, adj m) and builds the device
 * CSR + CSC, the in-degree relabelling and the pull schedule.  w is NULL or
 * m weights of w_dtype in the CSR's edge order. */
int pg_graph_create(int64_t n, int64_t m, const int64_t *off, const int32_t *adj,
                    const void *w, int w_dtype, uint32_t flags, pg_graph **out);
/* Deterministic recursive-matrix graph generated on the device: the first m
 * distinct non-self-loop pairs of the counter sequence of seed (no
 * duplicates, no self-loops: every app sees one graph, P3). */
int pg_graph_rmat(int scale, int64_t m, double a, double b, double c, uint64_t seed,
                  uint32_t flags, uint64_t wseed, pg_graph **out);
/* PBBS AdjacencyGraph / WeightedAdjacencyGraph text file. */
int pg_graph_load_adj(const char *path, uint32_t flags, pg_graph **out);
/* Binary graph <base>.config / .adj / .idx, replacing readGraphFromBinary and
 * readWghGraphFromBinary (IO-numa.h:543-720): n as text, m 32-bit neighbour ids,
 * n out-list offsets of 4 or 8 bytes (taken from the file's size; any other size
 * is an error).  PG_GRAPH_WEIGHT_I32: unit weights, as the reference's weighted
 * reader gives (IO-numa.h:683). */
int pg_graph_load_bin(const char *base, uint32_t flags, pg_graph **out);
/* The whole-graph dump of loadGraphFromBin / dumpGraphToBin (IO-numa.h:758-900):
 * int64 total size, int32 n, int64 m, then per vertex int32 -1, id, out-degree,
 * in-degree, out-neighbours, in-neighbours.  Loading checks the markers, ids,
 * totals and recorded in-degrees; dumping writes the graph in input ids. */
int pg_graph_load_dump(const char *path, uint32_t flags, pg_graph **out);
int pg_graph_dump(pg_graph *g, const char *path);
int pg_graph_destroy(pg_graph *g);
/* n, m (directed entries), max in-degree, has-weights (0 none, 1 i32, 2 f32) */
int pg_graph_info(const pg_graph *g, int64_t *n, int64_t *m, int64_t *max_in_degree,
                  int32_t *w_dtype);
/* CSR in input ids (out-lists sorted by id).  Any pointer may be NULL. */
int pg_graph_download(const pg_graph *g, int64_t *off, int32_t *adj, void *w);
/* Pull-schedule tuning: items per lane, block/multi thresholds.  0 = keep. */
int pg_graph_set_schedule(pg_graph *g, int items_per_lane, int64_t block_threshold,
                          int64_t multi_threshold, int64_t chunk);
/* Source-sorted tiles for the PageRank-family pull: enable = 1 on (the
 * default), 0 off, negative keeps; max_degree = in-degree up to which a vertex
 * is put in a tile (heavier ones keep the block pull), 0 keeps (default 2^30:
 * all); tile_edges = target edges per workgroup, 0 keeps (default 524288);
 * cold_from = first internal source id whose values a pre-pass streams into
 * tile order instead of the tile kernel gathering them (default 2^21), -1 =
 * none, 0 keeps.  Schedule tuning only (the reference partitions by NUMA node
 * instead, polymer.h:131-234); results agree to fp32 rounding. */
int pg_graph_set_tiles(pg_graph *g, int enable, int64_t max_degree, int64_t tile_edges,
                       int64_t cold_from);
/* Sizes of the tile structures of the last PageRank-family solve (built on
 * first use): tiles, workgroup parts, tiled in-edges, cold in-edges served by
 * the pre-pass, 64-edge groups whose sources are stored unpacked, and whether
 * the next iteration's cold pre-pass may run beside the finishing pass (1: every
 * tile with several parts lies below the first cold source).  Any pointer may
 * be NULL; all zero before the first tiled solve. */
int pg_graph_tile_stats(const pg_graph *g, int64_t *tiles, int64_t *parts, int64_t *tiled_edges,
                        int64_t *cold_edges, int64_t *escape_groups, int32_t *cold_overlap);
/* The current pull schedule and tile setting (any pointer may be NULL). */
int pg_graph_get_schedule(const pg_graph *g, int32_t *items_per_lane, int64_t *block_threshold,
                          int64_t *multi_threshold, int64_t *chunk, int32_t *tiles,
                          int64_t *max_degree, int64_t *tile_edges, int64_t *cold_from);
/* Times candidate schedules on PageRank iterations and keeps the fastest. */
int pg_graph_autotune(pg_graph *g, int32_t *items_per_lane, int64_t *block_threshold);
/* Fills m edge values in CSR order of the input ids (pair hash, see oracle). */
int pg_graph_bp_potentials(pg_graph *g, uint64_t seed, float lo, float hi);

/* Device view of the graph in input ids for code compiled against the
 * edgeMap/vertexMap header (polymer.h): CSR and CSC (graph.h:11-87
 * getOutNeighbor / getInNeighbor), neighbour lists sorted by id, weights
 * in the order of each adjacency array.  Built on first request and owned
 * by the graph; all pointers are device pointers.  For a symmetric graph
 * the in-arrays alias the out-arrays.  stream is the library's hipStream_t:
 * work queued there is ordered with every library call. */
typedef struct pg_view {
    int64_t n, m;
    int32_t symmetric, w_dtype;
    const int32_t *out_off; /* n + 1 */
    const int32_t *out_adj; /* m */
    const int32_t *in_off;  /* n + 1 */
    const int32_t *in_adj;  /* m */
    const void *out_w;      /* m, or NULL */
    const void *in_w;       /* m, or NULL */
    void *stream;
} pg_view;
int pg_graph_view(pg_graph *g, pg_view *v);

/* ---- vertex subsets (polymer.h:521-881 LocalFrontier / vertices) ------- */
int pg_subset_create(pg_graph *g, pg_subset **out);
int pg_subset_destroy(pg_subset *s);
int pg_subset_clear(pg_subset *s);
int pg_subset_fill(pg_subset *s);                                /* all vertices */
int pg_subset_from_dense(pg_subset *s, const uint8_t *flags);     /* host, n bools */
int pg_subset_from_sparse(pg_subset *s, const int32_t *ids, int64_t k); /* host */
/* toDense (polymer.h:594-601): host n bools. */
int pg_subset_to_dense(pg_subset *s, uint8_t *flags);
/* toSparse = packIndex (polymer.h:552-567): ascending ids, host. */
int pg_subset_to_sparse(pg_subset *s, int32_t *ids, int64_t cap, int64_t *k);
/* No counterpart in the reference: pg_subset_to_sparse for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Writes the ascending input ids of s to d_ids (device, at
 * least n entries) and their number to *d_count (device int64), asynchronously
 * on stream (a hipStream_t; NULL = the library stream).  Device memory beyond
 * the caller's comes from one static array per device (about 4 MiB, set up by
 * pg_init): n above 2^25 vertices is refused with PG_ERR_UNSUPPORTED.  Every call
 * on a device shares that array, so calls that overlap on one device race; the
 * subset must not be modified while the call runs. */
int pg_subset_to_sparse_noalloc(pg_subset *s, int32_t *d_ids, int64_t *d_count, void *stream);
/* vertexCounter (polymer.h:2312-2338): |F| and sum of out-degrees. */
int pg_subset_size(pg_subset *s, int64_t *count, int64_t *out_edges);

/* ---- single-graph applications (numa-*.C) ------------------------------ */
/* numa-PageRank.C:273-335: fixed iterations, shrinking frontier (P1), no
 * dangling redistribution (P2); fp32 values, fp32 accumulation. rank: host n. */
int pg_pagerank(pg_graph *g, int iters, double damping, float *rank, uint32_t flags,
                pg_stats *st);
/* numa-PageRankDelta.C:63-330: p = 0, Delta = 1/n, all active; each round the
 * active sources' Delta/outdeg is summed into every vertex, which then updates
 * p and stays active iff |Delta| > epsilon2 * p (the reference uses damping
 * 0.85, epsilon2 0.01).  fp64 throughout.  p: host n doubles.  st->active and
 * st->edges sum the frontier sizes and out-degrees over the rounds. */
int pg_pagerank_delta(pg_graph *g, int iters, double damping, double epsilon2, double *p,
                      uint32_t flags, pg_stats *st);
/* numa-SPMV.C:168-193: x_{k+1}[d] = sum_{s->d} x_k[s] w(s,d) over all
 * vertices; fp32 weights (w_dtype F32, or I32 converted), fp32 arithmetic.
 * Each iteration carries one power-of-two exponent shared by all vertices,
 * so values beyond fp32's range (the reference's fp64 iterates reach 1e42)
 * keep fp32 precision; y (host n) is returned in fp64 as mantissa * 2^exp.
 * x0: host n. */
int pg_spmv(pg_graph *g, int iters, const float *x0, double *y, uint32_t flags, pg_stats *st);
/* numa-BP.C:280-316 with the frontier of numa-BP.C:292-308.  vpot: host
 * 2n (phi_v[l]); epot: host 4m in CSR order of the input ids, W_e[k][l] at
 * 4e+2k+l, or NULL to use pg_graph_bp_potentials; prod0: host 2n initial
 * products; prod: host 2n result (fp64).  The unnormalised messages raise a
 * product to about its in-degree every iteration, so the product state and
 * the messages are kept in fp64 unless PG_BP_FP32 is set in flags. */
int pg_bp(pg_graph *g, int iters, const float *vpot, const float *epot, const float *prod0,
          double *prod, uint32_t flags, pg_stats *st);
/* numa-BFS.C:258-296.  threshold: < 0 -> direction optimising (bottom-up once
 * the frontier's out-degree sum exceeds 1/14 of the edges not yet reached,
 * top-down again when the frontier holds fewer than n/24 vertices);
 * 0 -> every round dense (the reference's setting, numa-BFS.C:268);
 * > 0 -> dense iff |F| + sum of out-degrees >= threshold (polymer.h:2247).
 * level/parent: host n (level -1 / parent -1 when unreached). */
int pg_bfs(pg_graph *g, int64_t src, int64_t threshold, int32_t *level, int32_t *parent,
           uint32_t flags, pg_stats *st);
/* numa-BFS-async-pipe.C:140-175 with edgeMapSparseAsyncPipe (polymer.h:1512-1751):
 * one asynchronous traversal to quiescence from src; every reachable vertex is
 * claimed once with a CAS of its parent from -1 by an already claimed in-neighbour
 * (a spanning tree of the reachable set, not in general a breadth-first one).
 * parent: host n (-1 unreached); visited: the number of reached vertices. */
int pg_bfs_async(pg_graph *g, int64_t src, int32_t *parent, int64_t *visited, uint32_t flags,
                 pg_stats *st);
/* numa-Components.C:233-281: min-label propagation, labels = min input id
 * of the component (symmetric graph).  threshold < 0 -> m/8 (:226). */
int pg_cc(pg_graph *g, int64_t threshold, int32_t *label, uint32_t flags, pg_stats *st);
/* numa-BellmanFord.C:166-211 relaxation with int32 weights; returns exact
 * shortest distances (unreached = 1073741823 = INT_MAX/2).  threshold < 0
 * -> m/20 (:182). */
int pg_sssp(pg_graph *g, int64_t src, int64_t threshold, int32_t *dist, uint32_t flags,
            pg_stats *st);

/* ---- batched forms -------------------------------------------------------- */
/* k <= 64 BFS sources as one 64-bit frontier mask per vertex.  level: host
 * k*n int32, source-major (level[c*n + v]). */
int pg_msbfs(pg_graph *g, const int64_t *src, int k, int32_t *level, uint32_t flags,
             pg_stats *st);
/* k <= 64 Bellman-Ford sources (int32 weights). dist: host k*n, source-major. */
int pg_mssssp(pg_graph *g, const int64_t *src, int k, int32_t *dist, uint32_t flags,
              pg_stats *st);
/* Personalized PageRank (all vertices active every iteration):
 *   r_0 = v, r_{t+1}[d] = damping * sum_{s->d} r_t[s]/outdeg(s) + (1-damping) v[d]
 * with one-hot teleports v = e_{src[c]}.  k <= 64 vectors in one pass as
 * k-wide rows; rank: host k*n, source-major. */
int pg_ppr_batched(pg_graph *g, const int64_t *src, int k, int iters, double damping,
                   float *rank, uint32_t flags, pg_stats *st);
/* The same iteration for one teleport vector at a time (the plain form). */
int pg_ppr(pg_graph *g, int64_t src, int iters, double damping, float *rank, uint32_t flags,
           pg_stats *st);

/* Many independent graphs in one device graph: graph i owns input ids
 * [vbase[i], vbase[i+1]).  PageRank of numa-PageRank.C per graph (its own n
 * in (1-damping)/n), all graphs advanced together, the iterations captured
 * in one hipGraph.  rank: host total-n. */
int pg_graph_multi(int num_graphs, const int64_t *vbase, int64_t m, const int64_t *off,
                   const int32_t *adj, pg_graph **out);
int pg_pagerank_multigraph(pg_graph *g, int iters, double damping, float *rank, uint32_t flags,
                           pg_stats *st);

/* ---- differentiation: weighted PageRank ---------------------------------
 * x = damping * W x + (1-damping) * v, W[d][s] = wn[e] for CSR edge e=(s->d)
 * (wn: host m fp32 in CSR order of the input ids, column-normalised by the
 * caller).  Solved by fixed-point iteration until ||x_{t+1}-x_t||_1 <= tol
 * (fp64 residual) or max_iters.  x: host n. */
int pg_wpr_forward(pg_graph *g, const float *wn, const float *tele, double damping,
                   int max_iters, double tol, float *x, uint32_t flags, pg_stats *st);
/* Adjoint of the last pg_wpr_forward on g for loss sum_i gvec_i x_i:
 * lambda = damping W^T lambda + gvec (pull over the transposed graph, same
 * fixed point), grad_w[e] = damping * lambda[d] * x[s] (CSR order, host m),
 * grad_tele = (1-damping) lambda (host n).  lambda may be NULL. */
int pg_wpr_adjoint(pg_graph *g, const float *gvec, int max_iters, double tol, float *lambda,
                   float *grad_w, float *grad_tele, uint32_t flags, pg_stats *st);

/* ---- differentiation: SpMV ------------------------------------------------
 * Adjoint of the last pg_spmv run with PG_SPMV_SAVE on g (numa-SPMV.C:56-97
 * iterated T times, :168-193: y_{t+1}[d] = sum_{e=(s->d)} y_t[s] w[e]) for the
 * loss L = sum_i gvec_i y_T[i]:
 *   mu_T = gvec,  mu_t[s] = sum_{e=(s->d)} w[e] mu_{t+1}[d]   (a pull over the CSR,
 *   i.e. the transposed graph), dL/dx0 = mu_0, dL/dw[e] = sum_{t<T} y_t[s] mu_{t+1}[d].
 * fp32 iterates with one power-of-two exponent per iteration (as pg_spmv), the
 * weight gradient accumulated in fp64.  gvec: host n; grad_x0: host n doubles;
 * grad_w: host m doubles in CSR order of the input ids.  Either output may be
 * NULL.  No counterpart in the reference, which has no derivatives. */
int pg_spmv_adjoint(pg_graph *g, const float *gvec, double *grad_x0, double *grad_w,
                    uint32_t flags, pg_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* POLYMER_GPU_H */
