/*
 * polymer_oracle.h -- CPU restatement of the Polymer edgeMap/vertexMap
 * algorithms.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check
 * and the CPU-baseline leg of the benchmark link this library as the checker.
 * The GPU product never calls it.
 *
 * Every routine cites the reference file:line whose behaviour it restates.
 * Vertex ids are int32, offsets int64 (n+1 entries, CSR/CSC), so that the
 * same arrays feed the GPU library unchanged.
 *
 * Parity pins (tests/test_oracle.py):
 *   or_pagerank_push     -- reference numa-PageRank binary output (golden),
 *                           closed forms (cycle, star, zero-in-degree parity).
 *   or_bfs_pull          -- textbook queue BFS (or_bfs_levels), reference
 *                           "Vert visited" / "Finished in" counts (golden).
 *   or_cc_labelprop      -- union-find (or_cc_unionfind), reference labels.
 *   or_dijkstra          -- brute-force Floyd-Warshall on tiny graphs.
 *   or_sssp_bf_reference -- the survey's 4-vertex counter-example.
 *   or_spmv              -- reference numa-SPMV output with integer weights
 *                           (golden) and scipy.sparse.
 *   or_bp                -- brute-force restatement in numpy on tiny graphs
 *                           plus the (1,1) invariant; parity unpinned against
 *                           the reference itself (it has no inputs/outputs).
 *   or_ppr / or_wpr_*    -- dense numpy solves (closed form).
 *   or_pagerank_delta    -- epsilon2 = 0 reduces to the power iteration (numpy),
 *                           a huge epsilon2 freezes p after round 1, and a
 *                           hand-derived 3-vertex known answer (3 rounds, the
 *                           frontier sizes 3, 2, 2).  The boundary of the
 *                           activity test (">" against the updated p) is pinned
 *                           by reading numa-PageRankDelta.C:98,114 only.
 *   or_read_bin          -- the file layout the reference reads and the PBBS
 *                           text reader on the same graph.
 */
#ifndef POLYMER_ORACLE_H
#define POLYMER_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- deterministic synthetic inputs -------------------------------- */
uint64_t or_mix64(uint64_t z);
void or_rmat_thresholds(double a, double b, double c, uint32_t thr[3]);
void or_rmat_pair(uint64_t seed, int scale, const uint32_t thr[3], uint64_t k,
                  uint32_t *src, uint32_t *dst);
/* First m distinct non-self-loop pairs of the counter sequence, returned
 * sorted by (src,dst).  Returns m on success, <0 on error. */
int64_t or_rmat_edges(int scale, int64_t m, double a, double b, double c,
                      uint64_t seed, int32_t *src, int32_t *dst);
/* Build CSR from (src,dst) sorted by src. off has n+1 entries. */
void or_csr_from_sorted(int64_t n, int64_t m, const int32_t *src,
                        const int32_t *dst, int64_t *off, int32_t *adj);
/* Transpose CSR -> CSC (in-lists ordered by source id). eperm[k] = CSR
 * index of the k-th CSC entry (may be NULL). */
void or_transpose(int64_t n, int64_t m, const int64_t *off, const int32_t *adj,
                  int64_t *toff, int32_t *tadj, int64_t *eperm);
/* Symmetrise a CSR graph (union with its transpose, de-duplicated). Returns
 * the new edge count; pass adj_out==NULL to query the count only. */
int64_t or_symmetrize(int64_t n, const int64_t *off, const int32_t *adj,
                      int64_t *off_out, int32_t *adj_out);
/* Per-edge values, functions of the vertex pair only (order independent). */
int32_t or_weight_i32(uint64_t seed, int32_t u, int32_t v);   /* [1,255], symmetric in (u,v) */
float or_value_f32(uint64_t seed, int32_t u, int32_t v);      /* (0,1], directed */
float or_unit_f32(uint64_t seed, uint64_t i);                 /* [0,1) */
float or_range_f32(uint64_t seed, uint64_t i, float lo, float hi); /* [lo,hi) */
float or_bp_pot(uint64_t seed, int32_t s, int32_t d, int k, float lo, float hi);
void or_fill_bp_pot(uint64_t seed, int64_t n, const int64_t *off, const int32_t *adj,
                    float lo, float hi, float *pot);
void or_fill_weights_i32(uint64_t seed, int64_t n, const int64_t *off,
                         const int32_t *adj, int32_t *w);
void or_fill_values_f32(uint64_t seed, int64_t n, const int64_t *off,
                        const int32_t *adj, float *w);
void or_fill_vertex_f32(uint64_t seed, int64_t n, float lo, float hi, float *x);

/* ---- PBBS text formats (IO-numa.h:112-190, IO.h:186-272) ------------ */
int or_write_adj(const char *path, int64_t n, int64_t m, const int64_t *off,
                 const int32_t *adj);
int or_write_wadj(const char *path, int64_t n, int64_t m, const int64_t *off,
                  const int32_t *adj, const int32_t *w);
/* Read: first call with adj==NULL fills *n,*m; second call fills arrays.
 * skip_self_loops follows readGraphFromFileSkipRing (IO-numa.h:225-268). */
int or_read_adj(const char *path, int skip_self_loops, int64_t *n, int64_t *m,
                int64_t *off, int32_t *adj, int32_t *w);
/* binary <base>.config / .adj / .idx (IO-numa.h:543-635); idx_bytes 4 or 8 */
int or_write_bin(const char *base, int64_t n, int64_t m, const int64_t *off,
                 const int32_t *adj, int idx_bytes);
int or_read_bin(const char *base, int64_t *n, int64_t *m, int64_t *off, int32_t *adj);

/* ---- the six applications ------------------------------------------ */
int or_pagerank_push(int64_t n, const int64_t *off, const int32_t *adj,
                     int iters, double damping, double *rank_out,
                     int64_t *edges_per_iter);
/* numa-PageRankDelta.C:63-127, 205-330: p (n), active[t] = frontier size
 * entering round t (iters entries; may be NULL). */
int or_pagerank_delta(int64_t n, const int64_t *off, const int32_t *adj, int iters,
                      double damping, double epsilon2, double *p, int64_t *active);
int or_pagerank_pull_omp(int64_t n, const int64_t *off, const int64_t *toff,
                         const int32_t *tadj, int iters, double damping,
                         int threads, double *rank_out);
int or_bfs_pull(int64_t n, const int64_t *toff, const int32_t *tadj,
                int32_t src, int32_t *parent, int32_t *level, int32_t *rounds);
int or_bfs_levels(int64_t n, const int64_t *off, const int32_t *adj,
                  int32_t src, int32_t *level);
int or_cc_labelprop(int64_t n, const int64_t *off, const int32_t *adj,
                    int32_t *label, int32_t *rounds);
int or_cc_unionfind(int64_t n, const int64_t *off, const int32_t *adj,
                    int32_t *label);
int or_sssp_bf_reference(int64_t n, const int64_t *off, const int32_t *adj,
                         const int32_t *w, int32_t src, int32_t *dist,
                         int32_t *rounds);
int or_sssp_bf(int64_t n, const int64_t *off, const int32_t *adj,
               const int32_t *w, int32_t src, int32_t *dist, int32_t *rounds);
int or_dijkstra(int64_t n, const int64_t *off, const int32_t *adj,
                const int32_t *w, int32_t src, int32_t *dist);
int or_spmv(int64_t n, const int64_t *off, const int32_t *adj, const double *w,
            const double *x0, int iters, double *y_out);
int or_spmv_adjoint(int64_t n, const int64_t *off, const int32_t *adj, const double *w,
                    const double *x0, const double *g, int iters, double *grad_x0,
                    double *grad_w);
int or_spmv_pull_omp(int64_t n, const int64_t *toff, const int32_t *tadj, const int64_t *teid,
                     const double *w, const double *x0, int iters, int threads,
                     double *y_out);
int or_bp_pull_omp(int64_t n, const int64_t *toff, const int32_t *tadj, const int64_t *teid,
                   const float *vpot, const float *epot, const double *prod0, int iters,
                   int threads, double *prod_out);
int or_bp(int64_t n, const int64_t *off, const int32_t *adj,
          const float *vpot, const float *epot, const double *prod0, int iters,
          double *prod_out);
int or_bp_state_f32(int64_t n, const int64_t *off, const int32_t *adj,
                    const float *vpot, const float *epot, const double *prod0, int iters,
                    double *prod_out);
int or_bp_f32(int64_t n, const int64_t *off, const int32_t *adj,
              const float *vpot, const float *epot, const float *prod0,
              int iters, float *prod_out);

/* ---- batched and differentiated forms ------------------------------- */
int or_ppr(int64_t n, const int64_t *off, const int32_t *adj, int iters,
           double damping, const double *tele, double *rank_out);
int or_wpr_forward(int64_t n, const int64_t *off, const int32_t *adj,
                   const double *wn, const double *tele, double damping,
                   int max_iters, double tol, double *x_out, int *iters_used);
int or_wpr_adjoint(int64_t n, const int64_t *off, const int32_t *adj,
                   const double *wn, const double *x, const double *g,
                   double damping, int max_iters, double tol,
                   double *lambda_out, double *grad_w, double *grad_tele,
                   int *iters_used);

#define OR_INF_DIST 1073741823 /* INT_MAX/2, numa-BellmanFord.C:283 */

#ifdef __cplusplus
}
#endif
#endif
