// pg_apps.hip -- the six Polymer applications on the MI355X engine.
//
// Each app is an edge operator (the reference's functor) plugged into the
// pull engine (k_pull) or the push engine (k_push) plus a host loop that
// mirrors the app's iteration protocol (the pthread SPMD loop of numa-*.C),
// with kernel boundaries in place of the reference's barriers.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pg_engine.h"
#include "pg_internal.h"

namespace pg {

#define PG_LAUNCH_K(kern, grid, ...)                                                  \
    do {                                                                              \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream(), __VA_ARGS__); \
        PG_CHECK_HIP(hipGetLastError());                                              \
    } while (0)

static inline int64_t resolve_threshold(int64_t thr, int64_t m, int64_t div) {
    return thr < 0 ? m / div : thr;
}

// ======================================================================================
// PageRank  (numa-PageRank.C:64-148 PR_F / PR_Vertex_F / PR_Vertex_Reset, loop :273-335)
//
// The reference pushes p_curr[s]/outdeg(s) from every active s into p_next
// (edgeMapDenseForwardOTHER :177-230) and marks every out-neighbour active;
// then, on the current frontier only, p_next = d*p_next + (1-d)/n and
// p_curr = 0; swap.  Pulled over the CSC this is, per active destination v:
//     sum    = sum_{s in in(v)} contrib[s],  contrib[s] = F[s] ? cur[s]/outdeg(s) : 0
//     nxt[v] = d*sum + (1-d)/n,   F'[v] = (sum > 0),   cur[v] = 0 if v leaves F
// and contrib' = F'[v] ? nxt[v]/outdeg(v) : 0 is produced in the same pass.
// sum > 0 is exactly "some in-neighbour is active": an active s has
// cur[s] >= (1-d)/n > 0.  Frontiers only shrink (N_out(F_t) is a subset of F_t), so an
// inactive vertex is never written (its value freezes, trap P1).
// ======================================================================================
struct PrOp {
    using Acc = float;
    using Ctx = int;
    using Stat = unsigned long long;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = true;
    const uint8_t *F;
    uint8_t *Fn;
    const float *contrib;
    float *contrib_n;
    float *cur;
    float *nxt;
    const int32_t *out_off;
    const float *addc_v;  // per-vertex (1-d)/n_g for many graphs, else null
    float damp, addc;
    __device__ bool active(int v) const { return F[v] != 0; }
    __device__ int ctx(int v) const { return v; }
    __device__ float identity() const { return 0.f; }
    __device__ float combine(float a, float b) const { return a + b; }
    __device__ float edge(int s, int, const int &) const { return contrib[s]; }
    __device__ bool found(float, const int &) const { return false; }
    __device__ Stat finish(int v, float sum, const int &) const {
        const float c = addc_v ? addc_v[v] : addc;
        const float r = damp * sum + c;
        nxt[v] = r;
        const bool any = sum > 0.f;
        const int od = out_off[v + 1] - out_off[v];
        Fn[v] = any ? 1 : 0;
        contrib_n[v] = (any && od > 0) ? r / (float)od : 0.f;
        if (!any) cur[v] = 0.f;  // PR_Vertex_Reset on a vertex leaving the frontier
        return any ? (Stat)od : 0ull;
    }
    __device__ void skip(int v) const {
        Fn[v] = 0;
        contrib_n[v] = 0.f;
    }
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a + b; }
    // spread over kStatSlots addresses so blocks do not serialise on one atomic
    __device__ static void stat_commit(Stat *p, Stat v) {
        atomicAdd(&p[blockIdx.x & (kStatSlots - 1)], v);
    }
    // Tiled pull: contributions in fixed point with 62 fraction bits.  Ranks of
    // one graph sum to at most 1 (no teleport mass is created beyond (1-d)), so
    // every per-vertex sum of contributions is <= 1 < 2^(64-62) and cannot
    // overflow; the truncation of each term is < 2^-62 absolute.
    static constexpr bool kTileWeighted = false;
    __device__ float tile_raw(int s) const { return contrib[s]; }
    __device__ unsigned long long tile_fix(float x) const { return to_fixed<62>(x); }
    __device__ float tile_sum(unsigned long long q) const { return (float)((double)q * 0x1p-62); }
};

__global__ void k_pr_init(int64_t n, const int32_t *out_off, const float *inv_n, float one_over_n,
                          float *cur, float *nxt, uint8_t *F, float *contrib) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const float x = inv_n ? inv_n[v] : one_over_n;
        const int od = out_off[v + 1] - out_off[v];
        cur[v] = x;
        nxt[v] = 0.f;
        F[v] = 1;
        contrib[v] = od > 0 ? x / (float)od : 0.f;
    }
}

__global__ void k_scale(const float *inv, int64_t n, float s, float *out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        out[v] = s * inv[v];
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---- PageRankDelta (numa-PageRankDelta.C) --------------------------------------------
// The reference pushes Delta[s]/outdeg(s) from the active sources into nghSum (PR_F
// :63-85, edgeMap DENSE_FORWARD at threshold 0, :223), then filters every vertex:
//   round 1 (PR_Vertex_F_FirstRound :87-102): Delta = d (p + nghSum) + (1-d)/n - p;
//            p += Delta;  Delta -= 1/n
//   later (PR_Vertex_F :104-117):            Delta = d nghSum;  p += Delta
// keeping a vertex active iff |Delta| > epsilon2 p, and resets nghSum (:119-127).
// Each source's contribution c[s] = Delta[s]/outdeg(s) is kept, zero for an
// inactive source.  A round whose frontier is large is a pull of every vertex's
// in-neighbours' c (the same sums as the push) with the filter as the pull's
// finish: one launch, nghSum never stored.  A round with a small frontier pushes
// the nonzero c with fp64 atomics into nghSum and then filters every vertex.  fp64
// throughout, as the reference.  Statistics of the next frontier: (count << 32) |
// out-degree sum (each below 2^31).
__device__ __forceinline__ unsigned long long prd_filter(int v, double sum, int first, double *p,
                                                         double *cn, const int32_t *out_off,
                                                         double damp, double addc,
                                                         double one_over_n, double eps2) {
#pragma clang fp contract(off)
    const double pv = p[v];
    double delta, pn;
    if (first) {
        delta = damp * (pv + sum) + addc - pv;
        pn = pv + delta;
        delta -= one_over_n;
    } else {
        delta = sum * damp;
        pn = pv + delta;
    }
    if (pn != pv) p[v] = pn;  // an unchanged rank is not stored again
    const bool act = fabs(delta) > eps2 * pn;
    const int od = act ? out_off[v + 1] - out_off[v] : 0;
    cn[v] = (act && od > 0) ? delta / (double)od : 0.0;
    return act ? ((1ull << 32) | (unsigned long long)od) : 0ull;
}

struct PrDeltaOp {
    using Acc = double;
    using Ctx = int;
    using Stat = unsigned long long;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = true;
    const double *c;
    double *cn;
    double *p;
    const int32_t *out_off;
    double damp, addc, one_over_n, eps2;
    int first;
    __device__ bool active(int) const { return true; }
    __device__ int ctx(int v) const { return v; }
    __device__ double identity() const { return 0.0; }
    __device__ double combine(double a, double b) const { return a + b; }
    __device__ double edge(int s, int, const int &) const { return c[s]; }
    __device__ bool found(double, const int &) const { return false; }
    __device__ Stat finish(int v, double sum, const int &) const {
        return prd_filter(v, sum, first, p, cn, out_off, damp, addc, one_over_n, eps2);
    }
    __device__ void skip(int v) const { cn[v] = 0.0; }
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a + b; }
    __device__ static void stat_commit(Stat *p, Stat v) {
        atomicAdd(&p[blockIdx.x & (kStatSlots - 1)], v);
    }
    // tiled pull: the fp64 contributions gathered in source order within a tile and
    // added in LDS as doubles (the cold ones staged as doubles by the pre-pass)
    static constexpr bool kTileWeighted = false;
    __device__ double tile_raw(int s) const { return c[s]; }
    __device__ double tile_fix(double x) const { return x; }
    __device__ double tile_sum(double q) const { return q; }
};

// a push round: the listed sources' contributions added into nghSum (the
// hardware fp64 add, not a compare-and-swap loop)
struct PrDeltaPushOp {
    const double *c;
    double *ngh;
    __device__ bool src_active(int) const { return true; }
    __device__ void push(int s, int d, int) const { unsafeAtomicAdd(&ngh[d], c[s]); }
};

// the filter after a push round (a later round): every vertex, nghSum read and
// cleared for the next round (only where the push added something: a -0.0 left in
// place adds like +0.0), and the bitmap of the next contributions' nonzero entries
// written for the next push round (a wave writes its two words)
__global__ __launch_bounds__(kBlock) void k_prd_filter(int64_t n, double *ngh, double *p,
                                                       double *cn, const int32_t *out_off,
                                                       double damp, double eps2,
                                                       unsigned long long *stats, uint32_t *bits) {
    __shared__ unsigned long long ss[kBlock / 64];
    unsigned long long st = 0;
    const int lane = lane_id();
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; base < n;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = base + lane;
        bool nz = false;
        if (v < n) {
            const double sum = ngh[v];
            if (sum != 0.0) ngh[v] = 0.0;
            st += prd_filter((int)v, sum, 0, p, cn, out_off, damp, 0.0, 0.0, eps2);
            nz = cn[v] != 0.0;
        }
        const unsigned long long m = __ballot(nz);
        if (lane == 0) bits[base >> 5] = (uint32_t)m;
        if (lane == 32 && base + 32 < n) bits[(base >> 5) + 1] = (uint32_t)(m >> 32);
    }
    for (int o = 32; o >= 1; o >>= 1) st += __shfl_xor(st, o, 64);
    if (lane_id() == 0) ss[threadIdx.x >> 6] = st;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kBlock / 64; i++) t += ss[i];
        if (t) atomicAdd(&stats[blockIdx.x & (kStatSlots - 1)], t);
    }
}

// bitmap of the sources with a nonzero contribution (a wave writes its two words)
__global__ void k_prd_bits(const double *c, int64_t n, uint32_t *bits) {
    const int lane = lane_id();
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; base < n;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = base + lane;
        const unsigned long long m = __ballot(v < n && c[v] != 0.0);
        if (lane == 0) bits[base >> 5] = (uint32_t)m;
        if (lane == 32) bits[(base >> 5) + 1] = (uint32_t)(m >> 32);
    }
}

// p = 0 and every vertex active with Delta = 1/n (numa-PageRankDelta.C:280-284)
__global__ void k_prd_init(int64_t n, const int32_t *out_off, double one_over_n, double *p,
                           double *c, double *ngh) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int od = out_off[v + 1] - out_off[v];
        p[v] = 0.0;
        c[v] = od > 0 ? one_over_n / (double)od : 0.0;
        ngh[v] = 0.0;
    }
}

struct PrBuffers {
    float *a, *b, *c0, *c1;
    uint8_t *f0, *f1;
    unsigned long long *stats;  // per iteration sum of out-degrees of F_{t+1}
};

static int pr_buffers(pg_graph *g, int iters, PrBuffers *B) {
    const size_t n = (size_t)g->n;
    PG_TRY(workspace_t(g, "pr_a", n, &B->a));
    PG_TRY(workspace_t(g, "pr_b", n, &B->b));
    PG_TRY(workspace_t(g, "pr_c0", n, &B->c0));
    PG_TRY(workspace_t(g, "pr_c1", n, &B->c1));
    PG_TRY(workspace_t(g, "pr_f0", n, &B->f0));
    PG_TRY(workspace_t(g, "pr_f1", n, &B->f1));
    PG_TRY(workspace_t(g, "pr_stats", ((size_t)iters + 1) * kStatSlots, &B->stats));
    return PG_OK;
}

// Enqueues the whole solve on the library stream (no host synchronisation, no
// allocation once the workspace exists: safe inside a stream capture).  On return
// *result points at the buffer that will hold p_curr.
int pagerank_enqueue(pg_graph *g, int iters, double damping, KernelTimer *kt, float **result,
                     unsigned long long **stats, bool serial_cold) {
    PrBuffers B;
    PG_TRY(pr_buffers(g, iters, &B));
    const int64_t n = g->n;
    PG_LAUNCH_K(k_pr_init, grid_for(n), n, g->out_off, g->inv_n, (float)(1.0 / (double)n), B.a,
                B.b, B.f0, B.c0);
    PG_LAUNCH_K(k_zero_u64, (unsigned)(((iters + 1) * kStatSlots + kBlock - 1) / kBlock), B.stats,
                (iters + 1) * kStatSlots);
    PrOp op{};
    op.out_off = g->out_off;
    op.addc_v = nullptr;
    op.damp = (float)damping;
    op.addc = (float)((1.0 - damping) * (1.0 / (double)n));
    float *inv_n_addc = nullptr;
    if (g->inv_n) {
        PG_TRY(workspace_t(g, "pr_addc", (size_t)n, &inv_n_addc));
        op.addc_v = inv_n_addc;
    }
    if (inv_n_addc)  // (1-d)/n_g per vertex
        PG_LAUNCH_K(k_scale, grid_for(n), g->inv_n, n, (float)(1.0 - damping), inv_n_addc);
    float *cur = B.a, *nxt = B.b, *c = B.c0, *cn = B.c1;
    uint8_t *F = B.f0, *Fn = B.f1;
    // the tiles are built before any capture (pagerank_prepare); inside one only
    // a built set is used
    const bool tiled = g->use_tiles && g->tiles.built && g->tiles.light_begin == g->in_sched.light_begin &&
                       g->tiles.light_end == g->in_sched.light_end;
    // the next iteration's cold pre-pass runs beside this one's finishing pass
    const bool ahead = tiled && g->tiles.cold_overlap && g->tiles.cold.nparts > 0 && !serial_cold;
    for (int t = 0; t < iters; t++) {
        op.F = F;
        op.Fn = Fn;
        op.contrib = c;
        op.contrib_n = cn;
        op.cur = cur;
        op.nxt = nxt;
        PrOp next = op;  // its cold pre-pass reads the contributions written now
        next.contrib = cn;
        const bool pre = ahead && t + 1 < iters;
        if (kt) PG_TRY(kt->begin());
        if (tiled)  // the zero-in-degree vertices are settled after two iterations
            PG_TRY(launch_pull_tiled(op, g->in_sched, g->tiles, &B.stats[(size_t)t * kStatSlots],
                                     t >= 2, pre ? &next : nullptr, ahead && t > 0, true));
        else
            PG_TRY(launch_pull(op, g->in_sched, &B.stats[(size_t)t * kStatSlots]));
        if (kt) PG_TRY(kt->end());
        std::swap(cur, nxt);
        std::swap(F, Fn);
        std::swap(c, cn);
    }
    *result = cur;
    if (stats) *stats = B.stats;
    return PG_OK;
}

// Runs the iterations; on return *result points at the buffer holding p_curr.
int pagerank_run(pg_graph *g, int iters, double damping, uint32_t flags, pg_stats *st,
                 float **result) {
    if (g->use_tiles) PG_TRY(ensure_tiles(g));  // once per schedule, before any capture
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    unsigned long long *stats = nullptr;
    PG_TRY(pagerank_enqueue(g, iters, damping, &kt, result, &stats,
                            (flags & PG_SERIAL_COLD) != 0));
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    PrBuffers B;
    B.stats = stats;
    if (st) {
        std::vector<unsigned long long> hs(((size_t)iters + 1) * kStatSlots), he((size_t)iters + 1);
        PG_CHECK_HIP(hipMemcpy(hs.data(), B.stats, 8 * hs.size(), hipMemcpyDeviceToHost));
        for (int t = 0; t <= iters; t++)
            for (int k = 0; k < kStatSlots; k++) he[t] += hs[(size_t)t * kStatSlots + k];
        st->rounds = iters;
        st->dense_rounds = iters;
        int64_t e = iters > 0 ? g->m : 0;  // F_1 = V
        for (int t = 0; t + 1 < iters; t++) e += (int64_t)he[t];
        st->edges = e;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    return PG_OK;
}

// ======================================================================================
// SpMV (numa-SPMV.C:56-108 SPMV_F / SPMV_Vertex_Reset, loop :168-193): frontier All is
// never switched, so every iteration is y[d] = sum_{s->d} x[s] w(s,d) over all vertices.
//
// The reference iterates in fp64, and its values leave fp32's range (integer weights
// reach 1e42 in 10 iterations on a scale-16 rMat).  The fp32 iteration therefore
// carries one power-of-two exponent per iteration, shared by all vertices: iteration k
// stores y_k * 2^-E_k with E_k chosen from the previous maximum (|y| ~ 2^64).  Scaling
// by a power of two is exact, so every product and sum rounds exactly as unscaled
// fp32 would; only the range grows.  The result is returned as fp64 y * 2^E.
// ======================================================================================
struct SpmvOp {
    using Acc = float;
    using Ctx = int;
    using Stat = unsigned;  // bits of max |y| (non-negative floats order as integers)
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = true;
    const float *x;
    const float *w;  // CSC order
    float *y;
    const int32_t *shift;   // this iteration's exponent shift
    const int32_t *tscale;  // tiled pull: exponent k of this iteration's fixed point
    // Tiled pull: a term x[s]·w_e is scaled by 2^k and accumulated as a signed
    // (two's complement) 64-bit fixed-point number with 32 fraction bits.  k is
    // chosen from max|x| and the largest in-weight sum W (k_spmv_tscale) so that
    // max|x|·W·2^k <= 2^28: no sum or partial sum reaches 2^61.  Each term is
    // truncated by < 2^-32 in scaled units, i.e. < 2^-60 of the largest possible
    // sum; the sum itself is converted to fp32 once.
    static constexpr bool kTileWeighted = true;
    __device__ float tile_raw(int s) const { return x[s]; }
    __device__ unsigned long long tile_fix(float p) const {
        const unsigned long long q = to_fixed<32>(ldexpf(fabsf(p), *tscale));
        return p < 0.f ? (unsigned long long)(-(long long)q) : q;
    }
    __device__ float tile_sum(unsigned long long q) const {
        return ldexpf((float)(long long)q, -32 - *tscale);
    }
    __device__ bool active(int) const { return true; }
    __device__ int ctx(int v) const { return v; }
    __device__ float identity() const { return 0.f; }
    __device__ float combine(float a, float b) const { return a + b; }
    __device__ float edge(int s, int e, const int &) const { return x[s] * w[e]; }
    __device__ bool found(float, const int &) const { return false; }
    __device__ Stat finish(int v, float sum, const int &) const {
        const float r = ldexpf(sum, -*shift);
        y[v] = r;
        return __float_as_uint(fabsf(r));
    }
    __device__ void skip(int) const {}
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a > b ? a : b; }
    __device__ static void stat_commit(Stat *p, Stat v) { atomicMax(p, v); }
};

// shift[k+1] from the maximum of iteration k; total[0] accumulates the exponent
__global__ void k_spmv_shift(const unsigned *maxbits, int32_t *shift_next, long long *total,
                             int32_t this_shift) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const float mx = __uint_as_float(*maxbits);
        int s = 0;
        if (mx > 0.f && isfinite(mx)) s = ilogbf(mx) - 64;
        *shift_next = s;
        total[0] += this_shift;
    }
}

// the tile scale exponent of an iteration whose input has max |x| = *maxbits (as
// float bits) on a graph whose largest in-weight sum is at most *wmax_bits
__global__ void k_spmv_tscale(const unsigned *maxbits, const unsigned *wmax_bits, int32_t *k) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const float mx = __uint_as_float(*maxbits), wm = __uint_as_float(*wmax_bits);
        int s = 0;
        if (mx > 0.f && isfinite(mx) && wm > 0.f && isfinite(wm))
            s = 28 - (ilogbf(mx) + 1) - (ilogbf(wm) + 1);
        *k = s;
    }
}

// largest sum of |w| over a vertex's in-edges (fp32 sums, times 1.5 for their
// rounding), as float bits in *out (non-negative floats order as integers)
__global__ void k_in_weight_sum_max(const int32_t *in_off, const float *w, int64_t n,
                                    unsigned *out) {
    unsigned m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int e = in_off[v]; e < in_off[v + 1]; e++) s += fabsf(w[e]);
        m = max(m, __float_as_uint(s * 1.5f));
    }
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor(m, o, 64));
    if (lane_id() == 0) atomicMax(out, m);
}

__global__ void k_absmax_bits(const float *x, int64_t n, unsigned *out) {
    unsigned m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fabsf(x[i])));
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor(m, o, 64));
    if (lane_id() == 0) atomicMax(out, m);
}

__global__ void k_ldexp_inplace(float *x, int64_t n, const int32_t *shift) {
    const int s = *shift;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = ldexpf(x[i], -s);
}

__global__ void k_to_f64_scaled(const float *a, const int32_t *perm, int64_t n,
                                const long long *total, double *out) {
    const int e = (int)total[0];
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x)
        out[u] = ldexp((double)a[perm[u]], e);
}

__global__ void k_i32_to_f32(const int32_t *a, int64_t m, float *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)a[i];
}

static int spmv_weights(pg_graph *g, const float **w) {
    if (g->w_dtype == PG_W_F32) {
        *w = g->in_w_f32;
        return PG_OK;
    }
    PG_REQUIRE(g->w_dtype == PG_W_I32, PG_ERR_STATE, "SpMV needs a weighted graph");
    float *wf = nullptr;
    auto it = g->ws.find("spmv_wf");
    const bool have = it != g->ws.end();
    PG_TRY(workspace_t(g, "spmv_wf", (size_t)g->m, &wf));
    if (!have) PG_LAUNCH_K(k_i32_to_f32, grid_for(g->m), g->in_w_i32, g->m, wf);
    *w = wf;
    return PG_OK;
}

// ======================================================================================
// Belief propagation (numa-BP.C:89-140 BP_F / BP_Vertex_Reset, push :179-230, loop
// :280-316).  For an edge e = s->d of an active s the reference multiplies
//     P_next[d][k] *= b_e[k],  b_e[k] = sum_l phi_d[l] W_e[k][l] P_cur[d][l],
// after P_next = (1,1) on the current frontier, and marks d active.  Pulled:
//     P_next[d] = prod over active in-neighbours of b_e   (F'[d] = any active).
// b_e does not depend on s (the reference's message uses the destination's
// potential and product), so q_l = phi_d[l] P_cur[d][l] is formed once per d.
// ======================================================================================
template <class T>
struct BpAcc {
    T p0, p1;
    int any;
    int pad;
};
template <class T>
__device__ __forceinline__ BpAcc<T> shfl_xor_t(BpAcc<T> v, int m) {
    BpAcc<T> r;
    r.p0 = __shfl_xor(v.p0, m, 64);
    r.p1 = __shfl_xor(v.p1, m, 64);
    r.any = __shfl_xor(v.any, m, 64);
    r.pad = 0;
    return r;
}

template <class T>
struct Vec2 {
    T x, y;
};

// T = double: the product state and the messages are fp64 (default); T = float:
// fp32 throughout (PG_BP_FP32).  Potentials are fp32 inputs either way
// (numa-BP.C:59-73 stores them as float).
//
// The frontier of the reference's loop depends on the graph alone: F_0 = V and
// F_{t+1} = the out-neighbours of F_t (every vertex the push reaches is set,
// numa-BP.C:122-126, whatever the values).  So F_{t+1} is a subset of F_t, and a
// vertex v leaves it for good at a fixed iteration D[v] (v is in F_t iff t < D[v]).
// kDeath: the per-edge source test reads D of the edge's source from a byte per
// CSC edge (de, 255 = never leaves), streamed beside the potentials, instead of
// looking the source up in a bitmap of F_t (a dependent random load per edge).
// The source id itself is not needed: the message uses the destination's state.
template <class T, bool kDeath>
struct BpOp {
    using Acc = BpAcc<T>;
    using Ctx = Vec2<T>;  // q_l = phi_d[l] * P_cur[d][l]
    using Stat = unsigned long long;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = true;  // sum of out-degrees of the next frontier
    static constexpr bool kEdgeToken = kDeath;
    const uint8_t *F;
    uint8_t *Fn;
    const float4 *pot;   // CSC order, W_e[0][0], W_e[0][1], W_e[1][0], W_e[1][1]
    const float2 *vpot;  // phi
    const Vec2<T> *cur;
    Vec2<T> *nxt;
    const int32_t *out_off;
    const uint32_t *Fb;  // !kDeath: F as a bitmap (n/8 bytes)
    const uint8_t *de;   // kDeath: D of the source of each CSC edge
    int t;               // kDeath: the iteration
    __device__ bool active(int v) const { return F[v] != 0; }
    __device__ Ctx ctx(int v) const {
        const float2 phi = vpot[v];
        const Vec2<T> p = cur[v];
        return Ctx{(T)phi.x * p.x, (T)phi.y * p.y};
    }
    __device__ Acc identity() const { return Acc{T(1), T(1), 0, 0}; }
    __device__ Acc combine(Acc a, Acc b) const {
        return Acc{a.p0 * b.p0, a.p1 * b.p1, a.any | b.any, 0};
    }
    __device__ int edge_token(int e) const {
        return (int)__builtin_nontemporal_load(de + e);
    }
    // tok: D of the source (kDeath) or the source id
    __device__ bool src_active(int tok) const {
        if constexpr (kDeath) return tok == 255 || tok > t;
        else return bit_test(Fb, tok);
    }
    // b_e[k] = sum_l W_e[k][l] q_l (numa-BP.C:118-121)
    __device__ void message(float4 w, const Ctx &q, T &b0, T &b1) const {
        b0 = (T)w.x * q.x + (T)w.y * q.y;
        b1 = (T)w.z * q.x + (T)w.w * q.y;
    }
    // the potential is loaded before the frontier test so that its (streamed,
    // coalesced) load is not held behind the test's load (a non-temporal load of
    // it measured 9% slower)
    __device__ Acc edge(int tok, int e, const Ctx &q) const {
        const float4 w = pot[e];
        if (!src_active(tok)) return identity();
        Acc a{T(1), T(1), 1, 0};
        message(w, q, a.p0, a.p1);
        return a;
    }
    __device__ bool found(Acc, const Ctx &) const { return false; }
    // v's product and next-frontier flag; the statistic is its out-degree if set
    __device__ Stat finish_deg(int v, Acc a, int outdeg) const {
        nxt[v] = Vec2<T>{a.p0, a.p1};
        Fn[v] = a.any ? 1 : 0;
        return a.any ? (Stat)outdeg : 0ull;
    }
    __device__ Stat finish(int v, Acc a, const Ctx &) const {
        return finish_deg(v, a, a.any ? out_off[v + 1] - out_off[v] : 0);
    }
    __device__ void skip(int v) const { Fn[v] = 0; }
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a + b; }
    __device__ static void stat_commit(Stat *p, Stat v) {
        atomicAdd(&p[blockIdx.x & (kStatSlots - 1)], v);
    }
};

// bits[w] = the 32 flags f[32w .. 32w+31] (zero past n): a wave reads 64
// consecutive flag bytes (one coalesced load) and its ballot is two words
__global__ void k_flags_to_bits(const uint8_t *f, int64_t n, uint32_t *bits) {
    const int64_t nw = (n + 31) / 32;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v0 = wave * 64; v0 < n; v0 += nwaves * 64) {  // uniform per wave
        const int64_t v = v0 + lane;
        const unsigned long long b = __ballot(v < n && f[v] != 0);
        const int64_t w = v0 / 32;
        if (lane == 0) bits[w] = (uint32_t)b;
        if (lane == 1 && w + 1 < nw) bits[w + 1] = (uint32_t)(b >> 32);
    }
}

__global__ void k_fill_u8(uint8_t *p, uint8_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ======================================================================================
// BFS (numa-BFS.C:59-79 BFS_F: cond = unvisited, updateAtomic = CAS(parent,-1,s);
// loop :258-296).  Dense rounds pull bottom-up (edgeMapDenseNoRep :113-159: the first
// in-neighbour found in the frontier becomes the parent, then break); sparse rounds push
// top-down from the frontier list (edgeMapSparseV3).  Levels are the round numbers.
// ======================================================================================
// The frontier of round r is {v : level[v] == r - 1}, held as a bitmap (n/8 bytes,
// small enough to stay in L2 while a bottom-up round probes it).
//
// Bottom-up round.  On rMat graphs almost every vertex found by a bottom-up round
// finds its parent at the first or second in-neighbour probed (the in-lists start
// with the highest-degree sources, which the frontier holds), so the round is
// per-vertex overhead, not edge traffic; one workgroup per 256 vertices made it
// dispatch-bound.  Here each wave walks chunks of 64 consecutive vertices with a
// grid-sized stride, the in-list probed in order so that the lowest hit position
// wins (the first frontier in-neighbour in in-list order).  The round's winners are
// also the next frontier: the chunk's ballot is its two bitmap words, and their
// number and out-degree sum go to spread counters.  Vertices without in-edges (the
// zero-degree tail of the in-degree order) are not visited.
// Almost every hit is the first in-neighbour, so that one is kept for every vertex
// in a separate array (first_in, 4 bytes per vertex, -1 without in-edges): the
// common case reads it, the visited bits and the out-offsets with coalesced loads
// and probes one bitmap word.  Only a miss reads the vertex's offsets and the rest
// of its in-list.  A wave takes J chunks of 64 at once and issues each stage's
// loads for all of them before waiting.
constexpr int kBuChunks = 4;
constexpr int kBuLaneScan = 64;
// a round's direction, chosen on the device (k_bfs_round_begin): every kernel of a
// round is enqueued and those of the other direction return at once
constexpr unsigned long long kBfsModeDone = 0, kBfsModeBU = 1, kBfsModeTD = 2;
__global__ void k_first_in(const int32_t *in_off, const int32_t *in_adj, int64_t n,
                           int32_t *first_in) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        first_in[v] = in_off[v] < in_off[v + 1] ? in_adj[in_off[v]] : -1;
}

// the statistics of a round's claims (count, out-degree sum), summed over the
// block and added to slot pair blockIdx mod 64 of the round's region; every thread
// of the block calls it
__device__ __forceinline__ void bfs_flush_stats(unsigned long long cnt, unsigned long long deg,
                                                unsigned long long *slots) {
    __shared__ unsigned long long sc[kBlock / 64], sd[kBlock / 64];
    for (int o = 32; o >= 1; o >>= 1) {
        cnt += __shfl_xor(cnt, o, 64);
        deg += __shfl_xor(deg, o, 64);
    }
    if (lane_id() == 0) {
        sc[threadIdx.x >> 6] = cnt;
        sd[threadIdx.x >> 6] = deg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0, d = 0;
        for (int i = 0; i < kBlock / 64; i++) {
            c += sc[i];
            d += sd[i];
        }
        if (c) {
            unsigned long long *p = slots + 2 * (blockIdx.x & 63);
            atomicAdd(&p[0], c);
            atomicAdd(&p[1], d);
        }
    }
}

// (the body of a bottom-up round, wave-uniform; the winners' count and out-degree
// sum accumulate in cnt and deg)
template <int J>
__device__ __forceinline__ void bfs_bottom_up(
    const int32_t *__restrict__ in_off, const int32_t *__restrict__ in_adj,
    const int32_t *__restrict__ first_in, const int32_t *__restrict__ out_off,
    const uint32_t *__restrict__ F, uint32_t *__restrict__ V, int32_t *level, int32_t *parent,
    int32_t round, int64_t nbody, uint32_t *__restrict__ Fn, unsigned long long &cnt,
    unsigned long long &deg) {
    const int lane = lane_id();
    const int64_t nchunks = (nbody + 64 * J - 1) / (64 * J);
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
    for (int64_t ch = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); ch < nchunks;
         ch += nwaves) {
        // visited word of this lane's vertex (all ones past the body)
        uint32_t vw[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int64_t v = (ch * J + j) * 64 + lane;
            vw[j] = v < nbody ? V[v >> 5] : 0xFFFFFFFFu;
        }
        int ps[J], od[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int64_t v = (ch * J + j) * 64 + lane;
            const bool open = !((vw[j] >> (lane & 31)) & 1u);
            ps[j] = open ? first_in[v] : -1;
            od[j] = open ? out_off[v + 1] - out_off[v] : 0;
        }
        bool miss[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const bool open = ps[j] >= 0;
            const bool h = open && bit_test(F, ps[j]);
            miss[j] = open && !h;
            if (!h) ps[j] = -1;
        }
        // a miss scans the rest of its in-list: four ids per step by its own lane
        // (all lanes of all J chunks stepping together) while at most kBuLaneScan
        // remain, else by the whole wave, 256 per step, one vertex at a time (an
        // edge-balanced scan of all the wave's misses measured slower: its prefix
        // sums cost every chunk, and it cannot stop a list at its first hit)
        int cur[J], end[J];
        bool act[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int64_t v = (ch * J + j) * 64 + lane;
            cur[j] = miss[j] ? in_off[v] + 1 : 0;
            end[j] = miss[j] ? in_off[v + 1] : 0;
            act[j] = miss[j] && end[j] > cur[j] && end[j] - cur[j] <= kBuLaneScan;
        }
        while (true) {
            bool any = false;
#pragma unroll
            for (int j = 0; j < J; j++) any = any || act[j];
            if (!__ballot(any)) break;
#pragma unroll
            for (int j = 0; j < J; j++) {
                if (!act[j]) continue;
                const int k4 = min(end[j] - cur[j], 4);
                int s[4];
#pragma unroll
                for (int k = 0; k < 4; k++) s[k] = k < k4 ? ld_stream(in_adj + cur[j] + k) : -1;
#pragma unroll
                for (int k = 3; k >= 0; k--)
                    if (s[k] >= 0 && bit_test(F, s[k])) ps[j] = s[k];
                cur[j] += 4;
                act[j] = ps[j] < 0 && cur[j] < end[j];
            }
        }
#pragma unroll
        for (int j = 0; j < J; j++) {
            unsigned long long rest = __ballot(miss[j] && end[j] - cur[j] > kBuLaneScan);
            while (rest) {
                const int l = __ffsll((long long)rest) - 1;
                rest &= rest - 1;
                const int b = __shfl(cur[j], l, 64), e = __shfl(end[j], l, 64);
                for (int base = b; base < e; base += 256) {
                    int sv[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int ee = base + 64 * k + lane;
                        sv[k] = ee < e ? ld_stream(in_adj + ee) : -1;
                    }
                    unsigned long long hm[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) hm[k] = __ballot(sv[k] >= 0 && bit_test(F, sv[k]));
                    int k = 0;
                    while (k < 4 && !hm[k]) k++;
                    if (k < 4) {
                        const int sf = __shfl(sv[k], __ffsll((long long)hm[k]) - 1, 64);
                        if (lane == l) ps[j] = sf;
                        break;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int64_t c64 = ch * J + j;
            const int64_t v = c64 * 64 + lane;
            const bool won = ps[j] >= 0;
            if (won) {
                parent[v] = ps[j];
                level[v] = round;
                cnt++;
                deg += (unsigned long long)od[j];
            }
            // the chunk's two words belong to this wave: the next frontier's are
            // written whole, the visited ones gain the winners
            const unsigned long long wm = __ballot(won);
            if (c64 * 64 < nbody && (lane & 31) == 0) {
                const uint32_t w = lane == 0 ? (uint32_t)wm : (uint32_t)(wm >> 32);
                Fn[2 * c64 + (lane >> 5)] = w;
                if (w) V[2 * c64 + (lane >> 5)] = vw[j] | w;
            }
        }
    }
}

// BFS counters and control: one region of kBfsRegion words per round (round r uses
// region r mod kBfsRegions; region 0 holds the source's statistics).  Words 0..127
// are 64 spread (count, out-degree) pairs of the frontier the round produces, 130
// and 131 the lengths of a top-down round's queues, kBfsCtlMode.. the round's
// direction and the running state after its decision (the direction policy's
// flag, totals, round counts, done).
constexpr int kBfsRegion = 256;
constexpr int kBfsRegions = 64;
constexpr int kBfsCtlMode = 132, kBfsCtlBu = 133, kBfsCtlEdges = 134, kBfsCtlActive = 135,
              kBfsCtlRounds = 136, kBfsCtlDense = 137, kBfsCtlDone = 138, kBfsCtlEnd = 139;
// rounds per enqueued (and captured) batch: the first batch of a search runs as
// many rounds as the graph's last search took (kBfsFirstBatch before any), the
// later ones kBfsAhead each; at most kBfsMaxBatch
constexpr int kBfsAhead = 4, kBfsFirstBatch = 8, kBfsMaxBatch = 32;
// direction policy of the default threshold (the top-down/bottom-up switch of
// direction-optimising BFS with the switch points as ratios)
constexpr int64_t kBfsAlpha = 14, kBfsBeta = 24;

// A batch of rounds is enqueued (captured once, replayed) with no host decision
// in between.  Each kernel finds its round as base[0] + k + 1 (base[0] = rounds of
// the batches before, advanced by k_bfs_advance at the end of a batch; k = the
// round's place in the batch) and its counters in the round's region.  Three
// frontier bitmaps rotate: round q reads F(q), writes N(q) and clears C(q), the
// bitmap round q+1 writes (it held the frontier of round q-1).
struct BfsRound {
    unsigned long long *regions;
    const unsigned long long *base;
    uint32_t *bits0, *bits1, *bits2;  // round 1 reads bits0 (the source)
    int k;
    __device__ int64_t r() const { return (int64_t)base[0] + k + 1; }
    __device__ unsigned long long *region(int64_t q) const {
        return regions + (size_t)(q % kBfsRegions) * kBfsRegion;
    }
    __device__ uint32_t *pick(int64_t i) const {
        return i == 0 ? bits0 : (i == 1 ? bits1 : bits2);
    }
    __device__ uint32_t *F(int64_t q) const { return pick((q + 2) % 3); }
    __device__ uint32_t *N(int64_t q) const { return pick(q % 3); }
    __device__ uint32_t *C(int64_t q) const { return pick((q + 1) % 3); }
};

// Top-down claims of edges (v, d), made by the whole wave together (every lane
// calls it) for J slices of 64 edges (d < 0: no edge), each stage's loads and
// atomics issued for every slice before the first wait.  V (n/8 bytes,
// L2-resident) screens out the vertices reached before; an OR on V's word that
// returns it picks the one claimer of a destination, which stores level and
// parent, sets the destination's bit in the next frontier and counts it with its
// out-degree for the round's statistics.
// Which frontier vertex becomes the parent is unspecified, as in the reference,
// whose claim is a CAS of the parent from -1.  Parents need no initialisation: the
// parent of an unreached vertex is set to -1 on output.
// The ORs are made per run of consecutive lanes whose destinations share a V word:
// neighbour lists are sorted and the ids are in degree order, so a hub's edges
// reach dense id ranges, and one OR per lane would queue up to 32 atomics on one
// word.  The run's first lane ORs the run's bits into V and the bits it won into
// the next frontier; a lane wins if its bit was clear before and no earlier lane
// of its run (a slice's runs) holds the same destination; two slices holding one
// destination are told apart by their ORs.
template <int J>
__device__ __forceinline__ void bfs_claim_wave(const int (&v)[J], const int (&d)[J], uint32_t *V,
                                               int32_t *level, int32_t *parent, uint32_t *Fn,
                                               int32_t round, const int32_t *__restrict__ out_off,
                                               unsigned long long &cnt,
                                               unsigned long long &deg) {
    const int lane = lane_id();
    uint32_t b[J], vw[J];
    int w[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        b[j] = d[j] >= 0 ? 1u << (d[j] & 31) : 0u;
        w[j] = d[j] >= 0 ? d[j] >> 5 : 0;
        vw[j] = d[j] >= 0 ? V[w[j]] : ~0u;
    }
    bool want[J], any = false;
#pragma unroll
    for (int j = 0; j < J; j++) {
        want[j] = d[j] >= 0 && !(vw[j] & b[j]);
        any = any || want[j];
    }
    if (!__ballot(any)) return;  // wave-uniform
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
    bool head[J];
    int start[J];
    uint32_t suf[J], before[J], old[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        // runs: a lane without a claim is a run of its own
        const int key = want[j] ? w[j] : -1 - lane;
        const int key_prev = __shfl_up(key, 1, 64);  // with every lane active
        head[j] = lane == 0 || key_prev != key;
        const unsigned long long H = __ballot(head[j]);
        start[j] = 63 - __clzll((long long)(H & upto));
        const unsigned long long after = H & ~upto;
        const int end = after ? __ffsll((long long)after) - 2 : 63;
        uint32_t s = want[j] ? b[j] : 0u, inc = s;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t ys = __shfl_down(s, o, 64);
            const uint32_t yp = __shfl_up(inc, o, 64);
            if (lane + o <= end) s |= ys;         // OR of lanes lane..end of the run
            if (lane - o >= start[j]) inc |= yp;  // OR of lanes start..lane
        }
        suf[j] = s;
        const uint32_t bp = __shfl_up(inc, 1, 64);
        before[j] = lane == start[j] ? 0u : bp;
    }
#pragma unroll
    for (int j = 0; j < J; j++) old[j] = head[j] && want[j] ? atomicOr(&V[w[j]], suf[j]) : 0u;
#pragma unroll
    for (int j = 0; j < J; j++) {
        if (head[j] && want[j] && (suf[j] & ~old[j])) atomicOr(&Fn[w[j]], suf[j] & ~old[j]);
        old[j] = __shfl(old[j], start[j], 64);
    }
#pragma unroll
    for (int j = 0; j < J; j++) {
        if (want[j] && !(old[j] & b[j]) && !(before[j] & b[j])) {
            level[d[j]] = round;
            parent[d[j]] = v[j];
            cnt++;
            deg += (unsigned long long)(out_off[d[j] + 1] - out_off[d[j]]);
        }
    }
}

// Top-down round over the frontier bitmap: no list and no scan, so the grid does
// not depend on the frontier and the round can be enqueued before its frontier is
// known.  A wave takes 64 bitmap words, a lane one word; every pass of the
// wave-uniform loop takes the next set bit of each lane's word.  A vertex of
// out-degree at most kBfsLaneDeg is pushed by its own lane; a larger one is queued
// (qbig up to kBfsWaveDeg, qhuge above) for k_bfs_push_queue, which gives a queued
// vertex a whole wave and a huge one the whole grid, edge-balanced, so a hub in the
// frontier is not one lane's (or one wave's) serial loop.
constexpr int kBfsLaneDeg = 8;
constexpr int kBfsWaveDeg = 1024;
// slices of 64 edges a wave claims at once: the small out-lists pushed by their
// lanes (bfs_push_bits), a queued big vertex's, the huge vertices' batches
constexpr int kBfsLaneJ = 1, kBfsBigJ = 2, kBfsHugeJ = 2;

__device__ __forceinline__ void bfs_push_bits(
    const uint32_t *__restrict__ F, int64_t nw, const int32_t *__restrict__ out_off,
    const int32_t *__restrict__ out_adj, uint32_t *V, int32_t *level, int32_t *parent,
    uint32_t *Fn, int32_t round, int32_t *qbig, int32_t *qhuge, unsigned long long *qctr,
    unsigned long long &cnt, unsigned long long &deg) {
    const int lane = lane_id();
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
    for (int64_t base = ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 64;
         base < nw; base += nwaves * 64) {
        uint32_t x = base + lane < nw ? F[base + lane] : 0u;
        while (__ballot(x != 0u)) {
            int v = -1, dg = 0, beg = 0;
            if (x) {
                v = (int)((base + lane) * 32 + __ffs(x) - 1);
                x &= x - 1;
                beg = out_off[v];
                dg = out_off[v + 1] - beg;
            }
            // a small out-list is pushed by its own lane, the wave stepping together
            const int own = v >= 0 && dg <= kBfsLaneDeg ? dg : 0;
            for (int k = 0; __ballot(k < own); k += kBfsLaneJ) {
                int vs[kBfsLaneJ], ds[kBfsLaneJ];
#pragma unroll
                for (int j = 0; j < kBfsLaneJ; j++) {
                    vs[j] = v;
                    ds[j] = k + j < own ? out_adj[beg + k + j] : -1;
                }
                bfs_claim_wave<kBfsLaneJ>(vs, ds, V, level, parent, Fn, round, out_off, cnt, deg);
            }
            if (v >= 0 && dg <= kBfsLaneDeg) v = -1;
            agg_append(qbig, &qctr[0], v >= 0 && dg <= kBfsWaveDeg, v);
            agg_append(qhuge, &qctr[1], v >= 0 && dg > kBfsWaveDeg, v);
        }
    }
}

// Round r (its first kernel; k_bfs_push_queue finishes a top-down round).  Every
// block takes the round's decision from region r-1 (the frontier's size c and
// out-degree sum e, the running state): done when c = 0 or already done; else
// threshold >= 0: bottom-up iff c + e >= threshold (P9); threshold < 0: bottom-up
// once e exceeds 1/alpha of the edges not yet reached, top-down again when the
// frontier holds fewer than n/beta vertices.  Block 0 writes the decision and the
// new state into region r and clears the counters of region r+1 (no kernel of
// round r touches it).  Then the grid clears C(r) and runs the round: bottom-up
// over the unvisited body, or the top-down scan of the frontier bitmap.
template <int J>
__global__ __launch_bounds__(kBlock) void k_bfs_round(
    BfsRound R, const int32_t *__restrict__ in_off, const int32_t *__restrict__ in_adj,
    const int32_t *__restrict__ first_in, const int32_t *__restrict__ out_off,
    const int32_t *__restrict__ out_adj, uint32_t *V, int32_t *level, int32_t *parent,
    int32_t *qbig, int32_t *qhuge, int64_t nbody, int64_t nbits, int64_t threshold, int64_t m,
    int64_t n, int64_t alpha, int64_t beta) {
    __shared__ unsigned long long s_mode;
    const int64_t r = R.r();
    const unsigned long long *prev = R.region(r - 1);
    unsigned long long *cur = R.region(r);
    if (threadIdx.x < 64) {
        const int t = threadIdx.x;
        unsigned long long c = prev[2 * t], e = prev[2 * t + 1];
        for (int o = 32; o >= 1; o >>= 1) {
            c += __shfl_xor(c, o, 64);
            e += __shfl_xor(e, o, 64);
        }
        if (t == 0) {
            unsigned long long st[kBfsCtlEnd - kBfsCtlBu];
            for (int i = 0; i < kBfsCtlEnd - kBfsCtlBu; i++) st[i] = prev[kBfsCtlBu + i];
            unsigned long long mode = kBfsModeDone;
            if (!st[kBfsCtlDone - kBfsCtlBu] && c != 0) {
                st[kBfsCtlRounds - kBfsCtlBu] += 1;
                st[kBfsCtlActive - kBfsCtlBu] += c;
                const unsigned long long tot = st[kBfsCtlEdges - kBfsCtlBu] + e;
                st[kBfsCtlEdges - kBfsCtlBu] = tot;
                bool dense;
                if (threshold >= 0) {
                    dense = (int64_t)(c + e) >= threshold;
                } else {
                    const int64_t unexplored = m > (int64_t)tot ? m - (int64_t)tot : 0;
                    dense = st[kBfsCtlBu - kBfsCtlBu] ? (int64_t)c >= n / beta
                                                       : (int64_t)e > unexplored / alpha;
                    st[kBfsCtlBu - kBfsCtlBu] = dense ? 1 : 0;
                }
                if (dense) st[kBfsCtlDense - kBfsCtlBu] += 1;
                mode = dense ? kBfsModeBU : kBfsModeTD;
            } else {
                st[kBfsCtlDone - kBfsCtlBu] = 1;
            }
            s_mode = mode;
            if (blockIdx.x == 0) {
                cur[kBfsCtlMode] = mode;
                for (int i = 0; i < kBfsCtlEnd - kBfsCtlBu; i++) cur[kBfsCtlBu + i] = st[i];
            }
        }
    }
    if (blockIdx.x == 0) {
        unsigned long long *nx = R.region(r + 1);
        for (int i = threadIdx.x; i < kBfsCtlMode; i += blockDim.x) nx[i] = 0ull;
    }
    __syncthreads();
    const unsigned long long mode = s_mode;
    if (mode == kBfsModeDone) return;  // block-uniform
    uint32_t *Cb = R.C(r);
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w <= nbits;
         w += (int64_t)gridDim.x * blockDim.x)
        Cb[w] = 0u;
    unsigned long long cnt = 0, deg = 0;
    if (mode == kBfsModeBU)
        bfs_bottom_up<J>(in_off, in_adj, first_in, out_off, R.F(r), V, level, parent, (int32_t)r,
                         nbody, R.N(r), cnt, deg);
    else
        bfs_push_bits(R.F(r), nbits, out_off, out_adj, V, level, parent, R.N(r), (int32_t)r, qbig,
                      qhuge, cur + 130, cnt, deg);
    bfs_flush_stats(cnt, deg, cur);
}

__global__ void k_bfs_advance(unsigned long long *base, int64_t rounds) {
    if (threadIdx.x == 0 && blockIdx.x == 0) base[0] += (unsigned long long)rounds;
}

// the queued vertices of a top-down round (bfs_push_bits): the huge ones
// edge-balanced over the whole grid (batches of kBlock vertices whose degree prefix
// every block forms in LDS; edge k of a batch goes to thread k mod grid size),
// every big one to one wave, four 64-edge slices in flight per step
__global__ __launch_bounds__(kBlock) void k_bfs_push_queue(
    BfsRound R, const int32_t *__restrict__ qbig, const int32_t *__restrict__ qhuge,
    const int32_t *__restrict__ out_off, const int32_t *__restrict__ out_adj, uint32_t *V,
    int32_t *level, int32_t *parent) {
    __shared__ int32_t s_v[kBlock], s_beg[kBlock];
    __shared__ int64_t s_pre[kBlock + 1];
    __shared__ int64_t s_wsum[kBlock / 64];
    const int64_t rr = R.r();
    unsigned long long *cur = R.region(rr);
    if (cur[kBfsCtlMode] != kBfsModeTD) return;  // block-uniform
    uint32_t *Fn = R.N(rr);
    const int32_t round = (int32_t)rr;
    const unsigned long long *qctr = cur + 130;
    unsigned long long cnt = 0, deg = 0;
    const int64_t nh = (int64_t)qctr[1], nb = (int64_t)qctr[0];
    const int64_t gt = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t gn = (int64_t)gridDim.x * kBlock;
    for (int64_t b0 = 0; b0 < nh; b0 += kBlock) {
        const int nbat = (int)min<int64_t>(kBlock, nh - b0);
        // the batch's degree prefix by a block scan (a wave scan, then the wave sums)
        long long x = 0;
        if ((int)threadIdx.x < nbat) {
            const int v = qhuge[b0 + threadIdx.x];
            s_v[threadIdx.x] = v;
            s_beg[threadIdx.x] = out_off[v];
            x = out_off[v + 1] - out_off[v];
        }
        for (int o = 1; o < 64; o <<= 1) {
            const long long y = __shfl_up(x, o, 64);
            if (lane_id() >= o) x += y;
        }
        if (lane_id() == 63) s_wsum[threadIdx.x >> 6] = x;
        __syncthreads();
        long long before = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); w++) before += s_wsum[w];
        s_pre[threadIdx.x + 1] = before + x;
        if (threadIdx.x == 0) s_pre[0] = 0;
        __syncthreads();
        const int64_t tot = s_pre[nbat];
        // a wave takes kBfsHugeJ consecutive 64-edge slices per step
        for (int64_t k0 = (gt >> 6) * (64 * kBfsHugeJ); k0 < tot; k0 += gn * kBfsHugeJ) {
            int vs[kBfsHugeJ], ds[kBfsHugeJ];
#pragma unroll
            for (int j = 0; j < kBfsHugeJ; j++) {
                const int64_t k = k0 + 64 * j + lane_id();
                int lo = 0, hi = nbat - 1;  // last i with s_pre[i] <= k
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= k) lo = mid; else hi = mid - 1;
                }
                vs[j] = s_v[lo];
                ds[j] = k < tot ? out_adj[s_beg[lo] + (int)(k - s_pre[lo])] : -1;
            }
            bfs_claim_wave<kBfsHugeJ>(vs, ds, V, level, parent, Fn, round, out_off, cnt, deg);
        }
        __syncthreads();
    }
    const int lane = lane_id();
    const int64_t nwaves = gn / 64;
    for (int64_t i = gt / 64; i < nb; i += nwaves) {
        const int v = qbig[i];
        const int beg = out_off[v], end = out_off[v + 1];
        for (int base = beg; base < end; base += 256) {
            int d[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int e = base + 64 * k + lane;
                d[k] = e < end ? out_adj[e] : -1;
            }
#pragma unroll
            for (int k = 0; k < 4; k += kBfsBigJ) {
                if (base + 64 * k >= end) break;  // wave-uniform
                int vs[kBfsBigJ], ds[kBfsBigJ];
#pragma unroll
                for (int j = 0; j < kBfsBigJ; j++) {
                    vs[j] = v;
                    ds[j] = d[k + j];
                }
                bfs_claim_wave<kBfsBigJ>(vs, ds, V, level, parent, Fn, round, out_off, cnt, deg);
            }
        }
    }
    bfs_flush_stats(cnt, deg, cur);
}

// A search's starting state in one pass: levels -1 but the source's 0, the source
// its own parent, the visited bitmap and bitmap 0 (round 1's frontier) holding the
// source alone, bitmaps 1 and 2 clear, every region clear but the source's
// statistics in region 0, the round base 0.  (Each word is written once, with its
// final value, so no ordering among threads is needed.)
__global__ void k_bfs_init(int32_t src, int64_t n, int64_t nw, const int32_t *out_off,
                           int32_t *level, int32_t *parent, uint32_t *V, uint32_t *b0,
                           uint32_t *b1, uint32_t *b2, unsigned long long *regions,
                           unsigned long long *base) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t gs = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = gt; v < n; v += gs) level[v] = v == src ? 0 : -1;
    const int64_t sw = src >> 5;
    const uint32_t sb = 1u << (src & 31);
    for (int64_t w = gt; w < nw; w += gs) {
        const uint32_t x = w == sw ? sb : 0u;
        V[w] = x;
        b0[w] = x;
        b1[w] = 0u;
        b2[w] = 0u;
    }
    for (int64_t i = gt; i < (int64_t)kBfsRegion * kBfsRegions; i += gs)
        regions[i] = i == 0 ? 1ull
                            : (i == 1 ? (unsigned long long)(out_off[src + 1] - out_off[src]) : 0ull);
    if (gt == 0) {
        parent[src] = src;
        base[0] = 0ull;
        base[1] = 0ull;
    }
}

__global__ void k_parent_mask(const int32_t *level, int64_t n, int32_t *parent) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        if (level[v] < 0) parent[v] = -1;
}

// ======================================================================================
// Connected components (numa-Components.C:55-119 CC_F / CC_Vertex_F, edgeMapCustom
// :121-175 with the switch at m/8, loop :233-281).  prev = ID on the frontier; dense
// rounds pull the minimum label of active in-neighbours (edgeMapDenseReduce); sparse
// rounds push writeMin (edgeMapSparseV3).  A vertex joins the next frontier on its
// first decrease in the round, i.e. when the value it replaced equals prev.
// ======================================================================================
struct CcPullOp {
    using Acc = int;
    using Ctx = int;
    using Stat = unsigned long long;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = false;
    const uint32_t *F;
    uint32_t *Fn;
    int32_t *label;
    __device__ bool active(int) const { return true; }
    __device__ int ctx(int v) const { return v; }
    __device__ int identity() const { return 0x7FFFFFFF; }
    __device__ int combine(int a, int b) const { return min(a, b); }
    __device__ int edge(int s, int, const int &) const {
        return bit_test(F, s) ? label[s] : 0x7FFFFFFF;
    }
    __device__ bool found(int, const int &) const { return false; }
    __device__ Stat finish(int v, int mn, const int &) const {
        if (mn < label[v]) {  // the next list is compacted from Fn after the round
            label[v] = mn;
            atomicOr(&Fn[v >> 5], 1u << (v & 31));
        }
        return 0ull;
    }
    __device__ void skip(int) const {}
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a + b; }
    // the tiled pull: the smallest label of an active in-neighbour, in 64-bit LDS
    // minima (an inactive source's all-ones value is not applied)
    static constexpr bool kTileWeighted = false;
    static constexpr int kTileReduce = kReduceMin;
    __device__ uint32_t tile_raw(int s) const {
        return bit_test(F, s) ? (uint32_t)label[s] : 0xFFFFFFFFu;
    }
    __device__ unsigned long long tile_fix(uint32_t x) const {
        return x == 0xFFFFFFFFu ? ~0ull : (unsigned long long)x;
    }
    __device__ int tile_sum(unsigned long long q) const { return q == ~0ull ? 0x7FFFFFFF : (int)q; }
};

struct CcPushOp {
    uint32_t *Fn;
    int32_t *list_n;
    unsigned long long *cnt_n;
    int32_t *label;
    const int32_t *prev;
    bool append;  // else the next list is compacted from Fn after the round
    __device__ bool src_active(int) const { return true; }
    __device__ void push(int s, int d, int) const {
        const int ls = label[s];
        bool act = false;
        if (ls < label[d]) {
            const int old = atomicMin(&label[d], ls);
            act = ls < old && old == prev[d];
            if (act) atomicOr(&Fn[d >> 5], 1u << (d & 31));
        }
        if (append) agg_append(list_n, cnt_n, act, d);
    }
};

__global__ void k_cc_prev(const int32_t *list, const unsigned long long *cnt, const int32_t *label,
                          int32_t *prev) {
    const int64_t c = (int64_t)*cnt;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int v = list[i];
        prev[v] = label[v];
    }
}

__global__ void k_iota_labels(const int32_t *iperm, int64_t n, int32_t *label, int32_t *list) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        label[v] = iperm[v];  // IDs[i] = i in input ids (numa-Components.C:387-391)
        list[v] = (int32_t)v;
    }
}

// ======================================================================================
// SSSP (numa-BellmanFord.C:61-95 BF_F: writeMin(dist[d], dist[s]+w) and activate d,
// loop :166-211, switch at m/20).  Both directions push (edgeMapDenseForward over the
// bitmap, edgeMapSparseV3 over the list).  Activation is "first improvement this round"
// (a new bit in the next frontier), the Ligra protocol, so distances are exactly the
// shortest ones; the reference's reset of Visited on the current frontier (P6) can stop
// short of them and is not reproduced.
// ======================================================================================
template <bool kDense>
struct SsspPushOp {
    const uint32_t *F;
    uint32_t *Fn;
    int32_t *list_n;
    unsigned long long *cnt_n;
    int32_t *dist;
    const int32_t *w;  // CSR order
    bool append;       // else the next list is compacted from Fn after the round
    __device__ bool src_active(int s) const { return kDense ? bit_test(F, s) : true; }
    __device__ void push(int s, int d, int e) const {
        const int nd = dist[s] + w[e];
        bool act = false;
        if (nd < dist[d]) {
            const int old = atomicMin(&dist[d], nd);
            if (nd < old) act = bit_set_new(Fn, d);
        }
        if (append) agg_append(list_n, cnt_n, act, d);
    }
};

// The dense rounds as a pull on the source-sorted tiles: every destination takes
// the smallest dist[s] + w over its active in-neighbours and, when that improves
// dist[d], joins the next frontier.  The distances reached are the same (relaxations
// from the round's frontier, each a true path length); a source's value is read as
// of the round's start for cold sources, possibly lowered within it for hot ones.
struct SsspPullOp {
    using Acc = int;
    using Ctx = int;
    using Stat = unsigned long long;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = false;
    static constexpr int32_t kInf = 0x7FFFFFFF;
    const uint32_t *F;
    uint32_t *Fn;
    int32_t *dist;
    const int32_t *w;  // CSC order (the heavy vertices' pull)
    __device__ bool active(int) const { return true; }
    __device__ int ctx(int v) const { return v; }
    __device__ int identity() const { return kInf; }
    __device__ int combine(int a, int b) const { return min(a, b); }
    __device__ int edge(int s, int e, const int &) const {
        return bit_test(F, s) ? dist[s] + w[e] : kInf;
    }
    __device__ bool found(int, const int &) const { return false; }
    __device__ Stat finish(int v, int mn, const int &) const {
        if (mn < dist[v]) {  // the next list is compacted from Fn after the round
            dist[v] = mn;
            atomicOr(&Fn[v >> 5], 1u << (v & 31));
        }
        return 0ull;
    }
    __device__ void skip(int) const {}
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a + b; }
    // the tiled pull: 64-bit LDS minima of dist[s] + w (an inactive source's
    // all-ones value is not applied); weights come in tile order as exact floats
    static constexpr bool kTileWeighted = true;
    static constexpr int kTileReduce = kReduceMin;
    __device__ uint32_t tile_raw(int s) const {
        return bit_test(F, s) ? (uint32_t)dist[s] : 0xFFFFFFFFu;
    }
    __device__ unsigned long long tile_fix(uint32_t x) const {
        return x == 0xFFFFFFFFu ? ~0ull : (unsigned long long)x;
    }
    __device__ unsigned long long tile_fix_w(uint32_t x, float wv) const {
        return x == 0xFFFFFFFFu ? ~0ull : (unsigned long long)x + (unsigned long long)wv;
    }
    __device__ int tile_sum(unsigned long long q) const { return q == ~0ull ? kInf : (int)q; }
};

// ---- traversal state ---------------------------------------------------------------
// ctr[0] = count, ctr[1] = out-edges; ctr[4 + 2k], ctr[5 + 2k] (k < kCtrSlots) are
// spread partial sums of the same, added by k_level_frontier's blocks (thousands of
// atomics on one address serialise)
constexpr int kCtrSlots = 64;
constexpr int kCtrWords = 4 + 2 * kCtrSlots;
struct Frontier {
    uint32_t *bits;
    int32_t *list;
    unsigned long long *ctr;
};

static int frontiers(pg_graph *g, Frontier *a, Frontier *b) {
    const size_t nw = (size_t)((g->n + 31) / 32) + 1;
    PG_TRY(workspace_t(g, "fr_bits0", nw, &a->bits));
    PG_TRY(workspace_t(g, "fr_bits1", nw, &b->bits));
    PG_TRY(workspace_t(g, "fr_list0", (size_t)g->n + 1, &a->list));
    PG_TRY(workspace_t(g, "fr_list1", (size_t)g->n + 1, &b->list));
    PG_TRY(workspace_t(g, "fr_ctr0", kCtrWords, &a->ctr));
    PG_TRY(workspace_t(g, "fr_ctr1", kCtrWords, &b->ctr));
    return PG_OK;
}

static int clear_frontier(pg_graph *g, Frontier &f, bool bits = true) {
    if (bits) PG_CHECK_HIP(hipMemsetAsync(f.bits, 0, 4 * ((g->n + 31) / 32 + 1), stream()));
    PG_CHECK_HIP(hipMemsetAsync(f.ctr, 0, 8 * kCtrWords, stream()));
    return PG_OK;
}

// |F| and the sum of its out-degrees (ctr[0..1] plus the slots), read back for the
// direction switch
static int read_counts(Frontier &f, int64_t *count, int64_t *edges) {
    unsigned long long h[kCtrWords];
    PG_TRY(fetch_words(h, f.ctr, sizeof(h)));
    unsigned long long c = h[0], d = h[1];
    for (int k = 0; k < kCtrSlots; k++) {
        c += h[4 + 2 * k];
        d += h[5 + 2 * k];
    }
    *count = (int64_t)c;
    *edges = (int64_t)d;
    return PG_OK;
}

// |F_next| and the sum of its out-degrees, read back for the direction switch
static int frontier_stats(pg_graph *g, Frontier &f, int64_t *count, int64_t *edges) {
    PG_TRY(list_degree_sum(g, f.list, &f.ctr[0], g->out_off, f.ctr + 5, 2));
    return read_counts(f, count, edges);
}

__global__ void k_seed_source(int32_t src, int32_t *level, int32_t *parent, int32_t *dist,
                              uint32_t *bits, int32_t *list, unsigned long long *ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (level) level[src] = 0;
        if (parent) parent[src] = src;
        if (dist) dist[src] = 0;
        bits[src >> 5] |= 1u << (src & 31);
        list[0] = src;
        ctr[0] = 1ull;
    }
}

static int lookup_internal(pg_graph *g, int64_t src, int32_t *out) {
    PG_REQUIRE(src >= 0 && src < g->n, PG_ERR_ARG, "source %lld out of range", (long long)src);
    PG_CHECK_HIP(hipMemcpy(out, g->perm + src, 4, hipMemcpyDeviceToHost));
    return PG_OK;
}

int bfs_run(pg_graph *g, int64_t src, int64_t threshold, uint32_t flags, pg_stats *st,
            int32_t **level_out, int32_t **parent_out) {
    int32_t s_int;
    PG_TRY(lookup_internal(g, src, &s_int));
    int32_t *level = nullptr, *parent = nullptr, *first_in = nullptr;
    uint32_t *V = nullptr;
    unsigned long long *slots = nullptr;
    const int64_t nw = (g->n + 31) / 32 + 1;
    PG_TRY(workspace_t(g, "bfs_level", (size_t)g->n, &level));
    PG_TRY(workspace_t(g, "bfs_parent", (size_t)g->n, &parent));
    PG_TRY(workspace_t(g, "bfs_visited", (size_t)nw, &V));
    PG_TRY(workspace_t(g, "bfs_slots", (size_t)kBfsRegion * kBfsRegions, &slots));
    unsigned long long *base = nullptr;  // rounds of the batches launched before
    PG_TRY(workspace_t(g, "bfs_base", 2, &base));
    // top-down queues: at most n vertices of out-degree above kBfsLaneDeg, at most
    // m / kBfsWaveDeg above kBfsWaveDeg
    int32_t *qbig = nullptr, *qhuge = nullptr;
    PG_TRY(workspace_t(g, "bfs_qbig", (size_t)g->n + 64, &qbig));
    PG_TRY(workspace_t(g, "bfs_qhuge", (size_t)(g->m / kBfsWaveDeg + 64), &qhuge));
    const bool have_first = g->ws.find("bfs_first_in") != g->ws.end();
    PG_TRY(workspace_t(g, "bfs_first_in", (size_t)g->n, &first_in));
    if (!have_first) PG_LAUNCH_K(k_first_in, grid_for(g->n), g->in_off, g->in_adj, g->n, first_in);
    Frontier A, B;
    PG_TRY(frontiers(g, &A, &B));
    const int64_t nbody = g->in_sched.body_end;
    const int64_t nbits = (g->n + 31) / 32;
    uint32_t *bits2 = nullptr;
    PG_TRY(workspace_t(g, "bfs_bits2", (size_t)nw, &bits2));
    const unsigned grid_round = grid_for(std::max<int64_t>(nbody / (4 * kBuChunks), nbits), 2048);
    // A batch of L rounds (k_bfs_round decides and runs a bottom-up round or a
    // top-down round's bitmap scan; k_bfs_push_queue finishes a top-down round),
    // then the base advances.  No host decision inside: the host reads the state
    // once per batch (rounds after the last frontier only record "done").
    auto enqueue_batch = [&](int L) -> int {
        for (int k = 0; k < L; k++) {
            const BfsRound R{slots, base, A.bits, B.bits, bits2, k};
            PG_LAUNCH_K(k_bfs_round<kBuChunks>, grid_round, R, g->in_off, g->in_adj, first_in,
                        g->out_off, g->out_adj, V, level, parent, qbig, qhuge, nbody, nbits,
                        threshold, g->m, g->n, kBfsAlpha, kBfsBeta);
            PG_LAUNCH_K(k_bfs_push_queue, 1024, R, qbig, qhuge, g->out_off, g->out_adj, V, level,
                        parent);
        }
        hipLaunchKernelGGL(k_bfs_advance, dim3(1), dim3(64), 0, stream(), base, (int64_t)L);
        PG_CHECK_HIP(hipGetLastError());
        return PG_OK;
    };
    const bool profile = (flags & PG_PROFILE_KERNELS) != 0;
    // the batch of L rounds, captured once per graph, threshold and L, replayed
    auto batch_exec = [&](int L, hipGraphExec_t *out) -> int {
        const std::string key = "bfs_batch_" + std::to_string(threshold) + "_" + std::to_string(L);
        auto it = g->graph_exec.find(key);
        if (it != g->graph_exec.end()) {
            *out = it->second.first;
            return PG_OK;
        }
        PG_CHECK_HIP(hipStreamSynchronize(stream()));
        PG_CHECK_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batch(L);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(stream(), &graph);
        if (rc != PG_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        PG_REQUIRE(e == hipSuccess, PG_ERR_HIP, "bfs capture failed: %s", hipGetErrorString(e));
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PG_REQUIRE(e == hipSuccess, PG_ERR_HIP, "bfs instantiate failed: %s", hipGetErrorString(e));
        g->graph_exec[key] = {exec, nullptr};
        *out = exec;
        return PG_OK;
    };
    const int first = std::min(kBfsMaxBatch, g->bfs_rounds_hint > 0 ? g->bfs_rounds_hint
                                                                      : kBfsFirstBatch);
    hipGraphExec_t exec_first = nullptr, exec_next = nullptr;
    if (!profile) {
        PG_TRY(batch_exec(first, &exec_first));
        PG_TRY(batch_exec(kBfsAhead, &exec_next));
    }
    Timer tm;
    KernelTimer kt;
    kt.on = profile;
    PG_TRY(tm.start());
    PG_LAUNCH_K(k_bfs_init, grid_for(g->n), s_int, g->n, nw, g->out_off, level, parent, V, A.bits,
                B.bits, bits2, slots, base);
    // the last round's region after a batch: its frontier's (count, out-degree)
    // pairs and the state after its decision
    unsigned long long hs[kBfsCtlEnd] = {};
    int64_t launched_rounds = 0;
    for (int L = first;; L = kBfsAhead) {
        if (profile) {
            PG_TRY(kt.begin());
            PG_TRY(enqueue_batch(L));
            PG_TRY(kt.end());
        } else {
            PG_CHECK_HIP(hipGraphLaunch(launched_rounds == 0 ? exec_first : exec_next,
                                        stream()));
        }
        launched_rounds += L;
        const unsigned long long *last =
            slots + (size_t)(launched_rounds % kBfsRegions) * kBfsRegion;
        PG_TRY(fetch_words(hs, last, sizeof(hs)));
        unsigned long long c = 0;
        for (int i = 0; i < 64; i++) c += hs[2 * i];
        if (hs[kBfsCtlDone] || c == 0) break;
        PG_REQUIRE(launched_rounds <= 4 * g->n + 64, PG_ERR_STATE, "bfs: no convergence");
    }
    // the next search's first batch: this one's rounds, lowered by at most one
    g->bfs_rounds_hint = std::max<int>((int)hs[kBfsCtlRounds], g->bfs_rounds_hint - 1);
    if (g->bfs_rounds_hint < 1) g->bfs_rounds_hint = 1;
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (st) {
        st->rounds = (int32_t)hs[kBfsCtlRounds];
        st->dense_rounds = (int32_t)hs[kBfsCtlDense];
        st->edges = (int64_t)hs[kBfsCtlEdges];
        st->active = (int64_t)hs[kBfsCtlActive];
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    *level_out = level;
    *parent_out = parent;
    return PG_OK;
}

int cc_run(pg_graph *g, int64_t threshold, uint32_t flags, pg_stats *st, int32_t **label_out) {
    const int64_t thr = resolve_threshold(threshold, g->m, 8);
    int32_t *label = nullptr, *prev = nullptr;
    PG_TRY(workspace_t(g, "cc_label", (size_t)g->n, &label));
    PG_TRY(workspace_t(g, "cc_prev", (size_t)g->n, &prev));
    Frontier A, B;
    PG_TRY(frontiers(g, &A, &B));
    // dense rounds on the source-sorted tiles when the graph has them
    const bool tiled = g->use_tiles;
    if (tiled) PG_TRY(ensure_tiles(g));
    Timer tm;
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    PG_TRY(tm.start());
    PG_TRY(clear_frontier(g, A));
    PG_TRY(fill_bits_all(g, A.bits));
    PG_LAUNCH_K(k_iota_labels, grid_for(g->n), g->iperm, g->n, label, A.list);
    unsigned long long hc[2] = {(unsigned long long)g->n, (unsigned long long)g->m};
    PG_CHECK_HIP(hipMemcpyAsync(A.ctr, hc, 16, hipMemcpyHostToDevice, stream()));
    int64_t cnt = g->n, edges = g->m;
    int rounds = 0, dense_rounds = 0;
    int64_t tot_active = 0, tot_edges = 0;
    Frontier *F = &A, *N = &B;
    while (cnt > 0 || rounds == 0) {
        rounds++;
        tot_active += cnt;
        tot_edges += edges;
        // vertexMap(CC_Vertex_F): prev = ID on the frontier
        PG_LAUNCH_K(k_cc_prev, grid_for(cnt), F->list, &F->ctr[0], label, prev);
        PG_TRY(clear_frontier(g, *N));
        const bool dense = cnt + edges >= thr;
        PG_TRY(kt.begin());
        // a round that may activate many vertices builds the next list by compaction
        const bool compact = dense || edges >= g->n / 64;
        if (dense) {
            dense_rounds++;
            CcPullOp op{F->bits, N->bits, label};
            if (tiled)
                PG_TRY(launch_pull_tiled(op, g->in_sched, g->tiles));
            else
                PG_TRY(launch_pull(op, g->in_sched));
        } else {
            CcPushOp op{N->bits, N->list, &N->ctr[0], label, prev, !compact};
            PG_TRY(launch_push_global(op, g, g->out_off, g->out_adj, F->list, cnt, edges));
        }
        if (compact) PG_TRY(bits_to_list(g, N->bits, N->list, &N->ctr[0]));
        PG_TRY(kt.end());
        PG_TRY(frontier_stats(g, *N, &cnt, &edges));
        std::swap(F, N);
        if (rounds > 4 * g->n + 8) break;
    }
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (st) {
        st->rounds = rounds;
        st->dense_rounds = dense_rounds;
        st->edges = tot_edges;
        st->active = tot_active;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    *label_out = label;
    return PG_OK;
}

int sssp_run(pg_graph *g, int64_t src, int64_t threshold, uint32_t flags, pg_stats *st,
             int32_t **dist_out) {
    PG_REQUIRE(g->w_dtype == PG_W_I32, PG_ERR_STATE, "SSSP needs int32 weights");
    int32_t s_int;
    PG_TRY(lookup_internal(g, src, &s_int));
    const int64_t thr = resolve_threshold(threshold, g->m, 20);
    int32_t *dist = nullptr;
    PG_TRY(workspace_t(g, "sssp_dist", (size_t)g->n, &dist));
    Frontier A, B;
    PG_TRY(frontiers(g, &A, &B));
    // dense rounds as a pull on the source-sorted tiles when the graph has them
    const bool tiled = g->use_tiles;
    if (tiled) PG_TRY(ensure_tiles(g));
    Timer tm;
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    PG_TRY(tm.start());
    PG_TRY(fill(dist, (int32_t)1073741823, g->n));  // INT_MAX/2, numa-BellmanFord.C:283
    PG_TRY(clear_frontier(g, A));
    PG_LAUNCH_K(k_seed_source, 1, s_int, (int32_t *)nullptr, (int32_t *)nullptr, dist, A.bits,
                A.list, A.ctr);
    int64_t cnt = 1, edges = 0;
    PG_TRY(frontier_stats(g, A, &cnt, &edges));
    int rounds = 0, dense_rounds = 0;
    int64_t tot_active = 0, tot_edges = 0;
    Frontier *F = &A, *N = &B;
    while (cnt > 0) {
        rounds++;
        tot_active += cnt;
        tot_edges += edges;
        PG_TRY(clear_frontier(g, *N));
        const bool dense = cnt + edges >= thr;
        PG_TRY(kt.begin());
        // a round that may activate many vertices builds the next list by compaction
        const bool compact = dense || edges >= g->n / 64;
        SsspPushOp<false> op{F->bits, N->bits, N->list, &N->ctr[0], dist, g->out_w_i32, !compact};
        // a dense round whose frontier holds at least half of the edges pulls on the
        // tiles (which read every in-edge); other dense rounds push (S22, nine sources:
        // 3.04 ms a source against 3.43 with every dense round pulled, 3.52 pushed)
        if (dense && tiled && edges >= g->m / 2) {
            dense_rounds++;
            PG_TRY(launch_pull_tiled(SsspPullOp{F->bits, N->bits, dist, g->in_w_i32}, g->in_sched,
                                     g->tiles));
        } else if (dense) {
            // edgeMapDenseForward visits the active vertices in id order: the
            // bitmap compacted in order (packIndex), then the same balanced push
            dense_rounds++;
            int32_t *ordered = nullptr;
            unsigned long long *oc = nullptr;
            PG_TRY(workspace_t(g, "sssp_ordered", (size_t)g->n + 1, &ordered));
            PG_TRY(workspace_t(g, "sssp_ordered_n", 2, &oc));
            PG_TRY(bits_to_list(g, F->bits, ordered, oc));
            PG_TRY(launch_push_global(op, g, g->out_off, g->out_adj, ordered, cnt, edges));
        } else {
            PG_TRY(launch_push_global(op, g, g->out_off, g->out_adj, F->list, cnt, edges));
        }
        if (compact) PG_TRY(bits_to_list(g, N->bits, N->list, &N->ctr[0]));
        PG_TRY(kt.end());
        PG_TRY(frontier_stats(g, *N, &cnt, &edges));
        std::swap(F, N);
        if (rounds > 4 * g->n + 8) break;
    }
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (st) {
        st->rounds = rounds;
        st->dense_rounds = dense_rounds;
        st->edges = tot_edges;
        st->active = tot_active;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    *dist_out = dist;
    return PG_OK;
}

}  // namespace pg

using namespace pg;

// ====================================================================================
// C ABI
// ====================================================================================
static int check_graph(pg_graph *g) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(g, PG_ERR_ARG, "null graph");
    return PG_OK;
}

extern "C" int pg_pagerank(pg_graph *g, int iters, double damping, float *rank, uint32_t flags,
                           pg_stats *st) {
    PG_TRY(check_graph(g));
    PG_REQUIRE(iters >= 0, PG_ERR_ARG, "iters < 0");
    PG_REQUIRE(damping >= 0.0 && damping <= 1.0, PG_ERR_ARG, "damping out of [0,1]");
    float *res = nullptr;
    PG_TRY(pagerank_run(g, iters, damping, flags, st, &res));
    return to_input_order_f32(g, res, rank, 1);
}

// rounds whose frontier has fewer out-edges than m / kPrdPushDiv push
constexpr int64_t kPrdPushDiv = 8;

extern "C" int pg_pagerank_delta(pg_graph *g, int iters, double damping, double epsilon2,
                                 double *p_out, uint32_t flags, pg_stats *st) {
    PG_TRY(check_graph(g));
    PG_REQUIRE(iters >= 0, PG_ERR_ARG, "iters < 0");
    PG_REQUIRE(damping >= 0.0 && damping <= 1.0, PG_ERR_ARG, "damping out of [0,1]");
    PG_REQUIRE(epsilon2 >= 0.0, PG_ERR_ARG, "epsilon2 < 0");
    const int64_t n = g->n;
    double *p = nullptr, *c = nullptr, *cn = nullptr, *ngh = nullptr;
    unsigned long long *stats = nullptr;
    PG_TRY(workspace_t(g, "prd_p", (size_t)n, &p));
    PG_TRY(workspace_t(g, "prd_c0", (size_t)n, &c));
    PG_TRY(workspace_t(g, "prd_c1", (size_t)n, &cn));
    PG_TRY(workspace_t(g, "prd_ngh", (size_t)n, &ngh));
    PG_TRY(workspace_t(g, "prd_stats", ((size_t)iters + 1) * kStatSlots, &stats));
    Frontier A, B;
    PG_TRY(frontiers(g, &A, &B));
    // pull rounds on the tiles when the graph's schedule has them: the tile set with
    // a cold value array of doubles
    const bool tiled = g->use_tiles;
    TileSet tset{};
    if (tiled) {
        PG_TRY(ensure_tiles(g));
        tset = g->tiles;
        if (tset.m_cold > 0) {
            double *cold64 = nullptr;
            PG_TRY(workspace_t(g, "prd_cold64", (size_t)tset.m_cold + 16, &cold64));
            tset.args.cold_val64 = cold64;
            tset.cold.val64 = cold64;
        }
    }
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    const double one_over_n = 1.0 / (double)n;
    PG_LAUNCH_K(k_prd_init, grid_for(n), n, g->out_off, one_over_n, p, c, ngh);
    PG_LAUNCH_K(k_zero_u64, (unsigned)(((iters + 1) * kStatSlots + kBlock - 1) / kBlock), stats,
                (iters + 1) * kStatSlots);
    PrDeltaOp op{};
    op.out_off = g->out_off;
    op.damp = damping;
    op.addc = (1.0 - damping) * one_over_n;
    op.one_over_n = one_over_n;
    op.eps2 = epsilon2;
    // round 1 runs on every vertex; round t+1 on the frontier round t left
    int64_t cnt = n, edges = g->m, tot_active = 0, tot_edges = 0;
    int dense_rounds = 0;
    bool bits_ready = false;  // A.bits holds the nonzero entries of c
    for (int t = 0; t < iters; t++) {
        tot_active += cnt;
        tot_edges += edges;
        unsigned long long *st_t = &stats[(size_t)t * kStatSlots];
        PG_TRY(kt.begin());
        const bool pull = t == 0 || edges >= g->m / kPrdPushDiv;
        if (pull) {
            dense_rounds++;
            op.c = c;
            op.cn = cn;
            op.p = p;
            op.first = t == 0 ? 1 : 0;
            if (tiled)
                PG_TRY(launch_pull_tiled(op, g->in_sched, tset, st_t));
            else
                PG_TRY(launch_pull(op, g->in_sched, st_t));
        } else {
            // the sources with a nonzero contribution, as an ordered list (their
            // bitmap written by the previous push round's filter, or formed here)
            PG_TRY(clear_frontier(g, A, false));
            if (!bits_ready) PG_LAUNCH_K(k_prd_bits, grid_for(n, 4096), c, n, A.bits);
            PG_TRY(bits_to_list(g, A.bits, A.list, &A.ctr[0]));
            int64_t lc = 0, le = 0;
            PG_TRY(frontier_stats(g, A, &lc, &le));
            PG_TRY(launch_push_global(PrDeltaPushOp{c, ngh}, g, g->out_off, g->out_adj, A.list, lc,
                                      le));
            PG_LAUNCH_K(k_prd_filter, grid_for(n, 4096), n, ngh, p, cn, g->out_off, damping,
                        epsilon2, st_t, A.bits);
        }
        bits_ready = !pull;
        PG_TRY(kt.end());
        std::swap(c, cn);
        if (t + 1 < iters) {  // the next frontier's size and out-degree sum
            unsigned long long hs[kStatSlots];
            PG_TRY(fetch_words(hs, st_t, sizeof(hs)));
            unsigned long long sum = 0;
            for (int k = 0; k < kStatSlots; k++) sum += hs[k];
            cnt = (int64_t)(sum >> 32);
            edges = (int64_t)(sum & 0xFFFFFFFFull);
        }
    }
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (st) {
        st->rounds = iters;
        st->dense_rounds = dense_rounds;
        st->active = tot_active;
        st->edges = tot_edges;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    // fp64 values travel as pairs of 32-bit words
    return to_input_order_f32(g, reinterpret_cast<const float *>(p), reinterpret_cast<float *>(p_out),
                              2);
}

__global__ void k_spmv_copy(const float *a, int64_t n, float *b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

extern "C" int pg_spmv(pg_graph *g, int iters, const float *x0, double *y, uint32_t flags,
                       pg_stats *st) {
    PG_TRY(check_graph(g));
    PG_REQUIRE(iters >= 0 && x0, PG_ERR_ARG, "bad argument");
    const float *w = nullptr;
    PG_TRY(spmv_weights(g, &w));
    const int64_t n = g->n;
    float *xa = nullptr, *xb = nullptr;
    unsigned *maxbits = nullptr;
    int32_t *shift = nullptr;
    long long *total = nullptr;
    const bool save = (flags & PG_SPMV_SAVE) != 0;
    g->spmv_saved_iters = 0;  // the kept iterates (if any) are overwritten below
    float *saved = nullptr;   // save: y_t (scaled) at saved + t n, t = 0..iters
    if (save) {
        PG_TRY(workspace_t(g, "spmv_saved", (size_t)n * (size_t)(iters + 1), &saved));
        xa = saved;
    } else {
        PG_TRY(workspace_t(g, "spmv_xa", (size_t)n, &xa));
        PG_TRY(workspace_t(g, "spmv_xb", (size_t)n, &xb));
    }
    PG_TRY(workspace_t(g, "spmv_max", (size_t)iters + 2, &maxbits));
    PG_TRY(workspace_t(g, "spmv_shift", (size_t)iters + 2, &shift));
    PG_TRY(workspace_t(g, "spmv_total", 2, &total));
    PG_TRY(from_input_order_f32(g, x0, xa, 1));
    PG_TRY(fill(maxbits, 0u, iters + 2));
    PG_TRY(fill(shift, 0, iters + 2));
    PG_TRY(fill(total, 0ll, 2));
    // exponent of x0 (slot 0), applied to x0 itself
    PG_LAUNCH_K(k_absmax_bits, grid_for(n, 2048), xa, n, maxbits);
    hipLaunchKernelGGL(k_spmv_shift, dim3(1), dim3(64), 0, stream(), maxbits, shift, total, 0);
    PG_CHECK_HIP(hipGetLastError());
    PG_LAUNCH_K(k_ldexp_inplace, grid_for(n), xa, n, shift);
    // tiled pull (the fixed-point tiles of the PageRank family, with the weights in
    // tile order): the largest in-weight sum once per graph, max |x0| as scaled, and
    // a fixed-point exponent per iteration
    const bool tiled = g->use_tiles;
    int32_t *tscale = nullptr;
    unsigned *wmax = nullptr, *x0max = nullptr;
    if (tiled) {
        PG_TRY(ensure_tiles(g));
        PG_REQUIRE(g->tiles.m_tiled == 0 || g->tiles.d_wt, PG_ERR_STATE,
                   "tiles carry no weights");
        const bool have_w = g->ws.find("spmv_wmax") != g->ws.end();
        PG_TRY(workspace_t(g, "spmv_wmax", 2, &wmax));
        if (!have_w) {
            PG_TRY(fill(wmax, 0u, 2));
            PG_LAUNCH_K(k_in_weight_sum_max, grid_for(n, 4096), g->in_off, w, n, wmax);
        }
        PG_TRY(workspace_t(g, "spmv_tscale", (size_t)iters + 2, &tscale));
        PG_TRY(workspace_t(g, "spmv_x0max", 2, &x0max));
        PG_TRY(fill(x0max, 0u, 2));
        PG_LAUNCH_K(k_absmax_bits, grid_for(n, 2048), xa, n, x0max);
    }
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    float *cur = xa, *nxt = xb;
    // the next iteration's cold pre-pass (raw values, scaled later in the tiles) runs
    // beside this one's finishing pass
    const bool ahead = tiled && g->tiles.cold_overlap && g->tiles.cold.nparts > 0 &&
                       (flags & PG_SERIAL_COLD) == 0;
    for (int t = 0; t < iters; t++) {
        if (save) {  // every iterate in its own slot
            cur = saved + (size_t)t * n;
            nxt = saved + (size_t)(t + 1) * n;
        }
        // iteration t scales its output by 2^-shift[t+1]; shift[t+1] comes from max(x_t)
        SpmvOp op{cur, w, nxt, shift + t + 1, tscale ? tscale + t : nullptr};
        const SpmvOp next{nxt, w, save ? nxt + n : cur, shift + t + 2,
                          tscale ? tscale + t + 1 : nullptr};
        PG_TRY(kt.begin());
        if (tiled) {
            hipLaunchKernelGGL(k_spmv_tscale, dim3(1), dim3(64), 0, stream(),
                               t == 0 ? x0max : maxbits + t, wmax, tscale + t);
            // the zero-in-degree tail holds zeros in both buffers from iteration 2
            // (a fresh slot of the kept iterates does not: it is written every time)
            // the finishing pass runs on the side stream beside the next cold pre-pass;
            // the exponent shift below reads its maxima, so it waits for it
            const bool aside = ahead && t + 1 < iters;
            PG_TRY(launch_pull_tiled(op, g->in_sched, g->tiles, maxbits + t + 1, t >= 2 && !save,
                                     aside ? &next : nullptr, ahead && t > 0, true));
            if (aside) PG_CHECK_HIP(hipStreamWaitEvent(stream(), g->tiles.ev_cold, 0));
        } else {
            PG_TRY(launch_pull(op, g->in_sched, maxbits + t + 1));
        }
        PG_TRY(kt.end());
        hipLaunchKernelGGL(k_spmv_shift, dim3(1), dim3(64), 0, stream(), maxbits + t + 1,
                           shift + t + 2, total, 0);
        PG_CHECK_HIP(hipGetLastError());
        std::swap(cur, nxt);  // SPMV_Vertex_Reset + swap
    }
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (save) g->spmv_saved_iters = iters > 0 ? iters : 0;
    if (st) {
        st->rounds = iters;
        st->dense_rounds = iters;
        st->edges = (int64_t)iters * g->m;
        st->active = (int64_t)iters * g->n;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    if (!y) return PG_OK;
    // total exponent = sum of the applied shifts shift[0..iters]
    std::vector<int32_t> hs((size_t)iters + 2);
    PG_CHECK_HIP(hipMemcpy(hs.data(), shift, 4 * (iters + 2), hipMemcpyDeviceToHost));
    long long E = 0;
    for (int t = 0; t <= iters; t++) E += hs[t];
    PG_CHECK_HIP(hipMemcpy(total, &E, 8, hipMemcpyHostToDevice));
    double *tmp = nullptr;
    PG_TRY(workspace_t(g, "spmv_out", (size_t)n, &tmp));
    PG_LAUNCH_K(k_to_f64_scaled, grid_for(n), cur, g->perm, n, total, tmp);
    PG_CHECK_HIP(hipMemcpyAsync(y, tmp, 8 * n, hipMemcpyDeviceToHost, stream()));
    PG_CHECK_HIP(hipStreamSynchronize(stream()));
    return PG_OK;
}

// ---- SpMV adjoint -----------------------------------------------------------------------
// Backward step t of L = g . y_T is the pull over the CSR (the transposed graph):
//     mu_t[s] = sum_{e=(s->d)} w[e] mu_{t+1}[d]
// run on the tiles of the CSR (every vertex a destination of its out-list, the
// out-neighbours sorted within a tile: the in-degree order puts the hot ones first),
// with SpmvOp's fixed point and power-of-two exponents: mu_t is stored scaled by
// 2^-F_t.  Every mu_t is kept; the weight gradient
//     grad[e] = sum_t y_t[s] mu_{t+1}[d] 2^(E_t + F_{t+1})   (fac[t] = 2^(E_t + F_{t+1}))
// is then one pass over the edges, in fp64, with y_t and mu_{t+1} of up to kGradRow
// steps transposed into one 128-byte row per vertex: a row gather per edge (the
// source's row is shared by its out-edges) instead of a read-modify-write of the
// gradient per edge and step.
constexpr int kGradRow = 32;

// Exponent bookkeeping of backward step t (one thread): shiftA[t] from the stored
// max of mu_{t+1}; F[t] = F[t+1] + shiftA[t]; fac[t] = 2^(E_t + F[t+1]) where
// E_t = shiftF[0] + ... + shiftF[t] (the forward's applied shifts).
__global__ void k_spmv_adj_step(const unsigned *maxbits_next, const int32_t *shiftF, int t,
                                int32_t *shiftA, long long *F, double *fac) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const float mx = __uint_as_float(*maxbits_next);
        int s = 0;
        if (mx > 0.f && isfinite(mx)) s = ilogbf(mx) - 64;
        shiftA[t] = s;
        long long E = 0;
        for (int k = 0; k <= t; k++) E += shiftF[k];
        F[t] = F[t + 1] + s;
        fac[t] = ldexp(1.0, (int)(E + F[t + 1]));
    }
}

// mu_T = g scaled: shiftA[T] from max |g|, F[T] = shiftA[T]
__global__ void k_spmv_adj_init(const unsigned *maxbits, int T, int32_t *shiftA, long long *F) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const float mx = __uint_as_float(*maxbits);
        int s = 0;
        if (mx > 0.f && isfinite(mx)) s = ilogbf(mx) - 64;
        shiftA[T] = s;
        F[T] = s;
    }
}

__global__ void k_scatter_f64(const double *in, const int32_t *idx, int64_t m, double *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[idx[i]] = in[i];
}

// src[e] = the source of CSR edge e (a wave per vertex)
__global__ void k_out_src(const int32_t *off, int64_t n, int32_t *src) {
    const int lane = lane_id();
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < n; v += nw)
        for (int e = off[v] + lane; e < off[v + 1]; e += 64) src[e] = (int32_t)v;
}

// rows[v * kGradRow + j] = a[(t0 + j) n + v] for j < cnt, 0 for cnt <= j < kGradRow: a
// wave reads its 64 vertices' values step by step (coalesced), stages the 64 rows in
// LDS and writes them as 8 KB of consecutive 16-byte pieces (a lane writing its own
// row's pieces touched 64 lines per store and wrote every line three times over)
__global__ __launch_bounds__(kBlock) void k_rows_from_steps(const float *a, int64_t n, int t0,
                                                            int cnt, float *rows) {
    __shared__ float st[kBlock / 64][64 * (kGradRow + 1)];  // +1: rows start on distinct banks
    const int lane = lane_id(), w = threadIdx.x >> 6;
    float *sw = st[w];
    // block-uniform steps of kBlock vertices (a wave takes 64 of them), so every wave
    // of the block reaches each barrier
    for (int64_t bb = (int64_t)blockIdx.x * kBlock; bb < n; bb += (int64_t)gridDim.x * kBlock) {
        const int64_t base = bb + 64 * w;
        const int64_t v = base + lane;
        for (int j = 0; j < kGradRow; j++)
            sw[lane * (kGradRow + 1) + j] = (j < cnt && v < n) ? a[(int64_t)(t0 + j) * n + v] : 0.f;
        __syncthreads();
        const int64_t nrows = max((int64_t)0, min((int64_t)64, n - base));
        float4 *out = reinterpret_cast<float4 *>(rows + base * kGradRow);
        for (int k = lane; k < nrows * (kGradRow / 4); k += 64) {
            const int r = k / (kGradRow / 4), q = k % (kGradRow / 4);
            const float *x = &sw[r * (kGradRow + 1) + 4 * q];
            out[k] = make_float4(x[0], x[1], x[2], x[3]);
        }
        __syncthreads();
    }
}

// grad[e] (+)= sum_{j < cnt} fac[j] Y[s][j] M[d][j] in fp64 for the CSR edge e = (s -> d)
// (first: store); a thread per edge, the rows' needed float4s loaded together (the
// source's row is shared by its out-edges, which sit side by side; reading the
// source's values from the kept iterates step by step instead took a line per step
// and source: 2.3x slower)
__global__ __launch_bounds__(kBlock) void k_spmv_grad_rows(const int32_t *__restrict__ src,
                                                           const int32_t *__restrict__ adj, int64_t m,
                                                           const float *__restrict__ Y,
                                                           const float *__restrict__ M,
                                                           const double *__restrict__ fac, int cnt,
                                                           int first, double *grad) {
    __shared__ double sf[kGradRow];
    if (threadIdx.x < kGradRow) sf[threadIdx.x] = (int)threadIdx.x < cnt ? fac[threadIdx.x] : 0.0;
    __syncthreads();
    const int nq = (cnt + 3) / 4;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        const float4 *ys = reinterpret_cast<const float4 *>(Y + (int64_t)__builtin_nontemporal_load(src + e) * kGradRow);
        const float4 *md = reinterpret_cast<const float4 *>(M + (int64_t)__builtin_nontemporal_load(adj + e) * kGradRow);
        float4 a[kGradRow / 4], b[kGradRow / 4];
#pragma unroll
        for (int q = 0; q < kGradRow / 4; q++) {
            a[q] = q < nq ? ys[q] : make_float4(0.f, 0.f, 0.f, 0.f);
            b[q] = q < nq ? md[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kGradRow / 4; q++) {
            acc += sf[4 * q] * ((double)a[q].x * (double)b[q].x);
            acc += sf[4 * q + 1] * ((double)a[q].y * (double)b[q].y);
            acc += sf[4 * q + 2] * ((double)a[q].z * (double)b[q].z);
            acc += sf[4 * q + 3] * ((double)a[q].w * (double)b[q].w);
        }
        grad[e] = first ? acc : grad[e] + acc;
    }
}

extern "C" int pg_spmv_adjoint(pg_graph *g, const float *gvec, double *grad_x0, double *grad_w,
                               uint32_t flags, pg_stats *st) {
    PG_TRY(check_graph(g));
    PG_REQUIRE(gvec, PG_ERR_ARG, "null gvec");
    PG_REQUIRE(g->spmv_saved_iters > 0, PG_ERR_STATE,
               "pg_spmv has not been run with PG_SPMV_SAVE on this graph");
    const int T = g->spmv_saved_iters;
    const int64_t n = g->n, m = g->m;
    // fp32 weights in internal CSR order (the largest out-weight sum)
    const float *w = nullptr;
    if (g->w_dtype == PG_W_F32) {
        w = g->out_w_f32;
    } else {
        PG_REQUIRE(g->w_dtype == PG_W_I32, PG_ERR_STATE, "SpMV needs a weighted graph");
        float *wf = nullptr;
        const bool have = g->ws.find("spmv_wf_out") != g->ws.end();
        PG_TRY(workspace_t(g, "spmv_wf_out", (size_t)m, &wf));
        if (!have) PG_LAUNCH_K(k_i32_to_f32, grid_for(m), g->out_w_i32, m, wf);
        w = wf;
    }
    // once per graph: the tiles of the CSR, the edges' sources, the largest out-weight sum
    PG_TRY(ensure_tiles_out(g));
    const TileSet &to = g->tiles_out;
    PG_REQUIRE(to.m_tiled == 0 || to.d_wt, PG_ERR_STATE, "CSR tiles carry no weights");
    int32_t *osrc = nullptr;
    unsigned *wmax = nullptr;
    const bool have_src = g->ws.find("spmv_adj_src") != g->ws.end();
    PG_TRY(workspace_t(g, "spmv_adj_src", (size_t)m + 1, &osrc));
    PG_TRY(workspace_t(g, "spmv_adj_wmax", 2, &wmax));
    if (!have_src) {
        PG_LAUNCH_K(k_out_src, grid_for(n * 64), g->out_off, n, osrc);
        PG_TRY(fill(wmax, 0u, 2));
        PG_LAUNCH_K(k_in_weight_sum_max, grid_for(n, 4096), g->out_off, w, n, wmax);
    }
    float *saved = nullptr, *mu = nullptr, *yrows = nullptr, *mrows = nullptr;
    int32_t *shiftF = nullptr, *shiftA = nullptr, *tsA = nullptr;
    unsigned *maxA = nullptr;
    long long *F = nullptr;
    double *fac = nullptr, *grad = nullptr;
    const int tb = std::min(T, kGradRow);  // steps per gradient pass
    PG_TRY(workspace_t(g, "spmv_saved", (size_t)n * (size_t)(T + 1), &saved));
    PG_TRY(workspace_t(g, "spmv_shift", (size_t)T + 2, &shiftF));
    PG_TRY(workspace_t(g, "spmv_adj_mu", (size_t)n * (size_t)(T + 1), &mu));
    PG_TRY(workspace_t(g, "spmv_adj_yrows", (size_t)n * kGradRow, &yrows));
    PG_TRY(workspace_t(g, "spmv_adj_mrows", (size_t)n * kGradRow, &mrows));
    PG_TRY(workspace_t(g, "spmv_adj_shift", (size_t)T + 2, &shiftA));
    PG_TRY(workspace_t(g, "spmv_adj_tscale", (size_t)T + 2, &tsA));
    PG_TRY(workspace_t(g, "spmv_adj_max", (size_t)T + 2, &maxA));
    PG_TRY(workspace_t(g, "spmv_adj_F", (size_t)T + 2, &F));
    PG_TRY(workspace_t(g, "spmv_adj_fac", (size_t)T + 1, &fac));
    PG_TRY(workspace_t(g, "spmv_adj_grad", (size_t)m, &grad));
    float *muT = mu + (size_t)T * n;
    PG_TRY(from_input_order_f32(g, gvec, muT, 1));
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    PG_TRY(fill(maxA, 0u, T + 2));
    // max |g| (slot T+1) sets mu_T's scale; slot T then holds the max of mu_T as
    // stored, from which the first backward step takes its shift
    PG_LAUNCH_K(k_absmax_bits, grid_for(n, 2048), muT, n, maxA + T + 1);
    hipLaunchKernelGGL(k_spmv_adj_init, dim3(1), dim3(64), 0, stream(), maxA + T + 1, T, shiftA,
                       F);
    PG_CHECK_HIP(hipGetLastError());
    PG_LAUNCH_K(k_ldexp_inplace, grid_for(n), muT, n, shiftA + T);
    PG_LAUNCH_K(k_absmax_bits, grid_for(n, 2048), muT, n, maxA + T);
    for (int t = T - 1; t >= 0; t--) {
        hipLaunchKernelGGL(k_spmv_adj_step, dim3(1), dim3(64), 0, stream(), maxA + t + 1, shiftF,
                           t, shiftA, F, fac);
        PG_CHECK_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_spmv_tscale, dim3(1), dim3(64), 0, stream(), maxA + t + 1, wmax,
                           tsA + t);
        PG_CHECK_HIP(hipGetLastError());
        // mu_t (stored scaled by 2^-shiftA[t]) from mu_{t+1}
        const SpmvOp op{mu + (size_t)(t + 1) * n, w, mu + (size_t)t * n, shiftA + t, tsA + t};
        PG_TRY(kt.begin());
        PG_TRY(launch_tiles(op, to, maxA + t));
        PG_TRY(kt.end());
    }
    // the weight gradient, kGradRow steps per pass
    for (int t0 = 0; t0 < T; t0 += tb) {
        const int cnt = std::min(tb, T - t0);
        PG_TRY(kt.begin());
        PG_LAUNCH_K(k_rows_from_steps, grid_for(n, 4096), saved, n, t0, cnt, yrows);   // y_t
        PG_LAUNCH_K(k_rows_from_steps, grid_for(n, 4096), mu + n, n, t0, cnt, mrows);  // mu_{t+1}
        PG_LAUNCH_K(k_spmv_grad_rows, grid_for(m), osrc, g->out_adj, m, yrows, mrows, fac + t0,
                    cnt, t0 == 0 ? 1 : 0, grad);
        PG_TRY(kt.end());
    }
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (st) {
        st->rounds = T;
        st->dense_rounds = T;
        st->edges = (int64_t)T * m;
        st->active = (int64_t)T * n;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    if (grad_x0) {  // mu_0 = stored * 2^F[0]
        double *tmp = nullptr;
        PG_TRY(workspace_t(g, "spmv_out", (size_t)n, &tmp));
        PG_LAUNCH_K(k_to_f64_scaled, grid_for(n), mu, g->perm, n, F, tmp);
        PG_CHECK_HIP(hipMemcpyAsync(grad_x0, tmp, 8 * n, hipMemcpyDeviceToHost, stream()));
    }
    if (grad_w) {
        double *tmp = nullptr;
        PG_TRY(workspace_t(g, "spmv_adj_gw_in", (size_t)m, &tmp));
        PG_LAUNCH_K(k_scatter_f64, grid_for(m), grad, g->out_eid, m, tmp);
        PG_CHECK_HIP(hipMemcpyAsync(grad_w, tmp, 8 * m, hipMemcpyDeviceToHost, stream()));
    }
    PG_CHECK_HIP(hipStreamSynchronize(stream()));
    return PG_OK;
}

__global__ void k_f32_to_float4(const float *in, int64_t m, float4 *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = make_float4(in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]);
}

__global__ void k_gather_pot(const float4 *in, const int32_t *eid, int64_t m, float4 *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[eid[i]];
}

template <class T>
__global__ void k_bp_state_init(const float *p0, int64_t n, Vec2<T> *cur, Vec2<T> *nxt) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        cur[v] = Vec2<T>{(T)p0[2 * v], (T)p0[2 * v + 1]};
        nxt[v] = Vec2<T>{T(0), T(0)};  // mmap'd zero pages (numa-BP.C:557)
    }
}

template <class T>
__global__ void k_bp_out(const Vec2<T> *a, const int32_t *perm, int64_t n, double *out) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x) {
        const Vec2<T> x = a[perm[u]];
        out[2 * u] = (double)x.x;
        out[2 * u + 1] = (double)x.y;
    }
}

// ---- edge-balanced pull of the light destinations -------------------------------
// The destinations of the lane-group classes up to kBpLightMaxG lanes (in-degree
// at most 16 * items per lane: positions [lb, be) of the CSC, all with at least one
// in-edge) are cut into chunks of about
// kBpChunkEdges consecutive in-edges, aligned to vertices: chunk c holds the
// vertices whose first in-edge lies in [off[lb] + c*K, off[lb] + (c+1)*K).  One
// wave walks a chunk 64 edges at a time, one edge per lane (the potentials are read
// as whole 1 KiB rows), and forms each destination's product with a segmented scan
// across the lanes: a bitmap over the CSC edges marks each light vertex's first
// in-edge, so a lane finds its destination with a popcount and its segment's start
// with a leading-zero count.  A destination whose edges run past a step carries its
// partial product into the next step.  The heavier destinations run on the
// degree-binned blocks of the same launch.
constexpr int kBpChunkEdges = 2048;
constexpr int kBpSteps = 2;     // 64-edge steps whose loads are issued together
constexpr int kBpLightMaxG = 16;  // the light range: the lane-group classes up to this one

struct BpLight {
    const int32_t *chunk_v;               // nchunks + 1 vertex positions
    const unsigned long long *heads;      // bit e set: e is a light vertex's first in-edge
    int nchunks;
};

// v moved across the lanes of a wave by a DPP control; the lanes it does not
// feed (outside the row mask, or without a source lane) get 1
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, ROWS, 0xF, false);
    const int hi =
        __builtin_amdgcn_update_dpp(0x3FF00000, (int)(unsigned)(u >> 32), CTRL, ROWS, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0x3F800000, __builtin_bit_cast(int, v), CTRL, ROWS, 0xF, false));
}

template <class T>
__device__ __forceinline__ T readlane_t(T v, int l) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    }
}

// Inclusive segmented product of (b0, b1) over the 64 lanes, a segment starting at
// lane ss (the last head at or below this lane, -1: none): Hillis-Steele steps
// within rows of 16 lanes (row_shr 1, 2, 4, 8), then the last lane of row 0 into
// row 1 and of row 2 into row 3 (row_bcast 15), then lane 31 into rows 2 and 3
// (row_bcast 31); a lane takes a value only from its own segment.
template <class T>
__device__ __forceinline__ void seg_scan_mul(T &b0, T &b1, int lane, int ss) {
#define PG_SEG_STAGE(CTRL, ROWS, COND)                                      \
    {                                                                       \
        const T u0 = dpp_move<CTRL, ROWS>(b0), u1 = dpp_move<CTRL, ROWS>(b1); \
        if (COND) {                                                         \
            b0 *= u0;                                                       \
            b1 *= u1;                                                       \
        }                                                                   \
    }
    PG_SEG_STAGE(0x111, 0xF, (lane & 15) >= 1 && lane - 1 >= ss)
    PG_SEG_STAGE(0x112, 0xF, (lane & 15) >= 2 && lane - 2 >= ss)
    PG_SEG_STAGE(0x114, 0xF, (lane & 15) >= 4 && lane - 4 >= ss)
    PG_SEG_STAGE(0x118, 0xF, (lane & 15) >= 8 && lane - 8 >= ss)
    PG_SEG_STAGE(0x142, 0xA, ((lane >> 4) & 1) && ss < (lane & ~15))
    PG_SEG_STAGE(0x143, 0xC, lane >= 32 && ss < 32)
#undef PG_SEG_STAGE
}

// the vertex protocol of k_pull for one destination, with its frontier flag and
// out-degree already loaded: skipped when it is not in the frontier, else
// finished with its product
template <class T, bool D>
__device__ __forceinline__ unsigned long long bp_close(const BpOp<T, D> &op, int v, T p0, T p1,
                                                       bool any, int in_frontier, int outdeg) {
    if (!in_frontier) {
        op.skip(v);
        return 0ull;
    }
    return op.finish_deg(v, BpAcc<T>{p0, p1, any ? 1 : 0, 0}, outdeg);
}

template <class T, bool D>
__device__ unsigned long long bp_light_chunk(const BpOp<T, D> &op, const int32_t *__restrict__ off,
                                             const int32_t *__restrict__ adj,
                                             const unsigned long long *__restrict__ heads, int vs,
                                             int ve) {
    unsigned long long st = 0;
    if (vs >= ve) return st;
    const int lane = lane_id();
    const int E0 = off[vs], E1 = off[ve];
    const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes <= lane
    int vprev = vs - 1;      // destination of the edge before the next step's first edge
    T c0 = T(1), c1 = T(1);  // the open destination's product so far (uniform)
    bool cany = false, open = false;
    int cfl = 0, cod = 0;  // the open destination's frontier flag and out-degree
    // the head words of the next batch, loaded one batch ahead
    unsigned long long hn[kBpSteps];
#pragma unroll
    for (int u = 0; u < kBpSteps; u++) hn[u] = heads[min((E0 & ~63) + 64 * u, E1 - 1) >> 6];
    for (int base = E0 & ~63; base < E1; base += 64 * kBpSteps) {
        int tok[kBpSteps], own[kBpSteps], vpb[kBpSteps], fl[kBpSteps], od[kBpSteps];
        float4 w[kBpSteps];
        float2 phi[kBpSteps];
        Vec2<T> pc[kBpSteps];
        unsigned long long hw[kBpSteps];
        int vp = vprev;
        // every load is issued unconditionally, from an index clamped into the
        // chunk (a lane outside it loads an edge or vertex of the chunk and its
        // values are not used), so that no load waits behind a branch
#pragma unroll
        for (int u = 0; u < kBpSteps; u++) {
            const int bu = base + 64 * u;
            const int ec = min(max(bu + lane, E0), E1 - 1);
            unsigned long long h = hn[u];
            if (bu >= E1) h = 0ull;
            if (bu < E0) h &= ~0ull << (E0 - bu);  // the previous chunk's vertices
            hw[u] = h;
            vpb[u] = vp;
            own[u] = vp + __popcll(h & le);
            vp += __popcll(h);
            const int oc = min(max(own[u], vs), ve - 1);
            tok[u] = edge_src(op, adj, ec);
            w[u] = op.pot[ec];
            phi[u] = op.vpot[oc];
            pc[u] = op.cur[oc];
            fl[u] = op.F[oc];
            od[u] = op.out_off[oc + 1] - op.out_off[oc];
        }
#pragma unroll
        for (int u = 0; u < kBpSteps; u++)
            hn[u] = heads[min(base + 64 * (kBpSteps + u), E1 - 1) >> 6];
        vprev = vp;
#pragma unroll
        for (int u = 0; u < kBpSteps; u++) {
            const int bu = base + 64 * u;
            if (bu >= E1) break;
            const unsigned long long h = hw[u];
            if (open && (h & 1ull)) {  // the carried destination ended at bu - 1
                if (lane == 0) st += bp_close(op, vpb[u], c0, c1, cany, cfl, cod);
                open = false;
                c0 = T(1);
                c1 = T(1);
                cany = false;
            }
            const int e = bu + lane;
            const bool valid = e >= E0 && e < E1;
            T b0 = T(1), b1 = T(1);
            const bool act = valid && op.src_active(tok[u]);
            if (act) {
                const Vec2<T> q{(T)phi[u].x * pc[u].x, (T)phi[u].y * pc[u].y};
                op.message(w[u], q, b0, b1);
            }
            // inclusive segmented product over the lanes; a segment starts at a head
            const unsigned long long hl = h & le;
            const int ss = hl ? 63 - __builtin_clzll(hl) : -1;
            seg_scan_mul(b0, b1, lane, ss);
            const unsigned long long A = __ballot(act);
            bool any = (A & le & (ss >= 0 ? ~((1ull << ss) - 1ull) : ~0ull)) != 0ull;
            if (ss < 0) {  // the lanes before the step's first head continue the carry
                b0 = c0 * b0;
                b1 = c1 * b1;
                any = any || cany;
            }
            const bool tail = valid && (e + 1 == E1 || (lane < 63 && ((h >> (lane + 1)) & 1ull)));
            if (tail) st += bp_close(op, own[u], b0, b1, any, fl[u], od[u]);
            // lane 63's destination may run on into the next step: carry it (the
            // carry taken in has been used by the lanes before the first head, or
            // is in lane 63's product when the step has no head)
            open = bu + 64 < E1;
            if (open) {
                c0 = readlane_t(b0, 63);
                c1 = readlane_t(b1, 63);
                cany = (__ballot(any) >> 63) != 0ull;
                cfl = __builtin_amdgcn_readlane(fl[u], 63);
                cod = __builtin_amdgcn_readlane(od[u], 63);
            }
        }
    }
    return st;
}

// blocks [0, hot_blocks): the degree-binned pull of the heavy destinations and the
// vertices without in-edges; the rest: one light chunk per wave
template <class T, bool D>
__global__ __launch_bounds__(kBlock) void k_bp_pull(BpOp<T, D> op, PullArgs pa, int hot_blocks,
                                                   BpLight L, unsigned long long *stat_out) {
    unsigned long long st = 0;
    if ((int)blockIdx.x < hot_blocks) {
        st = pull_dispatch(op, pa, (int)blockIdx.x);
    } else {
        const int c = __builtin_amdgcn_readfirstlane(((int)blockIdx.x - hot_blocks) * (kBlock / 64) +
                                                     (int)(threadIdx.x >> 6));
        if (c < L.nchunks)
            st = bp_light_chunk(op, pa.off, pa.adj, L.heads, L.chunk_v[c], L.chunk_v[c + 1]);
    }
    commit_block_stat<BpOp<T, D>>(st, stat_out);
}

// chunk c's first vertex: the first position v in [lb, be] with off[v] >= off[lb] + c*K
__global__ void k_bp_chunks(const int32_t *off, int lb, int be, int nch, int32_t *chunk_v) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nch) return;
    if (c == nch) {
        chunk_v[c] = be;
        return;
    }
    const long long target = (long long)off[lb] + (long long)c * kBpChunkEdges;
    int lo = lb, hi = be;  // off[hi] >= target
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if ((long long)off[mid] >= target) hi = mid;
        else lo = mid + 1;
    }
    chunk_v[c] = lo;
}

__global__ void k_bp_mark_heads(const int32_t *off, int lb, int be, unsigned long long *heads) {
    for (int64_t v = lb + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < be;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int e = off[v];
        atomicOr(&heads[e >> 6], 1ull << (e & 63));
    }
}

// The light range and its chunks, built once per schedule.  Returns the blocks of
// the degree-binned part in *hot (its segments are copied into *pa) and the
// light-range description in *L.
static int bp_light_setup(pg_graph *g, PullArgs *pa, int *hot, BpLight *L) {
    const PullSchedule &s = g->in_sched;
    const PullArgs &a = s.args;
    const int64_t n = g->n, be = s.body_end;
    int64_t lb = be;
    int i = 0;
    while (i < a.nseg && !(a.seg[i].g >= 1 && a.seg[i].g <= kBpLightMaxG)) i++;
    if (i < a.nseg) lb = std::min<int64_t>(a.seg[i].v_begin, be);
    *pa = a;
    pa->nseg = i;
    int64_t blocks = i < a.nseg ? a.seg[i].block_begin : s.nblocks;
    if (i == a.nseg) blocks = s.nblocks;  // no light class: the schedule unchanged
    if (i < a.nseg && be < n) {           // the vertices without in-edges, G = 1
        if (pa->nseg >= kMaxSegs) {
            set_error("bp: too many segments");
            return PG_ERR_STATE;
        }
        pa->seg[pa->nseg++] = PullSeg{(int32_t)be, (int32_t)n, 1, (int32_t)blocks};
        blocks += (n - be + kBlock - 1) / kBlock;
    }
    *hot = (int)blocks;
    if (i == a.nseg) {
        *L = BpLight{nullptr, nullptr, 0};
        return PG_OK;
    }
    if (!g->bp_light_built || g->bp_light_lb != lb || g->bp_light_be != be) {
        int32_t e2[2];
        PG_CHECK_HIP(hipMemcpyAsync(&e2[0], g->in_off + lb, 4, hipMemcpyDeviceToHost, stream()));
        PG_CHECK_HIP(hipMemcpyAsync(&e2[1], g->in_off + be, 4, hipMemcpyDeviceToHost, stream()));
        PG_CHECK_HIP(hipStreamSynchronize(stream()));
        const int64_t nch = ((int64_t)e2[1] - e2[0] + kBpChunkEdges - 1) / kBpChunkEdges;
        int32_t *cv = nullptr;
        unsigned long long *heads = nullptr;
        const int64_t nw = g->m / 64 + 2;
        PG_TRY(workspace_t(g, "bp_chunks", (size_t)nch + 1, &cv));
        PG_TRY(workspace_t(g, "bp_heads", (size_t)nw, &heads));
        PG_LAUNCH_K(k_zero_u64, grid_for(nw), heads, nw);
        if (be > lb)
            PG_LAUNCH_K(k_bp_mark_heads, grid_for(be - lb), g->in_off, (int)lb, (int)be, heads);
        PG_LAUNCH_K(k_bp_chunks, (unsigned)((nch + 1 + kBlock - 1) / kBlock), g->in_off, (int)lb,
                    (int)be, (int)nch, cv);
        g->bp_light_lb = lb;
        g->bp_light_be = be;
        g->bp_light_nch = (int)nch;
        g->bp_light_built = true;
    }
    int32_t *cv = nullptr;
    unsigned long long *heads = nullptr;
    PG_TRY(workspace_t(g, "bp_chunks", (size_t)g->bp_light_nch + 1, &cv));
    PG_TRY(workspace_t(g, "bp_heads", (size_t)(g->m / 64 + 2), &heads));
    *L = BpLight{cv, heads, g->bp_light_nch};
    const int64_t tot = blocks + ((int64_t)g->bp_light_nch + kBlock / 64 - 1) / (kBlock / 64);
    if (tot > 0x7FFFFFFF) {
        set_error("bp: grid too large");
        return PG_ERR_UNSUPPORTED;
    }
    return PG_OK;
}

// ---- the iteration at which each vertex leaves the frontier -----------------------
// One pull per iteration of the frontier alone (F_{t+1}[v] = some in-neighbour in
// F_t, stopping at the first), until no vertex leaves: those left never do (255).
// Built once per graph.  A graph whose frontier still shrinks after 254 steps
// keeps the bitmap test (bp_death_state = -1).
struct BpReachOp {
    using Acc = int;
    using Ctx = int;
    using Stat = unsigned long long;
    static constexpr bool kEarlyExit = true;
    static constexpr bool kStats = true;  // vertices that left
    const uint8_t *F;
    uint8_t *Fn;
    uint8_t *D;
    const uint32_t *Fb;
    int t1;  // the iteration whose frontier is formed
    __device__ bool active(int v) const { return F[v] != 0; }
    __device__ Ctx ctx(int) const { return 0; }
    __device__ Acc identity() const { return 0; }
    __device__ Acc combine(Acc a, Acc b) const { return a | b; }
    __device__ Acc edge(int s, int, const Ctx &) const { return bit_test(Fb, s) ? 1 : 0; }
    __device__ bool found(Acc a, const Ctx &) const { return a != 0; }
    __device__ Stat finish(int v, Acc a, const Ctx &) const {
        Fn[v] = a ? 1 : 0;
        if (a) return 0ull;
        D[v] = (uint8_t)t1;
        return 1ull;
    }
    __device__ void skip(int v) const { Fn[v] = 0; }
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a + b; }
    __device__ static void stat_commit(Stat *p, Stat v) {
        atomicAdd(&p[blockIdx.x & (kStatSlots - 1)], v);
    }
};

__global__ void k_gather_u8(const uint8_t *src, const int32_t *idx, int64_t m, uint8_t *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

static int bp_death_times(pg_graph *g) {
    if (g->bp_death_state != 0) return PG_OK;
    const int64_t n = g->n, m = g->m;
    uint8_t *D = nullptr, *fa = nullptr, *fb = nullptr, *de = nullptr;
    uint32_t *bits = nullptr;
    unsigned long long *cnt = nullptr;
    PG_TRY(workspace_t(g, "bp_dv", (size_t)n, &D));
    PG_TRY(workspace_t(g, "bp_de", (size_t)m, &de));
    PG_TRY(workspace_t(g, "bp_df0", (size_t)n, &fa));
    PG_TRY(workspace_t(g, "bp_df1", (size_t)n, &fb));
    PG_TRY(workspace_t(g, "bp_fbits", (size_t)(n + 31) / 32 + 1, &bits));
    PG_TRY(workspace_t(g, "bp_dcnt", (size_t)kStatSlots, &cnt));
    PG_LAUNCH_K(k_fill_u8, grid_for(n), D, (uint8_t)255, n);
    PG_LAUNCH_K(k_fill_u8, grid_for(n), fa, (uint8_t)1, n);
    int state = 0;
    for (int t = 0; t < 255 && state == 0; t++) {
        PG_LAUNCH_K(k_zero_u64, 1, cnt, kStatSlots);
        PG_LAUNCH_K(k_flags_to_bits, grid_for(n), fa, n, bits);
        PG_TRY(launch_pull(BpReachOp{fa, fb, D, bits, t + 1}, g->in_sched, cnt));
        unsigned long long h[kStatSlots];
        PG_TRY(fetch_words(h, cnt, sizeof(h)));
        unsigned long long left = 0;
        for (int k = 0; k < kStatSlots; k++) left += h[k];
        if (left == 0) state = 1;            // fixed point: the rest stay (255)
        else if (t + 1 == 255) state = -1;   // a 255th iteration would be needed
        std::swap(fa, fb);
    }
    if (state == 1) PG_LAUNCH_K(k_gather_u8, grid_for(m), D, g->in_adj, m, de);
    g->bp_death_state = state;
    return PG_OK;
}

template <class T, bool Death>
static int bp_iterations(pg_graph *g, int iters, const float4 *pot, const float2 *vp,
                         Vec2<T> *&cur, Vec2<T> *&nxt, uint8_t *&F, uint8_t *&Fn,
                         unsigned long long *stats, const PullArgs &pa, int hot, const BpLight &L,
                         KernelTimer &kt) {
    const int64_t n = g->n;
    uint32_t *Fb = nullptr;
    if (!Death) PG_TRY(workspace_t(g, "bp_fbits", (size_t)(n + 31) / 32 + 1, &Fb));
    const uint8_t *de = nullptr;
    if (Death) {
        uint8_t *p = nullptr;
        PG_TRY(workspace_t(g, "bp_de", (size_t)g->m, &p));
        de = p;
    }
    const int nb = hot + (L.nchunks + kBlock / 64 - 1) / (kBlock / 64);
    for (int t = 0; t < iters; t++) {
        BpOp<T, Death> op{F, Fn, pot, vp, cur, nxt, g->out_off, Fb, de, t};
        PG_TRY(kt.begin());
        if (!Death) PG_LAUNCH_K(k_flags_to_bits, grid_for(n), F, n, Fb);
        if (nb > 0) {
            hipLaunchKernelGGL((k_bp_pull<T, Death>), dim3(nb), dim3(kBlock), 0, stream(), op, pa, hot,
                               L, &stats[(size_t)t * kStatSlots]);
            PG_CHECK_HIP(hipGetLastError());
        }
        PG_TRY(kt.end());
        std::swap(cur, nxt);  // swap(vertD_curr, vertD_next), numa-BP.C:303-304
        std::swap(F, Fn);     // switchFrontier
    }
    return PG_OK;
}

template <class T>
static int bp_run(pg_graph *g, int iters, const float4 *pot, const float2 *vp, const float *p0,
                  uint8_t *f0, uint8_t *f1, double *prod, uint32_t flags, pg_stats *st) {
    const int64_t n = g->n;
    Vec2<T> *pa = nullptr, *pb = nullptr;
    PG_TRY(workspace_t(g, sizeof(T) == 8 ? "bp_pa64" : "bp_pa32", (size_t)n, &pa));
    PG_TRY(workspace_t(g, sizeof(T) == 8 ? "bp_pb64" : "bp_pb32", (size_t)n, &pb));
    PG_LAUNCH_K(k_bp_state_init<T>, grid_for(n), p0, n, pa, pb);
    PullArgs hpa{};
    int hot = 0;
    BpLight L{};
    PG_TRY(bp_light_setup(g, &hpa, &hot, &L));
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    Vec2<T> *cur = pa, *nxt = pb;
    uint8_t *F = f0, *Fn = f1;
    unsigned long long *stats = nullptr;  // per iteration: sum of out-degrees of F_{t+1}
    PG_TRY(workspace_t(g, "bp_stats", ((size_t)iters + 1) * kStatSlots, &stats));
    PG_LAUNCH_K(k_zero_u64, (unsigned)(((iters + 1) * kStatSlots + kBlock - 1) / kBlock), stats,
                (iters + 1) * kStatSlots);
    if (g->bp_death_state == 1)
        PG_TRY((bp_iterations<T, true>(g, iters, pot, vp, cur, nxt, F, Fn, stats, hpa, hot, L, kt)));
    else
        PG_TRY((bp_iterations<T, false>(g, iters, pot, vp, cur, nxt, F, Fn, stats, hpa, hot, L, kt)));
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (st) {
        std::vector<unsigned long long> hs(((size_t)iters + 1) * kStatSlots);
        PG_CHECK_HIP(hipMemcpy(hs.data(), stats, 8 * hs.size(), hipMemcpyDeviceToHost));
        int64_t e = iters > 0 ? g->m : 0;  // F_0 = V
        for (int t = 0; t + 1 < iters; t++)
            for (int k = 0; k < kStatSlots; k++) e += (int64_t)hs[(size_t)t * kStatSlots + k];
        st->rounds = iters;
        st->dense_rounds = iters;
        st->edges = e;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    if (!prod) return PG_OK;
    double *tmp = nullptr;
    PG_TRY(workspace_t(g, "bp_out", 2 * (size_t)n, &tmp));
    PG_LAUNCH_K(k_bp_out<T>, grid_for(n), cur, g->perm, n, tmp);
    PG_CHECK_HIP(hipMemcpyAsync(prod, tmp, 16 * n, hipMemcpyDeviceToHost, stream()));
    PG_CHECK_HIP(hipStreamSynchronize(stream()));
    return PG_OK;
}

extern "C" int pg_bp(pg_graph *g, int iters, const float *vpot, const float *epot,
                     const float *prod0, double *prod, uint32_t flags, pg_stats *st) {
    PG_TRY(check_graph(g));
    PG_REQUIRE(iters >= 0 && vpot && prod0, PG_ERR_ARG, "bad argument");
    PG_REQUIRE(epot || g->bp_pot, PG_ERR_STATE,
               "no edge potentials: pass epot or call pg_graph_bp_potentials");
    PG_TRY(bp_death_times(g));  // once per graph
    const int64_t n = g->n, m = g->m;
    float4 *pot = nullptr;
    if (epot) {
        float4 *tmp = nullptr;
        PG_TRY(workspace_t(g, "bp_pot_in", (size_t)m + 1, &tmp));
        PG_TRY(workspace_t(g, "bp_pot", (size_t)m + 1, &pot));
        PG_CHECK_HIP(hipMemcpyAsync(tmp, epot, 16 * m, hipMemcpyHostToDevice, stream()));
        PG_LAUNCH_K(k_gather_pot, grid_for(m), tmp, g->in_eid, m, pot);
    } else {
        pot = reinterpret_cast<float4 *>(g->bp_pot);
    }
    float2 *vp = nullptr;
    float *p0 = nullptr;
    uint8_t *f0 = nullptr, *f1 = nullptr;
    PG_TRY(workspace_t(g, "bp_vpot", (size_t)n, &vp));
    PG_TRY(workspace_t(g, "bp_p0", 2 * (size_t)n, &p0));
    PG_TRY(workspace_t(g, "bp_f0", (size_t)n, &f0));
    PG_TRY(workspace_t(g, "bp_f1", (size_t)n, &f1));
    PG_TRY(from_input_order_f32(g, vpot, reinterpret_cast<float *>(vp), 2));
    PG_TRY(from_input_order_f32(g, prod0, p0, 2));
    PG_LAUNCH_K(k_fill_u8, grid_for(n), f0, (uint8_t)1, n);
    if (flags & PG_BP_FP32)
        return bp_run<float>(g, iters, pot, vp, p0, f0, f1, prod, flags, st);
    return bp_run<double>(g, iters, pot, vp, p0, f0, f1, prod, flags, st);
}

// ---- asynchronous traversal (edgeMapSparseAsyncPipe, polymer.h:1512-1751;
// numa-BFS-async-pipe.C:140-175) -------------------------------------------------------
// The reference runs one edgeMap call to quiescence: workers take chunks of queued
// vertices, claim each unvisited out-neighbour with a CAS of its parent from -1
// (BFS_F), queue the claimed ones in chunks for the next worker, and stop when
// every queue is empty and no chunk is in flight.  The parents form a tree of the
// vertices reachable from the source (not, in general, a breadth-first one).
//
// Here one persistent grid works on a single queue of work items: a claimed vertex
// of out-degree D is queued as ceil(D / kAsyncItemEdges) items (vertex, chunk), so
// a hub's out-edges are spread over many waves instead of one wave's serial walk.
// tail counts the queued items, processed those whose edges have all been claimed
// or found visited (the vertices reached are counted from the visited bitmap
// afterwards).  A wave takes kAsyncChunk
// consecutive slots and, until each of its lanes' slots is settled, expands
// whichever of them are filled (so a wave never waits on work it holds); a slot is
// settled empty once processed == tail (nothing is being expanded and nothing is
// queued: the traversal is over) and it lies at or past the tail, or when it lies
// past the queue's capacity (n + m / kAsyncItemEdges + 1 bounds every item
// count).  A wave's wait is bounded: kAsyncPolls polls in a row without any item
// processed anywhere (the count it reads unchanged), after which it sets the
// error word and leaves, so a launch always ends.  (Waves that instead took only
// items already queued, by a compare-and-swap on head, all polled the same three
// control words: the traffic on those lines starved the working waves.)
constexpr int kAsyncChunk = 64;
constexpr unsigned kAsyncPolls = 1u << 20;
constexpr int kAsyncItemEdges = 256;
constexpr int kAsyncLaneDeg = 32;  // an item's edges its lane expands alone; more: the wave
constexpr unsigned long long kAsyncEmpty = ~0ull;
// control words, each on its own 64-byte line
constexpr int kAsyncHead = 0, kAsyncTail = 16, kAsyncProcessed = 32, kAsyncError = 48,
              kAsyncReached = 56, kAsyncCtlWords = 64;

__device__ __forceinline__ unsigned long long async_item(int v, int k) {
    return ((unsigned long long)(unsigned)v << 32) | (unsigned)k;
}

// claim d for v (want: this lane has an edge); winners queue their items
// The claim of d by v (BFS_F's CAS of parent[d] from -1, numa-BFS-async-pipe.C:
// 60-79): the visited bitmap (n/8 bytes, held in L2) is screened with an L2 load
// and claimed with an OR whose old word picks the one winner, which then stores
// parent[d]; parent is written only by winners and never read during the traversal.
// The winners' items are queued at once (holding them back to queue them in
// batches starved the other waves of work: 3x slower); the vertices reached are
// counted from the bitmap afterwards.
__device__ __forceinline__ void async_claim(int v, int d, int32_t *parent, uint32_t *visited,
                                            const int32_t *__restrict__ out_off,
                                            unsigned long long *q, unsigned *ctl, bool want) {
    bool won = false;
    if (want) {
        const uint32_t bit = 1u << (d & 31);
        uint32_t *word = &visited[d >> 5];
        if (!(__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit))
            won = (atomicOr(word, bit) & bit) == 0u;
    }
    if (!__ballot(won)) return;
    const int lane = lane_id();
    int cnt = 0;
    if (won) {
        parent[d] = v;
        cnt = (out_off[d + 1] - out_off[d] + kAsyncItemEdges - 1) / kAsyncItemEdges;
    }
    int x = cnt;  // inclusive scan of the item counts over the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    const int total = __shfl(x, 63, 64);
    if (!total) return;
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(ctl + kAsyncTail, (unsigned)total);
    base = __shfl(base, 0, 64);
    for (int j = 0; j < cnt; j++)
        __hip_atomic_store(&q[base + (unsigned)(x - cnt + j)], async_item(d, j), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBlock) void k_bfs_async(const int32_t *__restrict__ out_off,
                                                      const int32_t *__restrict__ out_adj,
                                                      int32_t *parent, uint32_t *visited,
                                                      unsigned long long *q, int64_t cap,
                                                      unsigned *ctl) {
    unsigned *head = ctl + kAsyncHead, *tail = ctl + kAsyncTail;
    unsigned *processed = ctl + kAsyncProcessed, *error = ctl + kAsyncError;
    const int lane = lane_id();
    while (true) {
        unsigned i0 = 0;
        if (lane == 0) i0 = atomicAdd(head, (unsigned)kAsyncChunk);
        i0 = __shfl(i0, 0, 64);
        if ((int64_t)i0 >= cap) return;  // every slot of the chunk lies past the queue
        const int64_t i = (int64_t)i0 + lane;
        bool pending = i < cap;
        bool over = false;
        unsigned polls = 0, seen = ~0u;
        while (__ballot(pending)) {
            unsigned long long it = kAsyncEmpty;
            if (pending)
                it = __hip_atomic_load(&q[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const bool ready = it != kAsyncEmpty;
            const unsigned long long rmask = __ballot(ready);
            if (rmask) {
                // expand the filled items: short ones by their own lane, longer ones
                // by the whole wave, one at a time
                int v = 0, beg = 0, end = 0;
                if (ready) {
                    v = (int)(it >> 32);
                    const int k = (int)(it & 0xFFFFFFFFull);
                    const int o0 = out_off[v], o1 = out_off[v + 1];
                    beg = o0 + k * kAsyncItemEdges;
                    end = min(o1, beg + kAsyncItemEdges);
                }
                const bool light = ready && end - beg <= kAsyncLaneDeg;
                for (int k = 0; __ballot(light && beg + k < end); k++) {
                    const bool act = light && beg + k < end;
                    async_claim(v, act ? out_adj[beg + k] : 0, parent, visited, out_off, q, ctl,
                                act);
                }
                unsigned long long heavy = __ballot(ready && !light);
                while (heavy) {
                    const int l = __ffsll((long long)heavy) - 1;
                    heavy &= heavy - 1;
                    const int hv = __shfl(v, l, 64);
                    const int hb = __shfl(beg, l, 64), he = __shfl(end, l, 64);
                    for (int e0 = hb; e0 < he; e0 += 64) {
                        const bool act = e0 + lane < he;
                        async_claim(hv, act ? out_adj[e0 + lane] : 0, parent, visited, out_off, q,
                                    ctl, act);
                    }
                }
                // the items this wave queued are visible before it counts its own
                // as processed
                __threadfence();
                if (lane == 0) atomicAdd(processed, (unsigned)__popcll(rmask));
                if (ready) pending = false;
                polls = 0;
                continue;
            }
            // nothing filled: settled empty once the traversal is over
            unsigned pr = 0, tl = 0;
            if (lane == 0) {
                pr = __hip_atomic_load(processed, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                tl = __hip_atomic_load(tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            }
            pr = __shfl(pr, 0, 64);
            tl = __shfl(tl, 0, 64);
            if (pr != seen) {  // progress elsewhere: the wait is not stuck
                seen = pr;
                polls = 0;
            }
            if (pr == tl) {
                over = true;
                if (pending && i >= (int64_t)tl) pending = false;
            }
            if (++polls > kAsyncPolls) {
                if (lane == 0) atomicOr(error, 1u);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (over) return;  // the traversal has ended: no slot will be filled again
    }
}

// the source claimed by itself, its items queued
__global__ void k_async_seed(int32_t src, const int32_t *out_off, int32_t *parent,
                             uint32_t *visited, unsigned long long *q, unsigned *ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        parent[src] = src;
        visited[src >> 5] |= 1u << (src & 31);
        const int cnt = (out_off[src + 1] - out_off[src] + kAsyncItemEdges - 1) / kAsyncItemEdges;
        for (int j = 0; j < cnt; j++) q[j] = async_item(src, j);
        ctl[kAsyncTail] = (unsigned)cnt;
    }
}

// the vertices reached: the set bits of the visited bitmap
__global__ void k_async_reached(const uint32_t *visited, int64_t nw, unsigned *out) {
    unsigned c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw;
         i += (int64_t)gridDim.x * blockDim.x)
        c += __popc(visited[i]);
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane_id() == 0 && c) atomicAdd(out, c);
}

extern "C" int pg_bfs_async(pg_graph *g, int64_t src, int32_t *parent, int64_t *visited,
                            uint32_t flags, pg_stats *st) {
    PG_TRY(check_graph(g));
    // queue capacity: every vertex's items, ceil(D / R) <= D / R + 1
    const int64_t cap = g->n + g->m / kAsyncItemEdges + 1;
    PG_REQUIRE(cap < 0x7FFFFFFF - (int64_t)kAsyncChunk * 65536, PG_ERR_UNSUPPORTED,
               "graph too large for the async queue");
    int32_t s_int;
    PG_TRY(lookup_internal(g, src, &s_int));
    int32_t *par = nullptr;
    unsigned long long *q = nullptr;
    unsigned *ctl = nullptr;
    uint32_t *vis = nullptr;
    const int64_t nw = (g->n + 31) / 32;
    PG_TRY(workspace_t(g, "async_parent", (size_t)g->n, &par));
    PG_TRY(workspace_t(g, "async_visited", (size_t)nw, &vis));
    PG_TRY(workspace_t(g, "async_queue", (size_t)cap, &q));
    PG_TRY(workspace_t(g, "async_ctl", (size_t)kAsyncCtlWords, &ctl));
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    PG_TRY(fill(par, -1, g->n));
    PG_TRY(fill(vis, 0u, nw));
    PG_TRY(fill(q, kAsyncEmpty, cap));
    PG_TRY(fill(ctl, 0u, kAsyncCtlWords));
    PG_LAUNCH_K(k_async_seed, 1, s_int, g->out_off, par, vis, q, ctl);
    PG_TRY(kt.begin());
    // one workgroup per CU: the waiting waves' polls contend, so fewer waves finish
    // sooner (S24, per source: 2048 workgroups 18.9-22.9 ms, 1024 17.5-19.4, 512
    // 17.1-17.7, 256 16.5-16.9, 128 21.1-21.3)
    PG_LAUNCH_K(k_bfs_async, 256, g->out_off, g->out_adj, par, vis, q, cap, ctl);
    PG_TRY(kt.end());
    PG_LAUNCH_K(k_async_reached, grid_for(nw, 1024), vis, nw, ctl + kAsyncReached);
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    unsigned h[kAsyncCtlWords];
    PG_CHECK_HIP(hipMemcpy(h, ctl, sizeof(h), hipMemcpyDeviceToHost));
    PG_REQUIRE(h[kAsyncError] == 0, PG_ERR_STATE, "async traversal: a wave's wait ran out");
    PG_REQUIRE(h[kAsyncProcessed] == h[kAsyncTail], PG_ERR_STATE,
               "async traversal ended with %u of %u queued items processed", h[kAsyncProcessed],
               h[kAsyncTail]);
    if (visited) *visited = (int64_t)h[kAsyncReached];
    if (st) {
        st->rounds = 1;
        st->dense_rounds = 0;
        st->active = (int64_t)h[kAsyncReached];
        st->edges = 0;
        st->ms = ms;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    return to_input_order_i32(g, par, parent, true);
}

extern "C" int pg_bfs(pg_graph *g, int64_t src, int64_t threshold, int32_t *level,
                      int32_t *parent, uint32_t flags, pg_stats *st) {
    PG_TRY(check_graph(g));
    int32_t *lv = nullptr, *pr = nullptr;
    PG_TRY(bfs_run(g, src, threshold, flags, st, &lv, &pr));
    if (parent) PG_LAUNCH_K(k_parent_mask, grid_for(g->n), lv, g->n, pr);
    PG_TRY(to_input_order_i32(g, lv, level, false));
    return to_input_order_i32(g, pr, parent, true);
}

extern "C" int pg_cc(pg_graph *g, int64_t threshold, int32_t *label, uint32_t flags,
                     pg_stats *st) {
    PG_TRY(check_graph(g));
    int32_t *lb = nullptr;
    PG_TRY(cc_run(g, threshold, flags, st, &lb));
    return to_input_order_i32(g, lb, label, false);
}

extern "C" int pg_sssp(pg_graph *g, int64_t src, int64_t threshold, int32_t *dist, uint32_t flags,
                       pg_stats *st) {
    PG_TRY(check_graph(g));
    int32_t *d = nullptr;
    PG_TRY(sssp_run(g, src, threshold, flags, st, &d));
    return to_input_order_i32(g, d, dist, false);
}
