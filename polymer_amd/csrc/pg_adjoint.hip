// pg_adjoint.hip -- weighted PageRank and its adjoint.
//
// Forward:  x = d W x + (1-d) v,  W[dst][s] = wn[e] for the edge e = s->dst, solved by the
// fixed-point iteration as a pull over the CSC (the transposed push of numa-PageRank.C).
// Adjoint of the loss L = g^T x:
//     lambda = d W^T lambda + g         -- the same fixed point, pulled over the CSR
//     dL/dwn[e] = d * lambda[dst] * x[s],   dL/dv = (1-d) lambda
// The backward of the forward's pull over in-edges is a pull over out-edges, so both
// directions run on the degree-binned engine.  Iterates are fp32 with fp32 accumulation;
// the convergence residual sum |x_{t+1} - x_t| is reduced in fp64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pg_engine.h"
#include "pg_internal.h"

namespace pg {

#define PG_LAUNCH_A(kern, grid, ...)                                                  \
    do {                                                                              \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream(), __VA_ARGS__); \
        PG_CHECK_HIP(hipGetLastError());                                              \
    } while (0)

// the residual statistic, and the gate of the device-side convergence test: a
// launch enqueued after the test found convergence returns at once
struct ResidualStat {
    const int *gate;
    static constexpr bool kGated = true;
    __device__ bool gated() const { return *gate != 0; }
    using Stat = double;
    __device__ static Stat stat_reduce(Stat a, Stat b) { return a + b; }
    __device__ static void stat_commit(Stat *p, Stat v) {
        atomicAdd(&p[blockIdx.x & (kStatSlots - 1)], v);
    }
};

// x_{t+1}[v] = d * sum_{e=(s->v)} w[e] x_t[s] + (1-d) tele[v]
struct WprOp : ResidualStat {
    using Acc = float;
    using Ctx = int;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = true;
    const float *x;
    float *xn;
    const float *w;  // CSC order
    const float *tele;
    float damp, omd;
    __device__ bool active(int) const { return true; }
    __device__ int ctx(int v) const { return v; }
    __device__ float identity() const { return 0.f; }
    __device__ float combine(float a, float b) const { return a + b; }
    __device__ float edge(int s, int e, const int &) const { return w[e] * x[s]; }
    __device__ bool found(float, const int &) const { return false; }
    __device__ double finish(int v, float sum, const int &) const {
        const float r = damp * sum + omd * tele[v];
        xn[v] = r;
        return fabs((double)r - (double)x[v]);
    }
    __device__ void skip(int) const {}
};

// lambda_{t+1}[s] = d * sum_{e=(s->dst)} w[e] lambda_t[dst] + g[s]   (pull over the CSR)
struct WprAdjOp : ResidualStat {
    using Acc = float;
    using Ctx = int;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = true;
    const float *lam;
    float *lamn;
    const float *w;  // CSR order
    const float *gv;
    float damp;
    __device__ bool active(int) const { return true; }
    __device__ int ctx(int v) const { return v; }
    __device__ float identity() const { return 0.f; }
    __device__ float combine(float a, float b) const { return a + b; }
    __device__ float edge(int dst, int e, const int &) const { return w[e] * lam[dst]; }
    __device__ bool found(float, const int &) const { return false; }
    __device__ double finish(int s, float sum, const int &) const {
        const float r = damp * sum + gv[s];
        lamn[s] = r;
        return fabs((double)r - (double)lam[s]);
    }
    __device__ void skip(int) const {}
};

// grad[e] = d * lambda[dst] * x[s] for every out-edge e of s (CSR order)
struct WprGradOp {
    using Acc = float;
    using Ctx = float;  // d * x[s]
    using Stat = double;
    static constexpr bool kEarlyExit = false;
    static constexpr bool kStats = false;
    const float *lam;
    const float *x;
    float *grad;
    float damp;
    __device__ bool active(int) const { return true; }
    __device__ float ctx(int s) const { return damp * x[s]; }
    __device__ float identity() const { return 0.f; }
    __device__ float combine(float a, float b) const { return a + b; }
    __device__ float edge(int dst, int e, const float &dx) const {
        grad[e] = dx * lam[dst];
        return 0.f;
    }
    __device__ bool found(float, const float &) const { return false; }
    __device__ double finish(int, float, const float &) const { return 0.0; }
    __device__ void skip(int) const {}
};

__global__ void k_copy_f32(const float *a, int64_t n, float *b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void k_scale_f32(const float *a, int64_t n, float s, float *b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = s * a[i];
}

// After an iteration: its residual (the 64 slots summed in order, as a host loop
// would), kept in hist[t]; the first t whose residual is <= tol sets the gate
// (state[0]) and records t + 1 iterations used (state[1]).
__global__ void k_res_check(const double *slots, double tol, int t, int *state, double *hist) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || state[0]) return;
    double s = 0;
    for (int i = 0; i < kStatSlots; i++) s += slots[i];
    hist[t] = s;
    if (s <= tol) {
        state[1] = t + 1;
        state[0] = 1;
    }
}

// Fixed-point loop shared by the forward and the adjoint.  Iterations are enqueued
// in batches of kFixBatch, each followed by the device-side test; the host reads the
// state once per batch (not the residual after every iteration).  Launches past
// convergence are gated off, so the result is the iterate of the first iteration
// whose residual is <= tol, as with a test on the host after every iteration.
constexpr int kFixBatch = 16;

template <class MakeOp>
static int fixed_point(pg_graph *g, const PullSchedule &sched, float *a, float *b, int max_iters,
                       double tol, KernelTimer &kt, MakeOp make, float **result, int *iters_used,
                       double *residual) {
    double *slots = nullptr, *hist = nullptr;
    int *state = nullptr;
    PG_TRY(workspace_t(g, "wpr_res", (size_t)kStatSlots, &slots));
    PG_TRY(workspace_t(g, "wpr_res_hist", (size_t)max_iters, &hist));
    PG_TRY(workspace_t(g, "wpr_state", 4, &state));
    PG_TRY(fill(state, 0, 4));
    float *cur = a, *nxt = b;
    int h[4] = {0, 0, 0, 0};
    for (int t0 = 0; t0 < max_iters && !h[0]; t0 += kFixBatch) {
        for (int t = t0; t < std::min(max_iters, t0 + kFixBatch); t++) {
            PG_TRY(fill(slots, 0.0, kStatSlots));
            auto op = make(state, cur, nxt);
            PG_TRY(kt.begin());
            PG_TRY(launch_pull(op, sched, slots));
            PG_TRY(kt.end());
            hipLaunchKernelGGL(k_res_check, dim3(1), dim3(64), 0, stream(), slots, tol, t, state, hist);
            PG_CHECK_HIP(hipGetLastError());
            std::swap(cur, nxt);
        }
        PG_TRY(fetch_words(h, state, sizeof(h)));
    }
    const int used = h[0] ? h[1] : max_iters;
    // iteration t writes b when t is even, a when it is odd
    *result = (used % 2 == 1) ? b : a;
    *iters_used = used;
    PG_CHECK_HIP(hipMemcpy(residual, hist + used - 1, sizeof(double), hipMemcpyDeviceToHost));
    return PG_OK;
}

}  // namespace pg

using namespace pg;

extern "C" int pg_wpr_forward(pg_graph *g, const float *wn, const float *tele, double damping,
                              int max_iters, double tol, float *x, uint32_t flags, pg_stats *st) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(g && wn && tele, PG_ERR_ARG, "null argument");
    PG_REQUIRE(max_iters >= 1 && damping >= 0 && damping < 1, PG_ERR_ARG, "bad iteration setup");
    const int64_t n = g->n, m = g->m;
    float *w_in = nullptr, *w_out = nullptr, *tv = nullptr, *xa = nullptr, *xb = nullptr,
          *xs = nullptr;
    PG_TRY(workspace_t(g, "wpr_w_in", (size_t)m, &w_in));
    PG_TRY(workspace_t(g, "wpr_w_out", (size_t)m, &w_out));
    PG_TRY(workspace_t(g, "wpr_tele", (size_t)n, &tv));
    PG_TRY(workspace_t(g, "wpr_xa", (size_t)n, &xa));
    PG_TRY(workspace_t(g, "wpr_xb", (size_t)n, &xb));
    PG_TRY(workspace_t(g, "wpr_x", (size_t)n, &xs));
    PG_TRY(edges_from_input_f32(g, wn, true, w_in));
    PG_TRY(edges_from_input_f32(g, wn, false, w_out));
    PG_TRY(from_input_order_f32(g, tele, tv, 1));
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    PG_LAUNCH_A(k_copy_f32, grid_for(n), tv, n, xa);  // x_0 = v
    const float d = (float)damping, omd = (float)(1.0 - damping);
    float *res_ptr = nullptr;
    int used = 0;
    double res = 0;
    PG_TRY(fixed_point(g, g->in_sched, xa, xb, max_iters, tol, kt,
                       [&](const int *gate, float *cur, float *nxt) {
                           return WprOp{{gate}, cur, nxt, w_in, tv, d, omd};
                       },
                       &res_ptr, &used, &res));
    PG_LAUNCH_A(k_copy_f32, grid_for(n), res_ptr, n, xs);  // kept for the adjoint
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    g->wpr_valid = true;
    g->wpr_damping = damping;
    if (st) {
        st->rounds = used;
        st->dense_rounds = used;
        st->edges = (int64_t)used * m;
        st->active = (int64_t)used * n;
        st->ms = ms;
        st->residual = res;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    return to_input_order_f32(g, xs, x, 1);
}

extern "C" int pg_wpr_adjoint(pg_graph *g, const float *gvec, int max_iters, double tol,
                              float *lambda, float *grad_w, float *grad_tele, uint32_t flags,
                              pg_stats *st) {
    PG_REQUIRE(initialised(), PG_ERR_STATE, "pg_init has not been called");
    PG_REQUIRE(g && gvec, PG_ERR_ARG, "null argument");
    PG_REQUIRE(g->wpr_valid, PG_ERR_STATE, "pg_wpr_forward has not been run on this graph");
    PG_REQUIRE(max_iters >= 1, PG_ERR_ARG, "max_iters < 1");
    const int64_t n = g->n, m = g->m;
    PG_TRY(ensure_out_schedule(g));
    float *w_out = nullptr, *xs = nullptr, *gv = nullptr, *la = nullptr, *lb = nullptr,
          *gw = nullptr, *gt = nullptr;
    PG_TRY(workspace_t(g, "wpr_w_out", (size_t)m, &w_out));
    PG_TRY(workspace_t(g, "wpr_x", (size_t)n, &xs));
    PG_TRY(workspace_t(g, "wpr_g", (size_t)n, &gv));
    PG_TRY(workspace_t(g, "wpr_la", (size_t)n, &la));
    PG_TRY(workspace_t(g, "wpr_lb", (size_t)n, &lb));
    PG_TRY(workspace_t(g, "wpr_gw", (size_t)m, &gw));
    PG_TRY(workspace_t(g, "wpr_gt", (size_t)n, &gt));
    PG_TRY(from_input_order_f32(g, gvec, gv, 1));
    const float d = (float)g->wpr_damping;
    KernelTimer kt;
    kt.on = (flags & PG_PROFILE_KERNELS) != 0;
    Timer tm;
    PG_TRY(tm.start());
    PG_LAUNCH_A(k_copy_f32, grid_for(n), gv, n, la);  // lambda_0 = g
    float *lam = nullptr;
    int used = 0;
    double res = 0;
    PG_TRY(fixed_point(g, g->out_sched, la, lb, max_iters, tol, kt,
                       [&](const int *gate, float *cur, float *nxt) {
                           return WprAdjOp{{gate}, cur, nxt, w_out, gv, d};
                       },
                       &lam, &used, &res));
    WprGradOp gop{lam, xs, gw, d};
    PG_TRY(launch_pull(gop, g->out_sched));
    PG_LAUNCH_A(k_scale_f32, grid_for(n), lam, n, (float)(1.0 - g->wpr_damping), gt);
    double ms = 0;
    PG_TRY(tm.stop(&ms));
    if (st) {
        st->rounds = used;
        st->dense_rounds = used;
        st->edges = (int64_t)(used + 1) * m;
        st->active = (int64_t)used * n;
        st->ms = ms;
        st->residual = res;
        PG_TRY(kt.total(&st->kernel_ms, &st->kernel_launches));
    }
    PG_TRY(to_input_order_f32(g, lam, lambda, 1));
    PG_TRY(to_input_order_f32(g, gt, grad_tele, 1));
    return edges_to_input_f32(g, gw, false, grad_w);
}
