"""ctypes view of liboracle.so -- the CPU restatement used as the checker.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke()
and by bench.py's CPU-baseline leg; the GPU product never imports it.
Each wrapper names the C routine (oracle/polymer_oracle.c) whose reference
citation it carries.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        i64, i32, f64, u64 = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64
        sig = {
            "or_rmat_edges": (i64, [ctypes.c_int, i64, f64, f64, f64, u64, I32P, I32P]),
            "or_csr_from_sorted": (None, [i64, i64, I32P, I32P, I64P, I32P]),
            "or_transpose": (None, [i64, i64, I64P, I32P, I64P, I32P, ctypes.c_void_p]),
            "or_symmetrize": (i64, [i64, I64P, I32P, ctypes.c_void_p, ctypes.c_void_p]),
            "or_fill_weights_i32": (None, [u64, i64, I64P, I32P, I32P]),
            "or_fill_values_f32": (None, [u64, i64, I64P, I32P, F32P]),
            "or_fill_bp_pot": (None, [u64, i64, I64P, I32P, ctypes.c_float, ctypes.c_float, F32P]),
            "or_fill_vertex_f32": (None, [u64, i64, ctypes.c_float, ctypes.c_float, F32P]),
            "or_write_adj": (ctypes.c_int, [ctypes.c_char_p, i64, i64, I64P, I32P]),
            "or_write_wadj": (ctypes.c_int, [ctypes.c_char_p, i64, i64, I64P, I32P, I32P]),
            "or_write_bin": (ctypes.c_int, [ctypes.c_char_p, i64, i64, I64P, I32P, ctypes.c_int]),
            "or_read_bin": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                           ctypes.c_void_p, ctypes.c_void_p]),
            "or_read_adj": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(i64),
                                           ctypes.POINTER(i64), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]),
            "or_pagerank_push": (ctypes.c_int, [i64, I64P, I32P, ctypes.c_int, f64, F64P, I64P]),
            "or_pagerank_delta": (ctypes.c_int, [i64, I64P, I32P, ctypes.c_int, f64, f64, F64P,
                                                 I64P]),
            "or_pagerank_pull_omp": (ctypes.c_int, [i64, I64P, I64P, I32P, ctypes.c_int, f64,
                                                    ctypes.c_int, F64P]),
            "or_bfs_pull": (ctypes.c_int, [i64, I64P, I32P, i32, I32P, I32P, I32P]),
            "or_bfs_levels": (ctypes.c_int, [i64, I64P, I32P, i32, I32P]),
            "or_cc_labelprop": (ctypes.c_int, [i64, I64P, I32P, I32P, I32P]),
            "or_cc_unionfind": (ctypes.c_int, [i64, I64P, I32P, I32P]),
            "or_sssp_bf_reference": (ctypes.c_int, [i64, I64P, I32P, I32P, i32, I32P, I32P]),
            "or_sssp_bf": (ctypes.c_int, [i64, I64P, I32P, I32P, i32, I32P, I32P]),
            "or_dijkstra": (ctypes.c_int, [i64, I64P, I32P, I32P, i32, I32P]),
            "or_spmv": (ctypes.c_int, [i64, I64P, I32P, F64P, F64P, ctypes.c_int, F64P]),
            "or_bp": (ctypes.c_int, [i64, I64P, I32P, F32P, F32P, F64P, ctypes.c_int, F64P]),
            "or_spmv_adjoint": (ctypes.c_int, [i64, I64P, I32P, F64P, F64P, F64P, ctypes.c_int,
                                               F64P, F64P]),
            "or_spmv_pull_omp": (ctypes.c_int, [i64, I64P, I32P, I64P, F64P, F64P, ctypes.c_int,
                                                ctypes.c_int, F64P]),
            "or_bp_pull_omp": (ctypes.c_int, [i64, I64P, I32P, I64P, F32P, F32P, F64P, ctypes.c_int,
                                              ctypes.c_int, F64P]),
            "or_bp_f32": (ctypes.c_int, [i64, I64P, I32P, F32P, F32P, F32P, ctypes.c_int, F32P]),
            "or_bp_state_f32": (ctypes.c_int, [i64, I64P, I32P, F32P, F32P, F64P, ctypes.c_int,
                                               F64P]),
            "or_ppr": (ctypes.c_int, [i64, I64P, I32P, ctypes.c_int, f64, F64P, F64P]),
            "or_wpr_forward": (ctypes.c_int, [i64, I64P, I32P, F64P, F64P, f64, ctypes.c_int, f64,
                                              F64P, ctypes.POINTER(ctypes.c_int)]),
            "or_wpr_adjoint": (ctypes.c_int, [i64, I64P, I32P, F64P, F64P, F64P, f64, ctypes.c_int,
                                              f64, F64P, F64P, F64P, ctypes.POINTER(ctypes.c_int)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class Graph:
    """Host CSR (+ lazily built CSC) in the int64-offset / int32-id layout."""

    def __init__(self, n, off, adj):
        self.n = int(n)
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.adj = np.ascontiguousarray(adj, dtype=np.int32)
        self.m = int(self.off[-1])
        self._t = None

    def transpose(self):
        if self._t is None:
            toff = np.zeros(self.n + 1, np.int64)
            tadj = np.zeros(max(self.m, 1), np.int32)
            eperm = np.zeros(max(self.m, 1), np.int64)
            lib().or_transpose(self.n, self.m, self.off, self.adj, toff, tadj,
                               eperm.ctypes.data)
            self._t = (toff, tadj[: self.m], eperm[: self.m])
        return self._t

    def out_degree(self):
        return np.diff(self.off)

    def symmetrized(self):
        L = lib()
        m2 = L.or_symmetrize(self.n, self.off, self.adj, None, None)
        off2 = np.zeros(self.n + 1, np.int64)
        adj2 = np.zeros(max(m2, 1), np.int32)
        L.or_symmetrize(self.n, self.off, self.adj, off2.ctypes.data, adj2.ctypes.data)
        return Graph(self.n, off2, adj2[:m2])


def rmat(scale, m, a=0.57, b=0.19, c=0.19, seed=1):
    """First m distinct non-self-loop rMat pairs of the counter sequence."""
    src = np.zeros(max(m, 1), np.int32)
    dst = np.zeros(max(m, 1), np.int32)
    got = lib().or_rmat_edges(scale, m, a, b, c, seed, src, dst)
    if got != m:
        raise RuntimeError("or_rmat_edges failed: %d" % got)
    n = 1 << scale
    off = np.zeros(n + 1, np.int64)
    adj = np.zeros(max(m, 1), np.int32)
    lib().or_csr_from_sorted(n, m, src, dst, off, adj)
    return Graph(n, off, adj[:m])


def from_edges(n, src, dst):
    """CSR from an edge list (sorted by (src,dst), duplicates kept)."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.int64)
    return Graph(n, off, dst.astype(np.int32))


def weights_i32(g, seed):
    w = np.zeros(max(g.m, 1), np.int32)
    lib().or_fill_weights_i32(seed, g.n, g.off, g.adj, w)
    return w[: g.m]


def values_f32(g, seed):
    w = np.zeros(max(g.m, 1), np.float32)
    lib().or_fill_values_f32(seed, g.n, g.off, g.adj, w)
    return w[: g.m]


def bp_potentials(g, seed, lo=0.1, hi=1.0):
    pot = np.zeros(max(4 * g.m, 1), np.float32)
    lib().or_fill_bp_pot(seed, g.n, g.off, g.adj, lo, hi, pot)
    return pot[: 4 * g.m]


def vertex_f32(n, seed, lo=0.0, hi=1.0):
    x = np.zeros(max(n, 1), np.float32)
    lib().or_fill_vertex_f32(seed, n, lo, hi, x)
    return x[:n]


def write_adj(path, g, w=None):
    if w is None:
        return lib().or_write_adj(path.encode(), g.n, g.m, g.off, g.adj)
    return lib().or_write_wadj(path.encode(), g.n, g.m, g.off, g.adj,
                               np.ascontiguousarray(w, np.int32))


def read_adj(path, skip_self_loops=False):
    L = lib()
    n = ctypes.c_int64()
    m = ctypes.c_int64()
    r = L.or_read_adj(path.encode(), int(skip_self_loops), ctypes.byref(n), ctypes.byref(m),
                      None, None, None)
    if r < 0:
        raise RuntimeError("or_read_adj failed: %d" % r)
    off = np.zeros(n.value + 1, np.int64)
    adj = np.zeros(max(m.value, 1), np.int32)
    w = np.zeros(max(m.value, 1), np.int32)
    L.or_read_adj(path.encode(), int(skip_self_loops), ctypes.byref(n), ctypes.byref(m),
                  off.ctypes.data, adj.ctypes.data, w.ctypes.data)
    return Graph(n.value, off, adj[: m.value]), (w[: m.value] if r == 1 else None)


def write_bin(base, g, idx_bytes=4):
    """or_write_bin: the binary graph triple base.config / .adj / .idx."""
    r = lib().or_write_bin(base.encode(), g.n, g.m, g.off, g.adj, int(idx_bytes))
    if r != 0:
        raise RuntimeError("or_write_bin failed: %d" % r)


def read_bin(base):
    """or_read_bin (readGraphFromBinary, IO-numa.h:543-635)."""
    L = lib()
    n = ctypes.c_int64()
    m = ctypes.c_int64()
    r = L.or_read_bin(base.encode(), ctypes.byref(n), ctypes.byref(m), None, None)
    if r < 0:
        raise RuntimeError("or_read_bin failed: %d" % r)
    off = np.zeros(n.value + 1, np.int64)
    adj = np.zeros(max(m.value, 1), np.int32)
    r = L.or_read_bin(base.encode(), ctypes.byref(n), ctypes.byref(m), off.ctypes.data,
                      adj.ctypes.data)
    if r < 0:
        raise RuntimeError("or_read_bin failed: %d" % r)
    return Graph(n.value, off, adj[: m.value])


def pagerank(g, iters=20, damping=0.85):
    out = np.zeros(g.n, np.float64)
    ept = np.zeros(max(iters, 1), np.int64)
    lib().or_pagerank_push(g.n, g.off, g.adj, iters, damping, out, ept)
    return out, ept[:iters]


def pagerank_delta(g, iters=10, damping=0.85, epsilon2=0.01):
    """or_pagerank_delta: (p, frontier size entering each round)."""
    p = np.zeros(max(g.n, 1), np.float64)
    act = np.zeros(max(iters, 1), np.int64)
    r = lib().or_pagerank_delta(g.n, g.off, g.adj, iters, damping, epsilon2, p, act)
    if r != 0:
        raise RuntimeError("or_pagerank_delta failed: %d" % r)
    return p[: g.n], act[:iters]


def pagerank_pull_omp(g, iters=20, damping=0.85, threads=0):
    toff, tadj, _ = g.transpose()
    out = np.zeros(g.n, np.float64)
    lib().or_pagerank_pull_omp(g.n, g.off, toff, tadj, iters, damping, threads, out)
    return out


def bfs_pull(g, src):
    """Reference-order dense-pull BFS; g must be the graph whose in-lists are
    pulled (for a symmetric graph, itself)."""
    toff, tadj, _ = g.transpose()
    parent = np.zeros(g.n, np.int32)
    level = np.zeros(g.n, np.int32)
    rounds = np.zeros(1, np.int32)
    lib().or_bfs_pull(g.n, toff, tadj, src, parent, level, rounds)
    return parent, level, int(rounds[0])


def bfs_levels(g, src):
    level = np.zeros(g.n, np.int32)
    lib().or_bfs_levels(g.n, g.off, g.adj, src, level)
    return level


def cc_labelprop(g):
    lab = np.zeros(g.n, np.int32)
    rounds = np.zeros(1, np.int32)
    lib().or_cc_labelprop(g.n, g.off, g.adj, lab, rounds)
    return lab, int(rounds[0])


def cc_unionfind(g):
    lab = np.zeros(g.n, np.int32)
    lib().or_cc_unionfind(g.n, g.off, g.adj, lab)
    return lab


def sssp_bf_reference(g, w, src):
    d = np.zeros(g.n, np.int32)
    r = np.zeros(1, np.int32)
    lib().or_sssp_bf_reference(g.n, g.off, g.adj, np.ascontiguousarray(w, np.int32), src, d, r)
    return d, int(r[0])


def sssp_bf(g, w, src):
    d = np.zeros(g.n, np.int32)
    r = np.zeros(1, np.int32)
    lib().or_sssp_bf(g.n, g.off, g.adj, np.ascontiguousarray(w, np.int32), src, d, r)
    return d, int(r[0])


def dijkstra(g, w, src):
    d = np.zeros(g.n, np.int32)
    lib().or_dijkstra(g.n, g.off, g.adj, np.ascontiguousarray(w, np.int32), src, d)
    return d


def spmv(g, w, x0, iters):
    y = np.zeros(g.n, np.float64)
    lib().or_spmv(g.n, g.off, g.adj, np.ascontiguousarray(w, np.float64),
                  np.ascontiguousarray(x0, np.float64), iters, y)
    return y


def spmv_adjoint(g, w, x0, gvec, iters):
    """Gradients of sum_i gvec_i y_T[i] for y_T = spmv(g, w, x0, iters):
    (dL/dx0 per vertex, dL/dw per CSR edge)."""
    gx = np.zeros(g.n, np.float64)
    gw = np.zeros(max(g.m, 1), np.float64)
    lib().or_spmv_adjoint(g.n, g.off, g.adj, np.ascontiguousarray(w, np.float64),
                          np.ascontiguousarray(x0, np.float64),
                          np.ascontiguousarray(gvec, np.float64), iters, gx, gw)
    return gx, gw[: g.m]


def spmv_pull_omp(g, w, x0, iters, threads=0):
    """or_spmv pulled over the CSC with OpenMP threads (bit-identical to spmv)."""
    toff, tadj, teid = g.transpose()
    y = np.zeros(g.n, np.float64)
    lib().or_spmv_pull_omp(g.n, toff, tadj, teid, np.ascontiguousarray(w, np.float64),
                           np.ascontiguousarray(x0, np.float64), iters, threads, y)
    return y


def bp_pull_omp(g, vpot, epot, prod0, iters, threads=0):
    """or_bp pulled over the CSC with OpenMP threads (bit-identical to bp)."""
    toff, tadj, teid = g.transpose()
    out = np.zeros(2 * g.n, np.float64)
    lib().or_bp_pull_omp(g.n, toff, tadj, teid, np.ascontiguousarray(vpot, np.float32),
                         np.ascontiguousarray(epot, np.float32),
                         np.ascontiguousarray(prod0, np.float64).ravel(), iters, threads, out)
    return out.reshape(g.n, 2)


def bp(g, vpot, epot, prod0, iters):
    out = np.zeros(2 * g.n, np.float64)
    lib().or_bp(g.n, g.off, g.adj, np.ascontiguousarray(vpot, np.float32),
                np.ascontiguousarray(epot, np.float32),
                np.ascontiguousarray(prod0, np.float64).ravel(), iters, out)
    return out.reshape(g.n, 2)


def bp_state_f32(g, vpot, epot, prod0, iters):
    """fp64 arithmetic, state rounded to fp32 after every iteration."""
    out = np.zeros(2 * g.n, np.float64)
    lib().or_bp_state_f32(g.n, g.off, g.adj, np.ascontiguousarray(vpot, np.float32),
                          np.ascontiguousarray(epot, np.float32),
                          np.ascontiguousarray(prod0, np.float64).ravel(), iters, out)
    return out.reshape(g.n, 2)


def bp_f32(g, vpot, epot, prod0, iters):
    out = np.zeros(2 * g.n, np.float32)
    lib().or_bp_f32(g.n, g.off, g.adj, np.ascontiguousarray(vpot, np.float32),
                    np.ascontiguousarray(epot, np.float32),
                    np.ascontiguousarray(prod0, np.float32).ravel(), iters, out)
    return out.reshape(g.n, 2)


def ppr(g, tele, iters=20, damping=0.85):
    out = np.zeros(g.n, np.float64)
    lib().or_ppr(g.n, g.off, g.adj, iters, damping, np.ascontiguousarray(tele, np.float64), out)
    return out


def wpr_forward(g, wn, tele, damping=0.85, max_iters=1000, tol=1e-14):
    x = np.zeros(g.n, np.float64)
    it = ctypes.c_int()
    lib().or_wpr_forward(g.n, g.off, g.adj, np.ascontiguousarray(wn, np.float64),
                         np.ascontiguousarray(tele, np.float64), damping, max_iters, tol, x,
                         ctypes.byref(it))
    return x, it.value


def wpr_adjoint(g, wn, x, gvec, damping=0.85, max_iters=1000, tol=1e-14):
    lam = np.zeros(g.n, np.float64)
    gw = np.zeros(max(g.m, 1), np.float64)
    gt = np.zeros(g.n, np.float64)
    it = ctypes.c_int()
    lib().or_wpr_adjoint(g.n, g.off, g.adj, np.ascontiguousarray(wn, np.float64),
                         np.ascontiguousarray(x, np.float64),
                         np.ascontiguousarray(gvec, np.float64), damping, max_iters, tol,
                         lam, gw, gt, ctypes.byref(it))
    return lam, gw[: g.m], gt, it.value
