"""polymer_amd -- Python view of libpolymer_gpu.so (MI355X edgeMap/vertexMap engine).

A thin ctypes layer over the C ABI in include/polymer_gpu.h.  Everything
computes on the GPU through the HIP library; there is no CPU fallback: if the
library is missing or no GPU is present the calls raise.

    import polymer_amd as pa
    pa.init(0)
    g = pa.Graph.rmat(16, 1 << 20, seed=1)          # device-generated rMat
    rank, st = g.pagerank(20)                          # numa-PageRank.C
    level, parent, st = g.bfs(0)                       # numa-BFS.C
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpolymer_gpu.so")

PG_OK = 0
PG_GRAPH_SYMMETRIC = 0x1
PG_GRAPH_SKIP_SELF_LOOPS = 0x2
PG_GRAPH_SYMMETRIZE = 0x4
PG_GRAPH_WEIGHT_I32 = 0x8
PG_GRAPH_VALUE_F32 = 0x10
PG_W_NONE, PG_W_I32, PG_W_F32 = 0, 1, 2
PG_PROFILE_KERNELS = 0x1
PG_BP_FP32 = 0x2
PG_SERIAL_COLD = 0x4
PG_SPMV_SAVE = 0x8
INF_DIST = 1073741823

# every symbol include/polymer_gpu.h declares
EXPORTS = [
    "pg_init", "pg_finalize", "pg_last_error", "pg_device_name",
    "pg_graph_create", "pg_graph_rmat", "pg_graph_load_adj", "pg_graph_load_bin",
    "pg_graph_load_dump", "pg_graph_dump", "pg_graph_destroy",
    "pg_graph_info", "pg_graph_download", "pg_graph_set_schedule", "pg_graph_set_tiles",
    "pg_graph_get_schedule", "pg_graph_tile_stats", "pg_graph_autotune",
    "pg_graph_bp_potentials", "pg_graph_view",
    "pg_subset_create", "pg_subset_destroy", "pg_subset_clear", "pg_subset_fill",
    "pg_subset_from_dense", "pg_subset_from_sparse", "pg_subset_to_dense",
    "pg_subset_to_sparse", "pg_subset_to_sparse_noalloc", "pg_subset_size",
    "pg_pagerank", "pg_pagerank_delta", "pg_spmv", "pg_bp", "pg_bfs", "pg_bfs_async", "pg_cc",
    "pg_sssp",
    "pg_msbfs", "pg_mssssp", "pg_ppr_batched", "pg_ppr",
    "pg_graph_multi", "pg_pagerank_multigraph",
    "pg_wpr_forward", "pg_wpr_adjoint", "pg_spmv_adjoint",
]


class Stats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int32),
        ("dense_rounds", ctypes.c_int32),
        ("edges", ctypes.c_int64),
        ("active", ctypes.c_int64),
        ("ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("kernel_launches", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("residual", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class PolymerError(RuntimeError):
    pass


_lib = None


def load():
    """Loads the HIP library (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PolymerError("%s is not built: run __graft_entry__.build() "
                           "(make -C polymer_amd/csrc)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i64, i32, u32, u64, f64 = (ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32,
                               ctypes.c_uint64, ctypes.c_double)
    PP = ctypes.POINTER(ctypes.c_void_p)
    SP = ctypes.POINTER(Stats)
    sig = {
        "pg_init": [ctypes.c_int],
        "pg_finalize": [],
        "pg_last_error": [ctypes.c_char_p, ctypes.c_size_t],
        "pg_device_name": [ctypes.c_char_p, ctypes.c_size_t],
        "pg_graph_create": [i64, i64, P, P, P, ctypes.c_int, u32, PP],
        "pg_graph_rmat": [ctypes.c_int, i64, f64, f64, f64, u64, u32, u64, PP],
        "pg_graph_load_adj": [ctypes.c_char_p, u32, PP],
        "pg_graph_load_bin": [ctypes.c_char_p, u32, PP],
        "pg_graph_load_dump": [ctypes.c_char_p, u32, PP],
        "pg_graph_dump": [P, ctypes.c_char_p],
        "pg_graph_destroy": [P],
        "pg_graph_info": [P, P, P, P, P],
        "pg_graph_download": [P, P, P, P],
        "pg_graph_set_schedule": [P, ctypes.c_int, i64, i64, i64],
        "pg_graph_set_tiles": [P, ctypes.c_int, i64, i64, i64],
        "pg_graph_get_schedule": [P, P, P, P, P, P, P, P, P],
        "pg_graph_tile_stats": [P, P, P, P, P, P, P],
        "pg_graph_autotune": [P, P, P],
        "pg_graph_bp_potentials": [P, u64, ctypes.c_float, ctypes.c_float],
        "pg_subset_create": [P, PP],
        "pg_subset_destroy": [P],
        "pg_subset_clear": [P],
        "pg_subset_fill": [P],
        "pg_subset_from_dense": [P, P],
        "pg_subset_from_sparse": [P, P, i64],
        "pg_subset_to_dense": [P, P],
        "pg_subset_to_sparse": [P, P, i64, P],
        "pg_subset_to_sparse_noalloc": [P, P, P, P],
        "pg_subset_size": [P, P, P],
        "pg_pagerank": [P, ctypes.c_int, f64, P, u32, SP],
        "pg_pagerank_delta": [P, ctypes.c_int, f64, f64, P, u32, SP],
        "pg_spmv": [P, ctypes.c_int, P, P, u32, SP],
        "pg_spmv_adjoint": [P, P, P, P, u32, SP],
        "pg_bp": [P, ctypes.c_int, P, P, P, P, u32, SP],
        "pg_bfs": [P, i64, i64, P, P, u32, SP],
        "pg_bfs_async": [P, i64, P, P, u32, SP],
        "pg_cc": [P, i64, P, u32, SP],
        "pg_sssp": [P, i64, i64, P, u32, SP],
        "pg_msbfs": [P, P, ctypes.c_int, P, u32, SP],
        "pg_mssssp": [P, P, ctypes.c_int, P, u32, SP],
        "pg_ppr_batched": [P, P, ctypes.c_int, ctypes.c_int, f64, P, u32, SP],
        "pg_ppr": [P, i64, ctypes.c_int, f64, P, u32, SP],
        "pg_graph_multi": [ctypes.c_int, P, i64, P, P, PP],
        "pg_pagerank_multigraph": [P, ctypes.c_int, f64, P, u32, SP],
        "pg_wpr_forward": [P, P, P, f64, ctypes.c_int, f64, P, u32, SP],
        "pg_wpr_adjoint": [P, P, ctypes.c_int, f64, P, P, P, u32, SP],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = L
    return L


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().pg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(rc, what):
    if rc != PG_OK:
        raise PolymerError("%s failed (%d): %s" % (what, rc, last_error()))


def init(device=0):
    _check(load().pg_init(device), "pg_init")


def finalize():
    _check(load().pg_finalize(), "pg_finalize")


def device_name():
    buf = ctypes.create_string_buffer(256)
    _check(load().pg_device_name(buf, 256), "pg_device_name")
    return buf.value.decode()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _flags(profile):
    return PG_PROFILE_KERNELS if profile else 0


class Graph:
    """A device-resident graph (CSR + CSC, relabelled by in-degree)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle
        n, m, mx, wd = (ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32())
        _check(load().pg_graph_info(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(mx),
                                    ctypes.byref(wd)), "pg_graph_info")
        self.n, self.m, self.max_in_degree, self.w_dtype = n.value, m.value, mx.value, wd.value

    # ---- construction -----------------------------------------------------------
    @classmethod
    def from_csr(cls, n, off, adj, w=None, symmetric=False, skip_self_loops=False):
        off = np.ascontiguousarray(off, np.int64)
        adj = np.ascontiguousarray(adj, np.int32)
        flags = (PG_GRAPH_SYMMETRIC if symmetric else 0) | (
            PG_GRAPH_SKIP_SELF_LOOPS if skip_self_loops else 0)
        wd = PG_W_NONE
        if w is not None:
            w = np.ascontiguousarray(w)
            if w.dtype == np.int32:
                wd = PG_W_I32
            else:
                w = np.ascontiguousarray(w, np.float32)
                wd = PG_W_F32
        h = ctypes.c_void_p()
        _check(load().pg_graph_create(n, int(off[-1]), _ptr(off), _ptr(adj), _ptr(w), wd, flags,
                                      ctypes.byref(h)), "pg_graph_create")
        return cls(h)

    @classmethod
    def rmat(cls, scale, m, a=0.57, b=0.19, c=0.19, seed=1, symmetrize=False, weights=None,
             wseed=7):
        flags = PG_GRAPH_SYMMETRIZE if symmetrize else 0
        if weights == "i32":
            flags |= PG_GRAPH_WEIGHT_I32
        elif weights == "f32":
            flags |= PG_GRAPH_VALUE_F32
        h = ctypes.c_void_p()
        _check(load().pg_graph_rmat(scale, m, a, b, c, seed, flags, wseed, ctypes.byref(h)),
               "pg_graph_rmat")
        return cls(h)

    @classmethod
    def load_adj(cls, path, symmetric=False, skip_self_loops=False):
        flags = (PG_GRAPH_SYMMETRIC if symmetric else 0) | (
            PG_GRAPH_SKIP_SELF_LOOPS if skip_self_loops else 0)
        h = ctypes.c_void_p()
        _check(load().pg_graph_load_adj(path.encode(), flags, ctypes.byref(h)),
               "pg_graph_load_adj")
        return cls(h)

    @classmethod
    def load_bin(cls, base, symmetric=False, weighted=False, skip_self_loops=False):
        """The binary graph base.config / base.adj / base.idx (unit weights when
        weighted, as the reference's binary reader gives)."""
        flags = (PG_GRAPH_SYMMETRIC if symmetric else 0) | (
            PG_GRAPH_SKIP_SELF_LOOPS if skip_self_loops else 0) | (
            PG_GRAPH_WEIGHT_I32 if weighted else 0)
        h = ctypes.c_void_p()
        _check(load().pg_graph_load_bin(base.encode(), flags, ctypes.byref(h)),
               "pg_graph_load_bin")
        return cls(h)

    @classmethod
    def load_dump(cls, path, symmetric=False):
        """The whole-graph dump of loadGraphFromBin (IO-numa.h:758-835)."""
        h = ctypes.c_void_p()
        _check(load().pg_graph_load_dump(path.encode(), PG_GRAPH_SYMMETRIC if symmetric else 0,
                                         ctypes.byref(h)), "pg_graph_load_dump")
        return cls(h)

    def dump(self, path):
        """Writes the graph as dumpGraphToBin does (IO-numa.h:837-900)."""
        _check(load().pg_graph_dump(self._h, path.encode()), "pg_graph_dump")

    @classmethod
    def multi(cls, vbase, off, adj):
        vbase = np.ascontiguousarray(vbase, np.int64)
        off = np.ascontiguousarray(off, np.int64)
        adj = np.ascontiguousarray(adj, np.int32)
        h = ctypes.c_void_p()
        _check(load().pg_graph_multi(len(vbase) - 1, _ptr(vbase), int(off[-1]), _ptr(off),
                                     _ptr(adj), ctypes.byref(h)), "pg_graph_multi")
        return cls(h)

    def close(self):
        if self._h:
            load().pg_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- structure ----------------------------------------------------------------
    def download(self):
        off = np.zeros(self.n + 1, np.int64)
        adj = np.zeros(max(self.m, 1), np.int32)
        w = None
        if self.w_dtype == PG_W_I32:
            w = np.zeros(max(self.m, 1), np.int32)
        elif self.w_dtype == PG_W_F32:
            w = np.zeros(max(self.m, 1), np.float32)
        _check(load().pg_graph_download(self._h, _ptr(off), _ptr(adj), _ptr(w)),
               "pg_graph_download")
        return off, adj[: self.m], (None if w is None else w[: self.m])

    def set_schedule(self, items_per_lane=0, block_threshold=0, multi_threshold=0, chunk=0):
        _check(load().pg_graph_set_schedule(self._h, items_per_lane, block_threshold,
                                            multi_threshold, chunk), "pg_graph_set_schedule")

    def set_tiles(self, enable=-1, tile_edges=0, cold_from=0, max_degree=0):
        """Source-sorted tiles for the PageRank pull (on by default); max_degree: the
        largest in-degree put in a tile; cold_from: first source id served by the cold
        pre-pass (-1 none).  0 keeps a setting."""
        _check(load().pg_graph_set_tiles(self._h, int(enable), int(max_degree), int(tile_edges),
                                         int(cold_from)), "pg_graph_set_tiles")

    def get_schedule(self):
        """(items_per_lane, block_threshold, multi_threshold, chunk, tiles, max_degree,
        tile_edges, cold_from)"""
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        v = [i32(), i64(), i64(), i64(), i32(), i64(), i64(), i64()]
        _check(load().pg_graph_get_schedule(self._h, *[ctypes.byref(x) for x in v]),
               "pg_graph_get_schedule")
        return tuple(x.value for x in v)

    def tile_stats(self):
        """dict: tiles, parts, tiled_edges, cold_edges, escape_groups, cold_overlap"""
        v = [ctypes.c_int64() for _ in range(5)] + [ctypes.c_int32()]
        _check(load().pg_graph_tile_stats(self._h, *[ctypes.byref(x) for x in v]),
               "pg_graph_tile_stats")
        keys = ("tiles", "parts", "tiled_edges", "cold_edges", "escape_groups", "cold_overlap")
        return dict(zip(keys, (x.value for x in v)))

    def autotune(self):
        ipl, bt = ctypes.c_int32(), ctypes.c_int64()
        _check(load().pg_graph_autotune(self._h, ctypes.byref(ipl), ctypes.byref(bt)),
               "pg_graph_autotune")
        return ipl.value, bt.value

    def bp_potentials(self, seed, lo=0.1, hi=1.0):
        _check(load().pg_graph_bp_potentials(self._h, seed, lo, hi), "pg_graph_bp_potentials")

    def subset(self):
        return Subset(self)

    # ---- applications -------------------------------------------------------------
    def pagerank(self, iters=20, damping=0.85, out=True, profile=False, serial_cold=False):
        rank = np.zeros(self.n, np.float32) if out else None
        st = Stats()
        flags = _flags(profile) | (PG_SERIAL_COLD if serial_cold else 0)
        _check(load().pg_pagerank(self._h, iters, damping, _ptr(rank), flags,
                                  ctypes.byref(st)), "pg_pagerank")
        return rank, st

    def pagerank_delta(self, iters=10, damping=0.85, epsilon2=0.01, out=True, profile=False):
        """numa-PageRankDelta.C: fp64 ranks (host n) and stats."""
        p = np.zeros(self.n, np.float64) if out else None
        st = Stats()
        _check(load().pg_pagerank_delta(self._h, iters, damping, epsilon2, _ptr(p),
                                        _flags(profile), ctypes.byref(st)), "pg_pagerank_delta")
        return p, st

    def spmv(self, x0, iters=20, out=True, profile=False, serial_cold=False, save=False):
        """numa-SPMV.C: fp64 y_T (host n) and stats; save keeps the iterates for
        spmv_adjoint."""
        x0 = np.ascontiguousarray(x0, np.float32)
        y = np.zeros(self.n, np.float64) if out else None
        st = Stats()
        flags = (_flags(profile) | (PG_SERIAL_COLD if serial_cold else 0) |
                 (PG_SPMV_SAVE if save else 0))
        _check(load().pg_spmv(self._h, iters, _ptr(x0), _ptr(y), flags,
                              ctypes.byref(st)), "pg_spmv")
        return y, st

    def spmv_adjoint(self, gvec, out=True, profile=False):
        """Gradients of sum_i gvec_i y_T[i] after spmv(..., save=True): (dL/dx0 host n
        fp64, dL/dw host m fp64 in input CSR order, stats)."""
        gvec = np.ascontiguousarray(gvec, np.float32)
        gx = np.zeros(self.n, np.float64) if out else None
        gw = np.zeros(self.m, np.float64) if out else None
        st = Stats()
        _check(load().pg_spmv_adjoint(self._h, _ptr(gvec), _ptr(gx), _ptr(gw), _flags(profile),
                                      ctypes.byref(st)), "pg_spmv_adjoint")
        return gx, gw, st

    def bp(self, vpot, prod0, iters=10, epot=None, out=True, profile=False, fp32=False):
        vpot = np.ascontiguousarray(vpot, np.float32).ravel()
        prod0 = np.ascontiguousarray(prod0, np.float32).ravel()
        if epot is not None:
            epot = np.ascontiguousarray(epot, np.float32).ravel()
        prod = np.zeros(2 * self.n, np.float64) if out else None
        st = Stats()
        flags = _flags(profile) | (PG_BP_FP32 if fp32 else 0)
        _check(load().pg_bp(self._h, iters, _ptr(vpot), _ptr(epot), _ptr(prod0), _ptr(prod),
                            flags, ctypes.byref(st)), "pg_bp")
        return (None if prod is None else prod.reshape(self.n, 2)), st

    def bfs(self, src, threshold=-1, out=True, profile=False):
        level = np.zeros(self.n, np.int32) if out else None
        parent = np.zeros(self.n, np.int32) if out else None
        st = Stats()
        _check(load().pg_bfs(self._h, src, threshold, _ptr(level), _ptr(parent),
                             _flags(profile), ctypes.byref(st)), "pg_bfs")
        return level, parent, st

    def bfs_async(self, src, out=True, profile=False):
        """numa-BFS-async-pipe.C: (parents of a spanning tree of the reachable set,
        reached count, stats)."""
        parent = np.zeros(self.n, np.int32) if out else None
        vis = ctypes.c_int64()
        st = Stats()
        _check(load().pg_bfs_async(self._h, src, _ptr(parent), ctypes.byref(vis), _flags(profile),
                                   ctypes.byref(st)), "pg_bfs_async")
        return parent, vis.value, st

    def cc(self, threshold=-1, out=True, profile=False):
        label = np.zeros(self.n, np.int32) if out else None
        st = Stats()
        _check(load().pg_cc(self._h, threshold, _ptr(label), _flags(profile), ctypes.byref(st)),
               "pg_cc")
        return label, st

    def sssp(self, src, threshold=-1, out=True, profile=False):
        dist = np.zeros(self.n, np.int32) if out else None
        st = Stats()
        _check(load().pg_sssp(self._h, src, threshold, _ptr(dist), _flags(profile),
                              ctypes.byref(st)), "pg_sssp")
        return dist, st

    def msbfs(self, sources, out=True, profile=False):
        src = np.ascontiguousarray(sources, np.int64)
        lev = np.zeros((len(src), self.n), np.int32) if out else None
        st = Stats()
        _check(load().pg_msbfs(self._h, _ptr(src), len(src), _ptr(lev), _flags(profile),
                               ctypes.byref(st)), "pg_msbfs")
        return lev, st

    def mssssp(self, sources, out=True, profile=False):
        src = np.ascontiguousarray(sources, np.int64)
        dist = np.zeros((len(src), self.n), np.int32) if out else None
        st = Stats()
        _check(load().pg_mssssp(self._h, _ptr(src), len(src), _ptr(dist), _flags(profile),
                                ctypes.byref(st)), "pg_mssssp")
        return dist, st

    def ppr_batched(self, sources, iters=20, damping=0.85, out=True, profile=False):
        src = np.ascontiguousarray(sources, np.int64)
        rank = np.zeros((len(src), self.n), np.float32) if out else None
        st = Stats()
        _check(load().pg_ppr_batched(self._h, _ptr(src), len(src), iters, damping, _ptr(rank),
                                     _flags(profile), ctypes.byref(st)), "pg_ppr_batched")
        return rank, st

    def ppr(self, source, iters=20, damping=0.85, out=True, profile=False):
        rank = np.zeros(self.n, np.float32) if out else None
        st = Stats()
        _check(load().pg_ppr(self._h, int(source), iters, damping, _ptr(rank), _flags(profile),
                             ctypes.byref(st)), "pg_ppr")
        return rank, st

    def pagerank_multigraph(self, iters=20, damping=0.85, out=True, profile=False):
        rank = np.zeros(self.n, np.float32) if out else None
        st = Stats()
        _check(load().pg_pagerank_multigraph(self._h, iters, damping, _ptr(rank),
                                             _flags(profile), ctypes.byref(st)),
               "pg_pagerank_multigraph")
        return rank, st

    def wpr_forward(self, wn, tele, damping=0.85, max_iters=500, tol=1e-7, out=True,
                    profile=False):
        wn = np.ascontiguousarray(wn, np.float32)
        tele = np.ascontiguousarray(tele, np.float32)
        x = np.zeros(self.n, np.float32) if out else None
        st = Stats()
        _check(load().pg_wpr_forward(self._h, _ptr(wn), _ptr(tele), damping, max_iters, tol,
                                     _ptr(x), _flags(profile), ctypes.byref(st)),
               "pg_wpr_forward")
        return x, st

    def wpr_adjoint(self, gvec, max_iters=500, tol=1e-7, out=True, profile=False):
        gvec = np.ascontiguousarray(gvec, np.float32)
        lam = np.zeros(self.n, np.float32) if out else None
        gw = np.zeros(max(self.m, 1), np.float32) if out else None
        gt = np.zeros(self.n, np.float32) if out else None
        st = Stats()
        _check(load().pg_wpr_adjoint(self._h, _ptr(gvec), max_iters, tol, _ptr(lam), _ptr(gw),
                                     _ptr(gt), _flags(profile), ctypes.byref(st)),
               "pg_wpr_adjoint")
        return lam, (None if gw is None else gw[: self.m]), gt, st


class Subset:
    """vertexSubset on the device: bitmap + ordered id list + counters."""

    def __init__(self, g):
        self.g = g
        self._h = ctypes.c_void_p()
        _check(load().pg_subset_create(g._h, ctypes.byref(self._h)), "pg_subset_create")

    def close(self):
        if self._h:
            load().pg_subset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        _check(load().pg_subset_clear(self._h), "pg_subset_clear")

    def fill(self):
        _check(load().pg_subset_fill(self._h), "pg_subset_fill")

    def from_dense(self, flags):
        flags = np.ascontiguousarray(flags, np.uint8)
        _check(load().pg_subset_from_dense(self._h, _ptr(flags)), "pg_subset_from_dense")

    def from_sparse(self, ids):
        ids = np.ascontiguousarray(ids, np.int32)
        _check(load().pg_subset_from_sparse(self._h, _ptr(ids), len(ids)),
               "pg_subset_from_sparse")

    def to_dense(self):
        out = np.zeros(self.g.n, np.uint8)
        _check(load().pg_subset_to_dense(self._h, _ptr(out)), "pg_subset_to_dense")
        return out.astype(bool)

    def to_sparse(self):
        k = ctypes.c_int64()
        cnt, _ = self.size()
        out = np.zeros(max(cnt, 1), np.int32)
        _check(load().pg_subset_to_sparse(self._h, _ptr(out), len(out), ctypes.byref(k)),
               "pg_subset_to_sparse")
        return out[: k.value]

    def size(self):
        c, e = ctypes.c_int64(), ctypes.c_int64()
        _check(load().pg_subset_size(self._h, ctypes.byref(c), ctypes.byref(e)),
               "pg_subset_size")
        return c.value, e.value
