"""Python binding of libcsm.so, the MI355X-native ClusterSM matching path.

A thin ctypes layer over include/csm.h: the product is the C ABI and its HIP
kernels; this module only moves arguments.  Loading fails loudly when the
library has not been built (``make -C clustersm_amd``): there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcsm.so")

MAX_QV = 32
MAX_QE = 512
MAX_CLUSTERS = 1024
MAX_COMPACT = 4096
NO_PARENT = 0xFFFFFFFF

STATUS = {
    0: "ok", 1: "invalid argument", 2: "cannot open file", 3: "malformed graph file",
    4: "query outside the supported envelope", 5: "planner limits exceeded", 6: "HIP runtime error",
    7: "device arena exhausted", 8: "no gfx950 device", 9: "serving queue full", 10: "unknown ticket",
    11: "result pending", 12: "graph exceeds the uint32 CSR layout", 13: "time limit reached",
    14: "item budget reached",
}
E_IO, E_FORMAT, E_QUERY, E_NOMEM, E_NO_GPU, E_BUSY, E_PENDING, E_LIMIT, E_TIMEOUT, E_BUDGET = 2, 3, 4, 7, 8, 9, 11, 12, 13, 14

K_FILTER, K_REFINE, K_EXPAND, K_MATERIALIZE, K_COUNT, K_AUX, K_BATCH = range(7)
K_NUM = 7
PATH_PLAIN, PATH_HOST, PATH_BATCH, PATH_BATCH_BUDGET, PATH_BATCH_ABORT = range(5)
KERNEL_NAMES = ["filter", "refine", "expand", "materialize", "count", "aux", "batch"]


class QueryPlan(C.Structure):
    _fields_ = [
        ("num_v", C.c_uint32), ("num_e", C.c_uint32),
        ("labels", C.c_uint32 * MAX_QV),
        ("degree", C.c_uint32 * MAX_QV),
        ("edge_src", C.c_uint32 * MAX_QE), ("edge_dst", C.c_uint32 * MAX_QE),
        ("tree_root", C.c_uint32),
        ("tree_degree", C.c_uint32 * MAX_QV),
        ("tree_offsets", C.c_uint32 * (MAX_QV + 1)),
        ("tree_adj", C.c_uint32 * (2 * MAX_QV)),
        ("cover_size", C.c_uint32),
        ("cover", C.c_uint32 * MAX_QV),
        ("num_clusters", C.c_uint32), ("num_cluster_us", C.c_uint32), ("num_layers", C.c_uint32),
        ("cluster_offsets", C.c_uint32 * (MAX_CLUSTERS + 1)),
        ("cluster_us", C.c_uint32 * MAX_COMPACT),
        ("cluster_valid", C.c_uint8 * MAX_CLUSTERS),
        ("root", C.c_uint32),
        ("order_len", C.c_uint32),
        ("order_src", C.c_uint32 * MAX_QE), ("order_dst", C.c_uint32 * MAX_QE),
        ("match_vertex", C.c_uint32 * MAX_QV),
        ("match_parent", C.c_uint32 * MAX_QV),
        ("num_back", C.c_uint32 * MAX_QV),
        ("back_pos", (C.c_uint32 * MAX_QV) * MAX_QV),
    ]

    # convenience views
    @property
    def clusters(self):
        off = list(self.cluster_offsets[: self.num_clusters + 1])
        us = list(self.cluster_us[: self.num_cluster_us])
        return [us[off[i]:off[i + 1]] for i in range(self.num_clusters)]

    @property
    def compact(self):
        return list(self.cluster_us[: self.num_cluster_us])

    @property
    def order(self):
        return [(self.order_src[i], self.order_dst[i]) for i in range(self.order_len)]

    @property
    def match_order(self):
        return list(self.match_vertex[: self.num_v])

    @property
    def parents(self):
        return [None] + [self.match_parent[i] for i in range(1, self.num_v)]

    @property
    def backs(self):
        return [sorted(self.back_pos[i][: self.num_back[i]]) for i in range(self.num_v)]

    @property
    def tree_csr(self):
        n = self.num_v
        off = list(self.tree_offsets[: n + 1])
        return off, list(self.tree_adj[: off[-1]]), list(self.tree_degree[:n])


class KernelStat(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("ns", C.c_uint64), ("bytes", C.c_uint64), ("bytes_ref", C.c_uint64)]


class MatchOpts(C.Structure):
    _fields_ = [("cap", C.c_uint64), ("refine", C.c_uint32), ("reuse_filter", C.c_uint32),
                ("list", C.POINTER(C.c_uint32)), ("list_capacity", C.c_uint64),
                ("chunk_items", C.c_uint64), ("time_limit_ms", C.c_uint32), ("reorder", C.c_uint32),
                ("item_budget", C.c_uint64), ("symmetry", C.c_uint32), ("lds_lists", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("count", C.c_uint64), ("capped", C.c_uint32), ("status", C.c_uint32),
                ("listed", C.c_uint64), ("partial_rows", C.c_uint64), ("items", C.c_uint64),
                ("ms_filter", C.c_double), ("ms_join", C.c_double), ("ms_total", C.c_double),
                ("ms_latency", C.c_double), ("path", C.c_uint32), ("anomalies", C.c_uint32),
                ("kernels", KernelStat * K_NUM), ("automorphisms", C.c_uint64)]

    def kernel(self, cls):
        k = self.kernels[cls]
        return {"launches": k.launches, "ns": k.ns, "bytes": k.bytes, "bytes_ref": k.bytes_ref}


class BatchOpts(C.Structure):
    _fields_ = [("cap", C.c_uint64), ("refine", C.c_uint32), ("use_graph", C.c_uint32),
                ("chunk_rows", C.c_uint32), ("wgs_per_cu", C.c_uint32), ("pool_bytes", C.c_uint64),
                ("block_kb", C.c_uint32), ("max_out", C.c_uint32), ("ilp", C.c_uint32),
                ("time_limit_ms", C.c_uint32), ("item_budget", C.c_uint64), ("reorder", C.c_uint32),
                ("reserved", C.c_uint32)]


class ServeOpts(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("max_inflight", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("cap", C.c_uint64), ("refine", C.c_uint32),
                ("chunk_rows", C.c_uint32), ("wgs_per_cu", C.c_uint32), ("reserved", C.c_uint32)]


class CtxInfo(C.Structure):
    _fields_ = [("device", C.c_int), ("num_cus", C.c_uint32), ("arena_bytes", C.c_uint64),
                ("arena_used", C.c_uint64), ("arena_peak", C.c_uint64), ("arch", C.c_char * 32)]


# Every exported symbol with its ctypes signature (argtypes, restype).
_vp = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
SIGNATURES = {
    "csm_host_graph_read": ([C.c_char_p, C.POINTER(_vp)], C.c_int),
    "csm_host_graph_from_edges": ([C.c_uint32, _u32p, C.c_uint64, _u32p, C.POINTER(_vp)], C.c_int),
    "csm_host_graph_free": ([_vp], None),
    "csm_host_graph_info": ([_vp, _u32p, _u64p], C.c_int),
    "csm_host_graph_views": ([_vp, C.POINTER(_u32p), C.POINTER(_u32p)], C.c_int),
    "csm_host_graph_write": ([_vp, C.c_char_p], C.c_int),
    "csm_host_graph_write_bin": ([_vp, C.c_char_p], C.c_int),
    "csm_host_graph_read_bin": ([C.c_char_p, C.POINTER(_vp)], C.c_int),
    "csm_host_last_error": ([], C.c_char_p),
    "csm_plan_build": ([_vp, C.POINTER(QueryPlan)], C.c_int),
    "csm_open": ([C.c_int, C.c_uint64, C.POINTER(_vp)], C.c_int),
    "csm_close": ([_vp], C.c_int),
    "csm_last_error": ([_vp], C.c_char_p),
    "csm_status_string": ([C.c_int], C.c_char_p),
    "csm_ctx_query": ([_vp, C.POINTER(CtxInfo)], C.c_int),
    "csm_set_profiling": ([_vp, C.c_int], C.c_int),
    "csm_copy_probe": ([_vp, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)], C.c_int),
    "csm_graph_upload": ([_vp, _vp, C.POINTER(_vp)], C.c_int),
    "csm_graph_upload_edges": ([_vp, C.c_uint32, _u32p, C.c_uint64, _u32p, C.POINTER(_vp)], C.c_int),
    "csm_graph_free": ([_vp, _vp], C.c_int),
    "csm_graph_info": ([_vp, _u32p, _u64p, _u32p], C.c_int),
    "csm_graph_download": ([_vp, _vp, _u32p, _u32p, _u32p], C.c_int),
    "csm_graph_to_host": ([_vp, _vp, C.POINTER(_vp)], C.c_int),
    "csm_query_prepare": ([_vp, _vp, C.POINTER(QueryPlan), C.POINTER(_vp)], C.c_int),
    "csm_query_free": ([_vp, _vp], C.c_int),
    "csm_filter": ([_vp, _vp, _vp, C.c_int, _u64p], C.c_int),
    "csm_candidates_download": ([_vp, _vp, _u64p], C.c_int),
    "csm_query_reorder": ([_vp, _vp, _vp, C.c_int], C.c_int),
    "csm_query_symmetry": ([C.POINTER(QueryPlan), _u64p, _u32p, _u32p], C.c_int),
    "csm_match": ([_vp, _vp, _vp, C.POINTER(MatchOpts), C.POINTER(Result)], C.c_int),
    "csm_match_noalloc": ([_vp, _vp, _vp, C.POINTER(MatchOpts), C.POINTER(Result)], C.c_int),
    "csm_batch_match": ([_vp, _vp, C.POINTER(_vp), C.c_uint32, C.POINTER(BatchOpts), C.POINTER(Result)], C.c_int),
    "csm_batch_autotune": ([_vp, _vp, C.POINTER(_vp), C.c_uint32, C.POINTER(BatchOpts), C.c_uint32,
                            C.POINTER(BatchOpts)], C.c_int),
    "csm_serve_start": ([_vp, _vp, C.POINTER(ServeOpts)], C.c_int),
    "csm_serve_submit": ([_vp, C.POINTER(QueryPlan), _u64p], C.c_int),
    "csm_serve_poll": ([_vp, C.c_uint64, C.c_int, C.POINTER(Result)], C.c_int),
    "csm_serve_stop": ([_vp], C.c_int),
    "csm_gen_erdos_renyi": ([C.c_uint32, C.c_uint64, C.c_uint32, C.c_double, C.c_uint64, C.POINTER(_vp)], C.c_int),
    "csm_gen_rmat": ([C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_double,
                      C.c_uint64, C.POINTER(_vp)], C.c_int),
    "csm_gen_chung_lu": ([C.c_uint32, C.c_uint64, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64,
                          C.POINTER(_vp)], C.c_int),
    "csm_gen_query": ([_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, C.POINTER(_vp)], C.c_int),
}

_lib = None


def lib():
    """The loaded libcsm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C clustersm_amd` "
                               "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class CsmError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}{': ' + msg if msg else ''}")


def _host_check(st):
    if st != 0:
        raise CsmError(st, (lib().csm_host_last_error() or b"").decode())


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p)


class HostGraph:
    """Graph in host memory (parsed file, generated, or downloaded)."""

    def __init__(self, handle):
        self.h = _vp(handle) if not isinstance(handle, _vp) else handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.csm_host_graph_free(self.h)
            self.h = None

    @classmethod
    def read(cls, path):
        h = _vp()
        _host_check(lib().csm_host_graph_read(os.fsencode(path), C.byref(h)))
        return cls(h)

    @classmethod
    def read_bin(cls, path):
        h = _vp()
        _host_check(lib().csm_host_graph_read_bin(os.fsencode(path), C.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, labels, edges):
        labels, lp = _u32(labels)
        edges, ep = _u32(np.asarray(edges, dtype=np.uint32).reshape(-1))
        h = _vp()
        _host_check(lib().csm_host_graph_from_edges(len(labels), lp, len(edges) // 2, ep, C.byref(h)))
        return cls(h)

    @property
    def num_v(self):
        n, m = C.c_uint32(), C.c_uint64()
        lib().csm_host_graph_info(self.h, C.byref(n), C.byref(m))
        return n.value

    @property
    def num_e(self):
        n, m = C.c_uint32(), C.c_uint64()
        lib().csm_host_graph_info(self.h, C.byref(n), C.byref(m))
        return m.value

    def arrays(self):
        """(labels, edges[m,2]) as numpy copies."""
        lp, ep = _u32p(), _u32p()
        lib().csm_host_graph_views(self.h, C.byref(lp), C.byref(ep))
        n, m = self.num_v, self.num_e
        labels = np.ctypeslib.as_array(lp, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
        edges = np.ctypeslib.as_array(ep, shape=(2 * m,)).copy().reshape(m, 2) if m else np.zeros((0, 2), np.uint32)
        return labels, edges

    def write(self, path):
        _host_check(lib().csm_host_graph_write(self.h, os.fsencode(path)))

    def write_bin(self, path):
        _host_check(lib().csm_host_graph_write_bin(self.h, os.fsencode(path)))

    def plan(self):
        p = QueryPlan()
        _host_check(lib().csm_plan_build(self.h, C.byref(p)))
        return p


def query_symmetry(plan):
    """(|Aut(Q)|, [(v, w), ...] with f(v) < f(w)) for a QueryPlan (csm_query_symmetry);
    raises CsmError(E_LIMIT) when the automorphism search is too costly."""
    aut, n = C.c_uint64(), C.c_uint32()
    pairs = (C.c_uint32 * (2 * MAX_QV * MAX_QV))()
    _host_check(lib().csm_query_symmetry(C.byref(plan), C.byref(aut), C.byref(n), pairs))
    return aut.value, [(pairs[2 * i], pairs[2 * i + 1]) for i in range(n.value)]


def gen_erdos_renyi(n, m, labels, seed, zipf_s=0.0):
    h = _vp()
    _host_check(lib().csm_gen_erdos_renyi(n, m, labels, zipf_s, seed, C.byref(h)))
    return HostGraph(h)


def gen_rmat(scale, samples, labels, zipf_s, seed, a=0.57, b=0.19, c=0.19):
    h = _vp()
    _host_check(lib().csm_gen_rmat(scale, samples, a, b, c, labels, zipf_s, seed, C.byref(h)))
    return HostGraph(h)


def gen_chung_lu(n, samples, exponent, max_degree, labels, zipf_s, seed):
    h = _vp()
    _host_check(lib().csm_gen_chung_lu(n, samples, exponent, max_degree, labels, zipf_s, seed, C.byref(h)))
    return HostGraph(h)


QUERY_INDUCED, QUERY_TREE, QUERY_EXTRA, QUERY_AVG_DEGREE = 0, 1, 2, 3


def gen_query(data, num_v, mode, seed, extra_edges=0, avg_degree=2.0):
    h = _vp()
    _host_check(lib().csm_gen_query(data.h, num_v, mode, extra_edges, avg_degree, seed, C.byref(h)))
    return HostGraph(h)


class Context:
    """Device + arena (csm_open)."""

    def __init__(self, device=0, arena_bytes=0):
        h = _vp()
        st = lib().csm_open(device, arena_bytes, C.byref(h))
        if st != 0:
            raise CsmError(st, "csm_open")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().csm_close(self.h)
            self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, st, what=""):
        if st != 0:
            raise CsmError(st, f"{what}: {(lib().csm_last_error(self.h) or b'').decode()}")

    def copy_probe(self, nbytes, iters=10):
        """ms per streaming copy of nbytes (csm_copy_probe); 2 * nbytes of HBM traffic each."""
        ms = C.c_double()
        self.check(lib().csm_copy_probe(self.h, nbytes, iters, C.byref(ms)), "csm_copy_probe")
        return ms.value

    def info(self):
        i = CtxInfo()
        self.check(lib().csm_ctx_query(self.h, C.byref(i)), "csm_ctx_query")
        return i

    def set_profiling(self, on=True):
        self.check(lib().csm_set_profiling(self.h, 1 if on else 0))

    def upload(self, host_graph):
        h = _vp()
        self.check(lib().csm_graph_upload(self.h, host_graph.h, C.byref(h)), "csm_graph_upload")
        return DeviceGraph(self, h)

    def upload_edges(self, labels, edges):
        labels, lp = _u32(labels)
        edges, ep = _u32(np.asarray(edges, dtype=np.uint32).reshape(-1))
        h = _vp()
        self.check(lib().csm_graph_upload_edges(self.h, len(labels), lp, len(edges) // 2, ep, C.byref(h)),
                   "csm_graph_upload_edges")
        return DeviceGraph(self, h)

    def serve_start(self, graph, budget_bytes=64 << 30, max_inflight=64, queue_capacity=4096, cap=0,
                    refine=False, chunk_rows=0, wgs_per_cu=0):
        """Continuous batching against `graph` (csm_serve_start); submit and
        poll from any thread, and no other matching call on this context
        until serve_stop."""
        o = ServeOpts(budget_bytes, max_inflight, queue_capacity, cap, 1 if refine else 0, chunk_rows, wgs_per_cu, 0)
        self.check(lib().csm_serve_start(self.h, graph.h, C.byref(o)), "csm_serve_start")

    def serve_submit(self, plan):
        t = C.c_uint64()
        st = lib().csm_serve_submit(self.h, C.byref(plan), C.byref(t))
        self.check(st, "csm_serve_submit")
        return t.value

    def serve_poll(self, ticket, wait=True):
        r = Result()
        st = lib().csm_serve_poll(self.h, ticket, 1 if wait else 0, C.byref(r))
        if st == E_PENDING:
            return None
        self.check(st, "csm_serve_poll")
        return r

    def serve_stop(self):
        self.check(lib().csm_serve_stop(self.h), "csm_serve_stop")


class DeviceGraph:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def free(self):
        if self.h:
            lib().csm_graph_free(self.ctx.h, self.h)
            self.h = None

    def info(self):
        n, m, d = C.c_uint32(), C.c_uint64(), C.c_uint32()
        lib().csm_graph_info(self.h, C.byref(n), C.byref(m), C.byref(d))
        return n.value, m.value, d.value

    def download(self):
        n, m, _ = self.info()
        off = np.zeros(n + 1, np.uint32)
        adj = np.zeros(max(2 * m, 1), np.uint32)
        lab = np.zeros(max(n, 1), np.uint32)
        self.ctx.check(lib().csm_graph_download(self.ctx.h, self.h, off.ctypes.data_as(_u32p),
                                                adj.ctypes.data_as(_u32p), lab.ctypes.data_as(_u32p)))
        return off, adj[: 2 * m], lab[:n]

    def to_host(self):
        h = _vp()
        self.ctx.check(lib().csm_graph_to_host(self.ctx.h, self.h, C.byref(h)), "csm_graph_to_host")
        return HostGraph(h)

    def prepare(self, plan):
        h = _vp()
        self.ctx.check(lib().csm_query_prepare(self.ctx.h, self.h, C.byref(plan), C.byref(h)), "csm_query_prepare")
        return Query(self, plan, h)


class Query:
    def __init__(self, graph, plan, h):
        self.graph, self.plan, self.h = graph, plan, h

    def free(self):
        if self.h:
            lib().csm_query_free(self.graph.ctx.h, self.h)
            self.h = None

    def filter(self, refine=True):
        counts = np.zeros(self.plan.num_v, np.uint64)
        self.graph.ctx.check(lib().csm_filter(self.graph.ctx.h, self.graph.h, self.h, 1 if refine else 0,
                                              counts.ctypes.data_as(_u64p)), "csm_filter")
        return counts

    def candidates(self):
        n = self.graph.info()[0]
        words = (n + 63) // 64
        bits = np.zeros(self.plan.num_v * words, np.uint64)
        self.graph.ctx.check(lib().csm_candidates_download(self.graph.ctx.h, self.h, bits.ctypes.data_as(_u64p)))
        return bits.reshape(self.plan.num_v, words)

    def reorder(self, refine=True):
        """Candidate-aware matching order (csm_query_reorder); counts are unchanged."""
        self.graph.ctx.check(lib().csm_query_reorder(self.graph.ctx.h, self.graph.h, self.h, 1 if refine else 0),
                             "csm_query_reorder")

    def match(self, cap=0, want_list=False, list_capacity=None, refine=True, reuse_filter=False, chunk_items=0,
              time_limit_ms=0, reorder=False, item_budget=0, symmetry=False, lds_lists=False):
        """Returns (Result, rows or None); rows are sorted, in query-id order.
        With time_limit_ms (item_budget), Result.status == E_TIMEOUT (E_BUDGET)
        marks a stopped search whose count is a lower bound.
        reorder: candidate-aware matching order after the filter (same counts).
        lds_lists: the count level searches LDS copies of its rows' short
        candidate lists (same counts)."""
        o = MatchOpts()
        o.cap, o.refine, o.reuse_filter, o.chunk_items = cap, 1 if refine else 0, 1 if reuse_filter else 0, chunk_items
        o.time_limit_ms = time_limit_ms
        o.item_budget = item_budget
        o.symmetry = 1 if symmetry else 0
        o.reorder = 1 if reorder else 0
        o.lds_lists = 1 if lds_lists else 0
        rows = None
        if want_list:
            lc = list_capacity if list_capacity is not None else (cap if cap else 1 << 20)
            rows = np.zeros((max(lc, 1), self.plan.num_v), np.uint32)
            o.list = rows.ctypes.data_as(_u32p)
            o.list_capacity = lc
        r = Result()
        self.graph.ctx.check(lib().csm_match(self.graph.ctx.h, self.graph.h, self.h, C.byref(o), C.byref(r)),
                             "csm_match")
        if want_list:
            rows = rows[: r.listed]
        return r, rows


    def match_noalloc(self, cap=0, refine=True, reuse_filter=False, time_limit_ms=0, reorder=False):
        """csm_match_noalloc: count only, every workspace buffer in the library's
        static device scratch (no allocation).  Returns the Result."""
        o = MatchOpts()
        o.cap, o.refine, o.reuse_filter = cap, 1 if refine else 0, 1 if reuse_filter else 0
        o.time_limit_ms = time_limit_ms
        o.reorder = 1 if reorder else 0
        r = Result()
        self.graph.ctx.check(lib().csm_match_noalloc(self.graph.ctx.h, self.graph.h, self.h, C.byref(o),
                                                     C.byref(r)), "csm_match_noalloc")
        return r


def batch_autotune(graph, queries, cap=0, refine=False, sample=32, time_limit_ms=0, item_budget=0):
    """Tuned BatchOpts for this graph and query mix (csm_batch_autotune)."""
    arr = (_vp * len(queries))(*[q.h for q in queries])
    base = BatchOpts(cap, 1 if refine else 0, 1, 0, 0, 0, 0, 0, 0, time_limit_ms, item_budget)
    out = BatchOpts()
    graph.ctx.check(lib().csm_batch_autotune(graph.ctx.h, graph.h, arr, len(queries), C.byref(base), sample,
                                             C.byref(out)), "csm_batch_autotune")
    return out


def batch_match(graph, queries, cap=0, refine=False, use_graph=True, chunk_rows=0, wgs_per_cu=0, pool_bytes=0,
                block_kb=0, max_out=0, ilp=0, time_limit_ms=0, item_budget=0, opts=None, reorder=False):
    """All queries in one persistent kernel (csm_batch_match); counts only.
    opts: a BatchOpts (e.g. from batch_autotune) used as given, except that
    reorder=True sets its reorder field."""
    arr = (_vp * len(queries))(*[q.h for q in queries])
    o = opts if opts is not None else BatchOpts(cap, 1 if refine else 0, 1 if use_graph else 0, chunk_rows,
                                                wgs_per_cu, pool_bytes, block_kb, max_out, ilp, time_limit_ms,
                                                item_budget)
    if reorder:
        o = BatchOpts.from_buffer_copy(o)
        o.reorder = 1
    res = (Result * len(queries))()
    graph.ctx.check(lib().csm_batch_match(graph.ctx.h, graph.h, arr, len(queries), C.byref(o), res),
                    "csm_batch_match")
    return list(res)
