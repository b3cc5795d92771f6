"""ctypes wrapper of the CPU oracle (checker) in csm_oracle.c.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DECYCLE_PATH = os.path.join(HERE, "_ref", "libref_decycle.so")

MAX_QV = 32
MAX_QE = 512
MAX_CLUSTERS = 4096
MAX_COMPACT = 16384


class OcQuery(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("m", C.c_uint32),
        ("label", C.c_uint32 * MAX_QV),
        ("deg", C.c_uint32 * MAX_QV),
        ("off", C.c_uint32 * (MAX_QV + 1)),
        ("adj", C.c_uint32 * (2 * MAX_QE)),
        ("eid", C.c_uint32 * (2 * MAX_QE)),
        ("esrc", C.c_uint32 * MAX_QE), ("edst", C.c_uint32 * MAX_QE),
    ]


class OcPlan(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("m", C.c_uint32),
        ("tree_root", C.c_uint32),
        ("tree_deg", C.c_uint32 * MAX_QV),
        ("tree_off", C.c_uint32 * (MAX_QV + 1)),
        ("tree_adj", C.c_uint32 * (2 * MAX_QV)),
        ("n_dropped", C.c_uint32),
        ("dropped_src", C.c_uint32 * MAX_QE), ("dropped_dst", C.c_uint32 * MAX_QE),
        ("n_cover", C.c_uint32),
        ("cover", C.c_uint32 * MAX_QV),
        ("n_clusters", C.c_uint32), ("n_compact", C.c_uint32), ("n_layers", C.c_uint32),
        ("cluster_off", C.c_uint32 * (MAX_CLUSTERS + 1)),
        ("cluster_valid", C.c_uint8 * MAX_CLUSTERS),
        ("compact", C.c_uint32 * MAX_COMPACT),
        ("root", C.c_uint32),
        ("order_len", C.c_uint32),
        ("order_src", C.c_uint32 * MAX_QE), ("order_dst", C.c_uint32 * MAX_QE),
        ("mvert", C.c_uint32 * MAX_QV),
        ("mpos", C.c_uint32 * MAX_QV),
        ("parent", C.c_uint32 * MAX_QV),
        ("nback", C.c_uint32 * MAX_QV),
        ("back", (C.c_uint32 * MAX_QV) * MAX_QV),
    ]


class OcResult(C.Structure):
    _fields_ = [("count", C.c_uint64), ("capped", C.c_uint32), ("timed_out", C.c_uint32),
                ("listed", C.c_uint64), ("seconds", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        u32p = C.POINTER(C.c_uint32)
        L.oc_query_read.argtypes = [C.c_char_p, C.POINTER(OcQuery)]
        L.oc_query_from_arrays.argtypes = [C.c_uint32, u32p, C.c_uint32, u32p, C.POINTER(OcQuery)]
        L.oc_data_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.oc_data_from_arrays.argtypes = [C.c_uint32, u32p, C.c_uint64, u32p, C.POINTER(C.c_void_p)]
        L.oc_data_free.argtypes = [C.c_void_p]
        L.oc_data_num_edges.argtypes = [C.c_void_p]
        L.oc_data_num_edges.restype = C.c_uint64
        L.oc_plan_build.argtypes = [C.POINTER(OcQuery), C.POINTER(OcPlan)]
        L.oc_filter_nlc.argtypes = [C.c_void_p, C.POINTER(OcQuery), C.POINTER(C.c_uint64)]
        L.oc_refine.argtypes = [C.c_void_p, C.POINTER(OcPlan), C.POINTER(C.c_uint64)]
        L.oc_match.argtypes = [C.c_void_p, C.POINTER(OcQuery), C.POINTER(OcPlan), C.c_uint64,
                               u32p, C.c_uint64, C.c_int, C.c_double, C.c_int, C.POINTER(OcResult)]
        _lib = L
    return _lib


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


class OracleError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} failed with status {rc}")
        self.rc = rc


def _check(fn, rc):
    if rc != 0:
        raise OracleError(fn, rc)


class Query:
    def __init__(self, q):
        self.q = q

    @classmethod
    def read(cls, path):
        q = OcQuery()
        _check("oc_query_read", lib().oc_query_read(path.encode(), C.byref(q)))
        return cls(q)

    @classmethod
    def from_arrays(cls, labels, edges):
        labels, lp = _u32(labels)
        edges, ep = _u32(np.asarray(edges, dtype=np.uint32).reshape(-1))
        q = OcQuery()
        _check("oc_query_from_arrays",
               lib().oc_query_from_arrays(len(labels), lp, len(edges) // 2, ep, C.byref(q)))
        return cls(q)

    @property
    def n(self):
        return self.q.n

    def plan(self):
        p = OcPlan()
        _check("oc_plan_build", lib().oc_plan_build(C.byref(self.q), C.byref(p)))
        return Plan(p)


class Plan:
    def __init__(self, p):
        self.p = p

    @property
    def root(self):
        return self.p.root

    @property
    def compact(self):
        return list(self.p.compact[: self.p.n_compact])

    @property
    def cluster_offsets(self):
        return list(self.p.cluster_off[: self.p.n_clusters + 1])

    @property
    def clusters(self):
        off = self.cluster_offsets
        c = self.compact
        return [c[off[i]:off[i + 1]] for i in range(self.p.n_clusters)]

    @property
    def cover(self):
        return list(self.p.cover[: self.p.n_cover])

    @property
    def order(self):
        return [(self.p.order_src[i], self.p.order_dst[i]) for i in range(self.p.order_len)]

    @property
    def match_order(self):
        return list(self.p.mvert[: self.p.n])

    @property
    def parents(self):
        return [None] + [self.p.parent[i] for i in range(1, self.p.n)]

    @property
    def backs(self):
        return [sorted(self.p.back[i][: self.p.nback[i]]) for i in range(self.p.n)]

    @property
    def tree_csr(self):
        n = self.p.n
        off = list(self.p.tree_off[: n + 1])
        return off, list(self.p.tree_adj[: off[-1]]), list(self.p.tree_deg[:n])


class Data:
    def __init__(self, h):
        self.h = h

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        _check("oc_data_read", lib().oc_data_read(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, labels, edges):
        labels, lp = _u32(labels)
        edges, ep = _u32(np.asarray(edges, dtype=np.uint32).reshape(-1))
        h = C.c_void_p()
        _check("oc_data_from_arrays",
               lib().oc_data_from_arrays(len(labels), lp, len(edges) // 2, ep, C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oc_data_free(self.h)
            self.h = None


def filter_nlc(data, nv, query):
    words = (nv + 63) // 64
    bits = np.zeros(query.n * words, dtype=np.uint64)
    _check("oc_filter_nlc", lib().oc_filter_nlc(data.h, C.byref(query.q),
                                               bits.ctypes.data_as(C.POINTER(C.c_uint64))))
    return bits.reshape(query.n, words)


def refine(data, nv, query, plan):
    bits = filter_nlc(data, nv, query)
    flat = np.ascontiguousarray(bits.reshape(-1))
    _check("oc_refine", lib().oc_refine(data.h, C.byref(plan.p),
                                        flat.ctypes.data_as(C.POINTER(C.c_uint64))))
    return flat.reshape(query.n, -1)


def bits_to_sets(bits, nv):
    out = []
    for row in bits:
        b = np.unpackbits(row.view(np.uint8), bitorder="little")[:nv]
        out.append(np.nonzero(b)[0].astype(np.uint32))
    return out


def match(data, query, plan, cap=(1 << 64) - 1, want_list=False, list_cap=None,
          threads=1, budget_s=0.0, refine=True):
    """Returns (count, capped, rows or None, timed_out, seconds)."""
    res = OcResult()
    if want_list:
        lc = min(cap, list_cap if list_cap is not None else (1 << 20))
        buf = np.zeros((max(lc, 1), query.n), dtype=np.uint32)
        ptr = buf.ctypes.data_as(C.POINTER(C.c_uint32))
    else:
        lc, buf, ptr = 0, None, None
    _check("oc_match", lib().oc_match(data.h, C.byref(query.q), C.byref(plan.p), C.c_uint64(cap),
                                      ptr, C.c_uint64(lc), threads, C.c_double(budget_s),
                                      1 if refine else 0, C.byref(res)))
    rows = buf[: res.listed] if want_list else None
    return res.count, bool(res.capped), rows, bool(res.timed_out), res.seconds


def ref_decycle_available():
    return os.path.exists(REF_DECYCLE_PATH)


def ref_decycle(labels, edges):
    """The reference's own decycle() (built into oracle/_ref from its source)."""
    L = C.CDLL(REF_DECYCLE_PATH)
    u32p = C.POINTER(C.c_uint32)
    L.ref_decycle.argtypes = [C.c_uint32, u32p, C.c_uint32, u32p, u32p, u32p, u32p, u32p]
    labels, lp = _u32(labels)
    edges, ep = _u32(np.asarray(edges, dtype=np.uint32).reshape(-1))
    n = len(labels)
    m = len(edges) // 2
    off = np.zeros(n + 1, np.uint32)
    adj = np.zeros(max(2 * m, 1), np.uint32)
    deg = np.zeros(n, np.uint32)
    tm = np.zeros(1, np.uint32)
    L.ref_decycle(n, lp, m, ep, off.ctypes.data_as(u32p), adj.ctypes.data_as(u32p),
                  deg.ctypes.data_as(u32p), tm.ctypes.data_as(u32p))
    return list(off), list(adj[: off[-1]]), list(deg), int(tm[0])
