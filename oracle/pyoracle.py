"""ctypes binding of the CPU oracle (oracle.c). Test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
product library. Builds liboracle.so with `make -C oracle` when it is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class OrGraph(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("ptrs", u64p),
        ("nbrs", u32p),
        ("wts", f32p),
        ("node_w", f64p),
        ("strength", f64p),
        ("total_weight", C.c_double),
        ("edge_count", C.c_uint64),
        ("any_self_loops", C.c_int),
    ]


class OrEdge(C.Structure):
    _fields_ = [("from_", C.c_uint64), ("to", C.c_uint64), ("w", C.c_double)]


class OrConfig(C.Structure):
    _fields_ = [
        ("objective", C.c_int),
        ("resolution", C.c_double),
        ("has_seed", C.c_int),
        ("seed", C.c_uint64),
        ("n_iterations", C.c_uint64),
        ("max_levels", C.c_uint64),
        ("tolerance", C.c_double),
        ("refine", C.c_int),
        ("refine_randomness", C.c_double),
        ("patience", C.c_uint64),
        ("has_max_community_weight", C.c_int),
        ("max_community_weight", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        gp = C.POINTER(OrGraph)
        L.or_graph_from_csr_parts.argtypes = [C.c_uint64, u64p, u32p, f32p, C.c_uint64, f64p,
                                              C.POINTER(gp)]
        L.or_graph_from_edges.argtypes = [C.c_uint64, C.POINTER(OrEdge), C.c_uint64, f64p,
                                          C.POINTER(gp)]
        L.or_graph_free.argtypes = [gp]
        L.or_graph_validate_symmetry.argtypes = [gp]
        L.or_self_loop_weight.argtypes = [gp, C.c_uint64]
        L.or_self_loop_weight.restype = C.c_double
        L.or_graph_aggregate.argtypes = [gp, u64p, C.c_uint64]
        L.or_graph_aggregate.restype = gp
        L.or_modularity.argtypes = [gp, u64p, C.c_double]
        L.or_modularity.restype = C.c_double
        L.or_quality.argtypes = [gp, u64p, C.c_int, C.c_double]
        L.or_quality.restype = C.c_double
        L.or_leiden.argtypes = [gp, C.POINTER(OrConfig), u64p, u64p]
        L.or_local_move_once.argtypes = [gp, u64p, C.c_int, C.c_double, C.c_uint64, C.c_int,
                                         C.c_double]
        L.or_local_move_once.restype = C.c_double
        L.or_disconnected_community_count.argtypes = [gp, u64p]
        L.or_disconnected_community_count.restype = C.c_uint64
        L.or_normalize_groups.argtypes = [u64p, C.c_uint64, u64p]
        L.or_normalize_groups.restype = C.c_uint64
        L.or_nmi.argtypes = [u64p, u64p, C.c_uint64]
        L.or_nmi.restype = C.c_double
        L.or_ari.argtypes = [u64p, u64p, C.c_uint64]
        L.or_ari.restype = C.c_double
        L.or_config_default.argtypes = [C.POINTER(OrConfig)]
        L.or_chacha_block.argtypes = [u32p, C.c_uint64, C.c_uint64, C.c_int, u32p]
        L.or_chacha8_seed_u64_stream.argtypes = [C.c_uint64, u32p, C.c_uint64]
        L.or_knn_bruteforce.argtypes = [f32p, C.c_uint64, C.c_uint32, C.c_uint32, u64p,
                                        C.c_uint64, u32p, f64p]
        L.or_knn_gaussian.argtypes = [u32p, f64p, C.c_uint64, C.c_uint32, C.c_int, u64p, u32p,
                                      f32p]
        L.or_knn_kdtree.argtypes = [f32p, C.c_uint64, C.c_uint32, C.c_uint32, u64p, C.c_uint64,
                                    C.c_uint32, u32p, f64p, f64p]
        L.or_knn_kdtree.restype = C.c_int
        L.or_knn_gaussian.restype = C.c_uint64
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleError(Exception):
    def __init__(self, code):
        super().__init__(f"oracle status {code}")
        self.code = code


class Graph:
    """CSRNetwork restatement (csr_network.rs)."""

    def __init__(self, ptr):
        self._g = ptr

    def __del__(self):
        if getattr(self, "_g", None):
            lib().or_graph_free(self._g)
            self._g = None

    @classmethod
    def from_csr_parts(cls, node_ptrs, neighbors, weights, node_weights=None):
        node_ptrs = np.ascontiguousarray(node_ptrs, dtype=np.uint64)
        neighbors = np.ascontiguousarray(neighbors, dtype=np.uint32)
        weights = np.ascontiguousarray(weights, dtype=np.float32)
        n = len(node_ptrs) - 1
        nw = None if node_weights is None else np.ascontiguousarray(node_weights, dtype=np.float64)
        out = C.POINTER(OrGraph)()
        st = lib().or_graph_from_csr_parts(
            n, _p(node_ptrs, u64p), _p(neighbors, u32p), _p(weights, f32p), len(neighbors),
            None if nw is None else _p(nw, f64p), C.byref(out))
        if st:
            raise OracleError(st)
        return cls(out)

    @classmethod
    def from_edges(cls, n, edges, node_weights=None):
        arr = (OrEdge * max(len(edges), 1))()
        for i, (a, b, w) in enumerate(edges):
            arr[i].from_ = a
            arr[i].to = b
            arr[i].w = w
        nw = None if node_weights is None else np.ascontiguousarray(node_weights, dtype=np.float64)
        out = C.POINTER(OrGraph)()
        st = lib().or_graph_from_edges(n, arr, len(edges), None if nw is None else _p(nw, f64p),
                                       C.byref(out))
        if st:
            raise OracleError(st)
        return cls(out)

    @property
    def n(self):
        return self._g.contents.n

    def node_count(self):
        return self.n

    def edge_count(self):
        return self._g.contents.edge_count

    def total_weight(self):
        return self._g.contents.total_weight

    def has_self_loops(self):
        return bool(self._g.contents.any_self_loops)

    def csr(self):
        g = self._g.contents
        n = g.n
        ptrs = np.ctypeslib.as_array(g.ptrs, shape=(n + 1,)).copy()
        e = int(ptrs[-1])
        nbrs = np.ctypeslib.as_array(g.nbrs, shape=(max(e, 1),))[:e].copy()
        wts = np.ctypeslib.as_array(g.wts, shape=(max(e, 1),))[:e].copy()
        return ptrs, nbrs, wts

    def strengths(self):
        g = self._g.contents
        return np.ctypeslib.as_array(g.strength, shape=(max(g.n, 1),))[:g.n].copy()

    def node_weights(self):
        g = self._g.contents
        return np.ctypeslib.as_array(g.node_w, shape=(max(g.n, 1),))[:g.n].copy()

    def self_loop_weight(self, v):
        return lib().or_self_loop_weight(self._g, v)

    def validate_symmetry(self):
        return lib().or_graph_validate_symmetry(self._g) == 0

    def aggregate(self, groups):
        groups = np.ascontiguousarray(groups, dtype=np.uint64)
        ng = int(groups.max()) + 1 if len(groups) else 0
        return Graph(lib().or_graph_aggregate(self._g, _p(groups, u64p), ng))


def default_config(**kw):
    cfg = OrConfig()
    lib().or_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k == "objective":
            cfg.objective = {"rb": 0, "cpm": 1}.get(v, v)
        elif k == "seed":
            if v is None:
                cfg.has_seed = 0
            else:
                cfg.has_seed = 1
                cfg.seed = v
        elif k == "max_community_weight":
            if v is None:
                cfg.has_max_community_weight = 0
            else:
                cfg.has_max_community_weight = 1
                cfg.max_community_weight = v
        else:
            setattr(cfg, k, v)
    return cfg


def leiden(g, **kw):
    cfg = default_config(**kw)
    labels = np.zeros(max(g.n, 1), dtype=np.uint64)
    nc = C.c_uint64(0)
    st = lib().or_leiden(g._g, C.byref(cfg), _p(labels, u64p), C.byref(nc))
    if st:
        raise OracleError(st)
    return labels[:g.n], nc.value


def modularity(g, labels, resolution=1.0):
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    return lib().or_modularity(g._g, _p(labels, u64p), resolution)


def quality(g, labels, objective="rb", resolution=1.0):
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    obj = {"rb": 0, "cpm": 1}.get(objective, objective)
    return lib().or_quality(g._g, _p(labels, u64p), obj, resolution)


def local_move_once(g, membership, objective="rb", resolution=1.0, seed=0, max_w=None):
    m = np.ascontiguousarray(membership, dtype=np.uint64).copy()
    obj = {"rb": 0, "cpm": 1}.get(objective, objective)
    gain = lib().or_local_move_once(g._g, _p(m, u64p), obj, resolution, seed,
                                    0 if max_w is None else 1, 0.0 if max_w is None else max_w)
    return m, gain


def disconnected_community_count(g, labels):
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    return lib().or_disconnected_community_count(g._g, _p(labels, u64p))


def normalize_groups(labels):
    """VectorGrouping::from_assignments normalization: (ids compacted ascending, count)."""
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    out = np.zeros(len(labels), dtype=np.uint64)
    k = lib().or_normalize_groups(_p(labels, u64p), len(labels), _p(out, u64p))
    return out, k


def nmi(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    return lib().or_nmi(_p(a, u64p), _p(b, u64p), len(a))


def ari(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    return lib().or_ari(_p(a, u64p), _p(b, u64p), len(a))


def chacha_block(key_words, counter, stream, rounds):
    key = np.ascontiguousarray(key_words, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().or_chacha_block(_p(key, u32p), counter, stream, rounds, _p(out, u32p))
    return out


def chacha8_stream(seed, count):
    out = np.zeros(count, dtype=np.uint32)
    lib().or_chacha8_seed_u64_stream(seed, _p(out, u32p), count)
    return out


def knn_bruteforce(data, k, rows=None):
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, d = data.shape
    if rows is None:
        nq = n
        rp = None
    else:
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        nq = len(rows)
        rp = _p(rows, u64p)
    idx = np.zeros((nq, k), dtype=np.uint32)
    d2 = np.zeros((nq, k), dtype=np.float64)
    lib().or_knn_bruteforce(_p(data, f32p), n, d, k, rp, nq, _p(idx, u32p), _p(d2, f64p))
    return idx, d2


def knn_kdtree(data, k, rows=None, threads=1, return_build_seconds=False):
    """Exact kNN through the kd-tree restatement of the reference's n <= 250 000 path
    (neighborhood/mod.rs:51-98, 233-314), on `threads` workers; equals knn_bruteforce."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, d = data.shape
    if rows is None:
        nq, rp = n, None
    else:
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        nq, rp = len(rows), _p(rows, u64p)
    idx = np.zeros((nq, k), dtype=np.uint32)
    d2 = np.zeros((nq, k), dtype=np.float64)
    tb = C.c_double(0.0)
    rc = lib().or_knn_kdtree(_p(data, f32p), n, d, k, rp, nq, int(threads), _p(idx, u32p), _p(d2, f64p),
                             C.byref(tb))
    if rc:
        raise OracleError(rc)
    return (idx, d2, tb.value) if return_build_seconds else (idx, d2)


def knn_gaussian(idx, d2, mode=0):
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    d2 = np.ascontiguousarray(d2, dtype=np.float64)
    n, k = idx.shape
    ptrs = np.zeros(n + 1, dtype=np.uint64)
    nbrs = np.zeros(2 * n * k + 1, dtype=np.uint32)
    wts = np.zeros(2 * n * k + 1, dtype=np.float32)
    nnz = lib().or_knn_gaussian(_p(idx, u32p), _p(d2, f64p), n, k, mode, _p(ptrs, u64p),
                                _p(nbrs, u32p), _p(wts, f32p))
    return ptrs, nbrs[:nnz].copy(), wts[:nnz].copy()
