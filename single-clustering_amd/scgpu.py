"""Host-side mirror of single-clustering's public API over libscgpu's C ABI.

Same names, argument meaning and error behaviour as the Rust crate
(SingleRust/single-clustering v0.7.0): ``CSRNetwork.from_csr_parts`` (csr_network.rs:185),
``leiden(graph, LeidenConfig)`` (leiden/mod.rs:168), ``ObjectiveKind`` (objective.rs:188),
``modularity`` (objective.rs:221), ``Clustering`` (clustering.rs:11) and ``ClusteringError``
(error.rs:7). All computation runs in the HIP library on the GPU; there is no CPU fallback:
if libscgpu.so is missing or fails to load, importing this module raises.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscgpu.so")

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class ScError(C.Structure):
    _fields_ = [
        ("code", C.c_int32),
        ("node", C.c_uint64),
        ("n_nodes", C.c_uint64),
        ("edge_u", C.c_uint64),
        ("edge_v", C.c_uint64),
        ("weight", C.c_double),
        ("got", C.c_uint64),
        ("expected", C.c_uint64),
        ("degree_sum", C.c_double),
        ("expected_sum", C.c_double),
        ("msg", C.c_char * 256),
    ]


class ScLeidenConfig(C.Structure):
    _fields_ = [
        ("objective", C.c_int32),
        ("resolution", C.c_double),
        ("has_seed", C.c_int32),
        ("seed", C.c_uint64),
        ("n_iterations", C.c_uint64),
        ("max_levels", C.c_uint64),
        ("tolerance", C.c_double),
        ("refine", C.c_int32),
        ("refine_randomness", C.c_double),
        ("patience", C.c_uint64),
        ("has_max_community_weight", C.c_int32),
        ("max_community_weight", C.c_double),
    ]


class ScLeidenStats(C.Structure):
    _fields_ = [
        ("passes", C.c_uint64),
        ("levels", C.c_uint64),
        ("move_iterations", C.c_uint64),
        ("refine_rounds", C.c_uint64),
        ("vertex_visits", C.c_uint64),
        ("edge_visits", C.c_uint64),
        ("moves", C.c_uint64),
        ("algorithmic_bytes", C.c_double),
        ("move_kernel_ms", C.c_double),
        ("move_kernel_bytes", C.c_double),
        ("move_kernel_launches", C.c_uint64),
        ("total_ms", C.c_double),
        ("phase_ms", C.c_double * 6),
    ]


class ScKnnOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("mode", C.c_int32)]


class ScKnnStats(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("candidates_rescored", C.c_uint64),
        ("rows_fallback", C.c_uint64),
        ("tile_kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("flops", C.c_double),
        ("pairs_computed", C.c_double),
    ]


class ScFilterStats(C.Structure):
    _fields_ = [("skipped", C.c_uint64), ("judged", C.c_uint64), ("judge_moves", C.c_uint64),
                ("skipped_at_judge", C.c_uint64), ("skipped_moved", C.c_uint64), ("skip_error_rate", C.c_double)]


# Every symbol include/scgpu.h declares, with its ctypes signature.
_GP = C.c_void_p
SIGNATURES = {
    "sc_version": ([], C.c_char_p),
    "sc_shutdown": ([], None),
    "sc_pool_bytes": ([u64p, u64p], None),
    "sc_set_profiling": ([C.c_int32], None),
    "sc_set_sweep_mode": ([C.c_int32], C.c_int),
    "sc_leiden_config_default": ([], ScLeidenConfig),
    "sc_graph_from_csr_parts": ([u64p, C.c_uint64, u32p, f32p, C.c_uint64, f64p, C.POINTER(_GP),
                                 C.POINTER(ScError)], C.c_int),
    "sc_graph_from_csr_export": ([C.c_char_p, C.POINTER(_GP), C.POINTER(ScError)], C.c_int),
    "sc_graph_from_csr_parts_device": ([C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint64,
                                        C.c_void_p, C.POINTER(_GP), C.POINTER(ScError)], C.c_int),
    "sc_graph_from_edges": ([C.c_uint64, u64p, u64p, f64p, C.c_uint64, f64p, C.POINTER(_GP),
                             C.POINTER(ScError)], C.c_int),
    "sc_graph_from_csr_matrix": ([C.c_uint64, u64p, u64p, f64p, C.c_uint64, f64p, C.POINTER(_GP), C.POINTER(ScError)],
                                 C.c_int),
    "sc_graph_to_csr_matrix": ([_GP, u64p, u64p, f64p, C.POINTER(ScError)], C.c_int),
    "sc_graph_free": ([_GP], None),
    "sc_graph_node_count": ([_GP], C.c_uint64),
    "sc_graph_edge_count": ([_GP], C.c_uint64),
    "sc_graph_nnz": ([_GP], C.c_uint64),
    "sc_graph_total_weight": ([_GP], C.c_double),
    "sc_graph_has_self_loops": ([_GP], C.c_int32),
    "sc_graph_to_csr_parts": ([_GP, u64p, u32p, f32p, f64p, f64p, C.POINTER(ScError)], C.c_int),
    "sc_graph_validate_symmetry": ([_GP, C.POINTER(ScError)], C.c_int),
    "sc_graph_aggregate": ([_GP, u64p, C.POINTER(_GP), C.POINTER(ScError)], C.c_int),
    "sc_leiden": ([_GP, C.POINTER(ScLeidenConfig), u64p, u64p, C.POINTER(ScError)], C.c_int),
    "sc_autotune_report": ([C.c_char_p, C.c_size_t], C.c_size_t),
    "sc_autotune_reset": ([], None),
    "sc_mlp_param_count": ([], C.c_uint32),
    "sc_mlp_init": ([C.c_uint64, C.POINTER(C.c_float)], None),
    "sc_mlp_grad": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8), C.c_uint64, C.c_float,
                     C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(ScError)], C.c_int),
    "sc_mlp_train": ([C.POINTER(C.c_float), C.POINTER(C.c_uint8), C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                      C.c_float, C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(ScError)], C.c_int),
    "sc_mlp_predict": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint64, C.c_int32, C.POINTER(C.c_float),
                        C.POINTER(ScError)], C.c_int),
    "sc_leiden_trace": ([_GP, C.POINTER(ScLeidenConfig), C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                         u64p, C.POINTER(ScError)], C.c_int),
    "sc_leiden_filtered": ([_GP, C.POINTER(ScLeidenConfig), C.POINTER(C.c_float), C.c_double, C.c_int32, u64p, u64p,
                            C.POINTER(ScFilterStats), C.POINTER(ScError)], C.c_int),
    "sc_leiden_batch": ([C.POINTER(_GP), C.c_uint64, C.POINTER(C.c_double), C.c_uint64, C.POINTER(ScLeidenConfig),
                         u64p, u64p, C.POINTER(ScError)], C.c_int),
    "sc_leiden_last_stats": ([C.POINTER(ScLeidenStats)], C.c_int),
    "sc_modularity": ([_GP, u64p, C.c_double], C.c_double),
    "sc_quality": ([_GP, u64p, C.c_int32, C.c_double, f64p, C.POINTER(ScError)], C.c_int),
    "sc_renumber": ([u64p, C.c_uint64, u64p, u64p, C.POINTER(ScError)], C.c_int),
    "sc_quality_noalloc": ([_GP, C.c_void_p, C.c_int32, C.c_double, f64p, C.POINTER(ScError)], C.c_int),
    "sc_disconnected_community_count": ([_GP, u64p, u64p, C.POINTER(ScError)], C.c_int),
    "sc_knn": ([f32p, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(ScKnnOpts), u32p, f64p,
                C.POINTER(ScKnnStats), C.POINTER(ScError)], C.c_int),
    "sc_knn_graph": ([f32p, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(ScKnnOpts), C.POINTER(_GP),
                      C.POINTER(ScKnnStats), C.POINTER(ScError)], C.c_int),
    "sc_knn_graph_device": ([C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(ScKnnOpts),
                             C.POINTER(_GP), C.POINTER(ScKnnStats), C.POINTER(ScError)], C.c_int),
    "sc_knn_gaussian_csr": ([u32p, f64p, C.c_uint64, C.c_uint32, C.c_int32, u64p, u32p, f32p, u64p,
                             C.POINTER(ScError)], C.c_int),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python __graft_entry__.py` or "
            "`python single-clustering_amd/build.py` (there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = _load()

VARIANTS = {
    1: "NodeIndexOutOfRange",
    2: "NonFiniteWeight",
    3: "NegativeWeight",
    4: "InvalidNodeWeight",
    5: "NodeWeightLengthMismatch",
    6: "MembershipLengthMismatch",
    7: "InvalidConfig",
    8: "TooManyNodes",
    9: "InvalidCsr",
    10: "AsymmetricGraph",
    100: "DeviceError",
    101: "InvalidArgument",
    102: "OutOfMemory",
}


class ClusteringError(Exception):
    """ClusteringError (error.rs:7-68): ``variant`` names the enum variant, payload fields
    carry its data."""

    def __init__(self, err):
        self.code = err.code
        self.variant = VARIANTS.get(err.code, f"Unknown{err.code}")
        self.message = err.msg.decode(errors="replace")
        self.node = err.node
        self.n_nodes = err.n_nodes
        self.edge = (err.edge_u, err.edge_v)
        self.weight = err.weight
        self.got = err.got
        self.expected = err.expected
        self.degree_sum = err.degree_sum
        self.expected_sum = err.expected_sum
        super().__init__(f"{self.variant}: {self.message}")


def _check(status, err):
    if status != 0:
        if err.code == 0:
            err.code = status
        raise ClusteringError(err)


def _p(a, t):
    return a.ctypes.data_as(t)


def lib():
    return _lib


def version():
    return _lib.sc_version().decode()


def pool_bytes():
    """(live, cached) device bytes of the calling thread's allocator (diagnostic)."""
    live, cached = C.c_uint64(), C.c_uint64()
    _lib.sc_pool_bytes(C.byref(live), C.byref(cached))
    return live.value, cached.value


def set_profiling(on):
    _lib.sc_set_profiling(1 if on else 0)


SWEEP_MODES = {"multi": 0, "persistent": 1, "graph": 2}


def set_sweep_mode(mode):
    """"persistent" (or True): one cooperative launch per local-moving / refinement phase;
    "multi" (or False): one launch per sub-sweep kernel; "graph": the multi-launch iteration
    captured into a hipGraph once per level and replayed per iteration. Labels are identical."""
    code = SWEEP_MODES.get(mode) if isinstance(mode, str) else (1 if mode else 0)
    if code is None or _lib.sc_set_sweep_mode(code) != 0:
        raise ValueError(f"invalid sweep mode {mode!r}")


@dataclass(frozen=True)
class ObjectiveKind:
    """ObjectiveKind::{Rb, Cpm} (objective.rs:188-199)."""

    kind: str = "rb"
    resolution: float = 1.0

    @staticmethod
    def Rb(resolution=1.0):
        return ObjectiveKind("rb", float(resolution))

    @staticmethod
    def Cpm(resolution=0.05):
        return ObjectiveKind("cpm", float(resolution))

    @property
    def code(self):
        return 0 if self.kind == "rb" else 1


@dataclass
class LeidenConfig:
    """LeidenConfig (leiden/mod.rs:49-106) with the reference's defaults."""

    objective: ObjectiveKind = field(default_factory=ObjectiveKind.Rb)
    seed: object = 42
    n_iterations: int = 2
    max_levels: int = 64
    tolerance: float = 1e-9
    refine: bool = True
    refine_randomness: float = 0.01
    patience: int = 2
    max_community_weight: object = None

    @staticmethod
    def with_resolution(resolution):
        return LeidenConfig(objective=ObjectiveKind.Rb(resolution))

    def to_c(self):
        c = _lib.sc_leiden_config_default()
        c.objective = self.objective.code
        c.resolution = self.objective.resolution
        c.has_seed = 0 if self.seed is None else 1
        c.seed = 0 if self.seed is None else int(self.seed)
        c.n_iterations = int(self.n_iterations)
        c.max_levels = int(self.max_levels)
        c.tolerance = float(self.tolerance)
        c.refine = 1 if self.refine else 0
        c.refine_randomness = float(self.refine_randomness)
        c.patience = int(self.patience)
        c.has_max_community_weight = 0 if self.max_community_weight is None else 1
        c.max_community_weight = 0.0 if self.max_community_weight is None else float(self.max_community_weight)
        return c


class Clustering:
    """Clustering (clustering.rs:11-108): consecutive labels and the cluster count."""

    def __init__(self, labels, n_clusters):
        self._labels = labels
        self._n = int(n_clusters)

    def labels(self):
        return self._labels

    def n_clusters(self):
        return self._n

    def __len__(self):
        return len(self._labels)

    def cluster_sizes(self):
        return np.bincount(self._labels.astype(np.int64), minlength=self._n)

    def __eq__(self, other):
        return (isinstance(other, Clustering) and self._n == other._n
                and np.array_equal(self._labels, other._labels))


class CSRNetwork:
    """Device-resident CSRNetwork (csr_network.rs:38-69)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.sc_graph_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def from_csr_parts(cls, node_ptrs, neighbors, weights, node_weights=None):
        node_ptrs = np.ascontiguousarray(node_ptrs, dtype=np.uint64)
        neighbors = np.ascontiguousarray(neighbors, dtype=np.uint32)
        weights = np.ascontiguousarray(weights, dtype=np.float32)
        if len(node_ptrs) == 0:
            err = ScError()
            err.code = 9
            err.msg = b"malformed CSR input: node_ptrs must have at least one element"
            raise ClusteringError(err)
        if len(neighbors) != len(weights):
            err = ScError()
            err.code = 9
            err.msg = (f"malformed CSR input: neighbors has {len(neighbors)} entries but weights has "
                       f"{len(weights)}").encode()
            raise ClusteringError(err)
        n = len(node_ptrs) - 1
        nw = None
        if node_weights is not None:
            nw = np.ascontiguousarray(node_weights, dtype=np.float64)
            if len(nw) != n:
                err = ScError()
                err.code = 5
                err.got = len(nw)
                err.expected = n
                err.msg = f"node weight vector has length {len(nw)}, expected {n}".encode()
                raise ClusteringError(err)
        h = C.c_void_p()
        err = ScError()
        st = _lib.sc_graph_from_csr_parts(_p(node_ptrs, u64p), n, _p(neighbors, u32p), _p(weights, f32p),
                                          len(neighbors), None if nw is None else _p(nw, f64p),
                                          C.byref(h), C.byref(err))
        _check(st, err)
        return cls(h)

    @classmethod
    def from_csr_export(cls, directory):
        """The real-data example's reader (examples/real_data.rs:22-44, 129-136): the
        connectivities matrix as `indptr.bin` (u64), `indices.bin` (u32) and `data.bin` (f32),
        raw little-endian, read by the library and handed to from_csr_parts."""
        h = C.c_void_p()
        err = ScError()
        st = _lib.sc_graph_from_csr_export(os.fsencode(os.fspath(directory)), C.byref(h), C.byref(err))
        _check(st, err)
        return cls(h)

    @classmethod
    def from_edges(cls, n_nodes, edges, node_weights=None):
        """from_edges / from_edges_with_node_weights (csr_network.rs:104-167)."""
        m = len(edges)
        a = np.zeros(max(m, 1), dtype=np.uint64)
        b = np.zeros(max(m, 1), dtype=np.uint64)
        w = np.zeros(max(m, 1), dtype=np.float64)
        for i, (x, y, z) in enumerate(edges):
            a[i], b[i], w[i] = x, y, z
        nw = None if node_weights is None else np.ascontiguousarray(node_weights, dtype=np.float64)
        if nw is not None:
            n_nodes = len(nw)
        h = C.c_void_p()
        err = ScError()
        st = _lib.sc_graph_from_edges(n_nodes, _p(a, u64p), _p(b, u64p), _p(w, f64p), m,
                                      None if nw is None else _p(nw, f64p), C.byref(h), C.byref(err))
        _check(st, err)
        return cls(h)

    @classmethod
    def from_csr_matrix(cls, matrix, node_weights=None):
        """from_csr_matrix / from_csr_matrix_with_node_weights (csr_network.rs:360-391): a square
        sparse matrix (scipy.sparse CSR, or a (row_ptrs, col_indices, values) tuple) read as
        symmetric (entries with row <= col, zeros skipped)."""
        if isinstance(matrix, tuple):
            indptr, indices, data = matrix
            n = len(indptr) - 1
        else:
            indptr, indices, data = matrix.indptr, matrix.indices, matrix.data
            n = matrix.shape[0]
        indptr = np.ascontiguousarray(indptr, dtype=np.uint64)
        indices = np.ascontiguousarray(indices, dtype=np.uint64)
        data = np.ascontiguousarray(data, dtype=np.float64)
        nw = None
        if node_weights is not None:
            nw = np.ascontiguousarray(node_weights, dtype=np.float64)
            if len(nw) != n:
                err = ScError()
                err.code = 5
                err.got = len(nw)
                err.expected = n
                err.msg = f"node weight vector has length {len(nw)}, expected {n}".encode()
                raise ClusteringError(err)
        h = C.c_void_p()
        err = ScError()
        st = _lib.sc_graph_from_csr_matrix(n, _p(indptr, u64p), _p(indices, u64p), _p(data, f64p), len(indices),
                                           None if nw is None else _p(nw, f64p), C.byref(h), C.byref(err))
        _check(st, err)
        return cls(h)

    def to_csr_matrix(self):
        """to_csr_matrix (csr_network.rs:610-618): (row_ptrs u64, col_indices u64, values f64)."""
        n, e = self.node_count(), self.nnz()
        ptrs = np.zeros(n + 1, dtype=np.uint64)
        cols = np.zeros(max(e, 1), dtype=np.uint64)
        vals = np.zeros(max(e, 1), dtype=np.float64)
        err = ScError()
        _check(_lib.sc_graph_to_csr_matrix(self._h, _p(ptrs, u64p), _p(cols, u64p), _p(vals, f64p), C.byref(err)), err)
        return ptrs, cols[:e], vals[:e]

    def node_count(self):
        return _lib.sc_graph_node_count(self._h)

    def edge_count(self):
        return _lib.sc_graph_edge_count(self._h)

    def nnz(self):
        return _lib.sc_graph_nnz(self._h)

    def total_weight(self):
        return _lib.sc_graph_total_weight(self._h)

    def has_self_loops(self):
        return bool(_lib.sc_graph_has_self_loops(self._h))

    def to_csr_parts(self):
        n, e = self.node_count(), self.nnz()
        ptrs = np.zeros(n + 1, dtype=np.uint64)
        nbrs = np.zeros(max(e, 1), dtype=np.uint32)
        wts = np.zeros(max(e, 1), dtype=np.float32)
        err = ScError()
        _check(_lib.sc_graph_to_csr_parts(self._h, _p(ptrs, u64p), _p(nbrs, u32p), _p(wts, f32p), None, None,
                                          C.byref(err)), err)
        return ptrs, nbrs[:e], wts[:e]

    def node_weights(self):
        n = self.node_count()
        out = np.zeros(max(n, 1), dtype=np.float64)
        err = ScError()
        _check(_lib.sc_graph_to_csr_parts(self._h, None, None, None, _p(out, f64p), None, C.byref(err)), err)
        return out[:n]

    def strengths(self):
        n = self.node_count()
        out = np.zeros(max(n, 1), dtype=np.float64)
        err = ScError()
        _check(_lib.sc_graph_to_csr_parts(self._h, None, None, None, None, _p(out, f64p), C.byref(err)), err)
        return out[:n]

    def strength(self, v):
        return float(self.strengths()[v])

    def validate_symmetry(self):
        err = ScError()
        _check(_lib.sc_graph_validate_symmetry(self._h, C.byref(err)), err)

    def aggregate(self, groups):
        groups = np.ascontiguousarray(groups, dtype=np.uint64)
        h = C.c_void_p()
        err = ScError()
        _check(_lib.sc_graph_aggregate(self._h, _p(groups, u64p), C.byref(h), C.byref(err)), err)
        return CSRNetwork(h)


def leiden(graph, config=None):
    """leiden(&CSRNetwork, &LeidenConfig) -> Clustering (leiden/mod.rs:168-204)."""
    config = config or LeidenConfig()
    c = config.to_c()
    n = graph.node_count()
    labels = np.zeros(max(n, 1), dtype=np.uint64)
    nc = C.c_uint64(0)
    err = ScError()
    _check(_lib.sc_leiden(graph.handle, C.byref(c), _p(labels, u64p), C.byref(nc), C.byref(err)), err)
    return Clustering(labels[:n], nc.value)


def leiden_batch(graphs, resolutions, config=None):
    """Every graph at every resolution in one batched run; returns the Clustering of job
    g * len(resolutions) + r, each equal to leiden(graphs[g], config.with_resolution(r))."""
    config = config or LeidenConfig()
    graphs = list(graphs)
    res = np.ascontiguousarray(resolutions, dtype=np.float64)
    c = config.to_c()
    sizes = [g.node_count() for g in graphs for _ in range(len(res))]
    labels = np.zeros(max(sum(sizes), 1), dtype=np.uint64)
    nc = np.zeros(max(len(sizes), 1), dtype=np.uint64)
    handles = (_GP * len(graphs))(*[g.handle for g in graphs])
    err = ScError()
    _check(_lib.sc_leiden_batch(handles, len(graphs), res.ctypes.data_as(C.POINTER(C.c_double)), len(res),
                                C.byref(c), _p(labels, u64p), _p(nc, u64p), C.byref(err)), err)
    out, off = [], 0
    for j, n in enumerate(sizes):
        out.append(Clustering(labels[off:off + n].copy(), int(nc[j])))
        off += n
    return out


_f32p = C.POINTER(C.c_float)


def _fp(a):
    return a.ctypes.data_as(_f32p)


def autotune_report():
    """The online autotuner's table: one line per tuned shape."""
    n = _lib.sc_autotune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    _lib.sc_autotune_report(buf, n + 1)
    return buf.value.decode()


def autotune_reset():
    _lib.sc_autotune_reset()


def mlp_param_count():
    return int(_lib.sc_mlp_param_count())


def mlp_init(seed=0):
    """He-uniform initial parameters of the learned filter (8 -> 64 -> 64 -> 1)."""
    w = np.zeros(mlp_param_count(), dtype=np.float32)
    _lib.sc_mlp_init(int(seed), _fp(w))
    return w


def _mlp_data(X, y):
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
    y = np.ascontiguousarray(y, dtype=np.uint8).reshape(-1)
    assert len(X) == len(y)
    return X, y


def mlp_grad(w, X, y, pos_weight=1.0):
    """(mean loss, gradient) of the filter's loss on (X, y), from the HIP backward pass."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    X, y = _mlp_data(X, y)
    g = np.zeros_like(w)
    loss = C.c_double(0.0)
    err = ScError()
    _check(_lib.sc_mlp_grad(_fp(w), _fp(X), y.ctypes.data_as(C.POINTER(C.c_uint8)), len(y), float(pos_weight),
                            _fp(g), C.byref(loss), C.byref(err)), err)
    return loss.value, g


def mlp_train(X, y, epochs=10, batch=4096, lr=2e-3, pos_weight=1.0, seed=0, w=None):
    """Minibatch Adam on the GPU; returns (weights, per-epoch mean loss)."""
    w = mlp_init(seed) if w is None else np.array(w, dtype=np.float32)
    X, y = _mlp_data(X, y)
    hist = np.zeros(epochs, dtype=np.float64)
    err = ScError()
    _check(_lib.sc_mlp_train(_fp(X), y.ctypes.data_as(C.POINTER(C.c_uint8)), len(y), int(epochs), int(batch),
                             float(lr), float(pos_weight), int(seed), _fp(w),
                             hist.ctypes.data_as(C.POINTER(C.c_double)), C.byref(err)), err)
    return w, hist


def mlp_predict(w, X, bf16=True):
    """Logits of the filter (BF16-rounded weights by default, as local moving evaluates them)."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
    out = np.zeros(len(X), dtype=np.float32)
    err = ScError()
    _check(_lib.sc_mlp_predict(_fp(w), _fp(X), len(X), 1 if bf16 else 0, _fp(out), C.byref(err)), err)
    return out


def leiden_trace(graph, config=None, cap=1 << 22):
    """Runs leiden() and returns up to cap local-moving decisions: (features[n][8], moved[n])."""
    config = config or LeidenConfig()
    c = config.to_c()
    X = np.zeros((max(cap, 1), 8), dtype=np.float32)
    y = np.zeros(max(cap, 1), dtype=np.uint8)
    n = C.c_uint64(0)
    err = ScError()
    _check(_lib.sc_leiden_trace(graph.handle, C.byref(c), int(cap), _fp(X), y.ctypes.data_as(C.POINTER(C.c_uint8)),
                                C.byref(n), C.byref(err)), err)
    return X[:n.value].copy(), y[:n.value].copy()


JUDGES = {False: 0, True: 1, "none": 0, "full": 1, "skipped": 2}


def leiden_filtered(graph, weights, threshold=0.02, config=None, judge="full"):
    """leiden() with the learned active-vertex filter in local moving; returns (Clustering,
    stats) where stats holds the skip count, the judge iteration's size and moves, the skipped
    vertices it moved and the skip-error rate. A skipped vertex leaves the active set until a
    neighbour moves. At each level's convergence the judge runs plain iterations from every
    vertex ("full" or True), from the vertices last skipped ("skipped"), or not at all ("none"
    or False: only the skip count is then reported)."""
    config = config or LeidenConfig()
    c = config.to_c()
    w = np.ascontiguousarray(weights, dtype=np.float32)
    n = graph.node_count()
    labels = np.zeros(max(n, 1), dtype=np.uint64)
    nc = C.c_uint64(0)
    st = ScFilterStats()
    err = ScError()
    if judge not in JUDGES:
        raise ValueError(f"invalid judge {judge!r}")
    _check(_lib.sc_leiden_filtered(graph.handle, C.byref(c), _fp(w), float(threshold), JUDGES[judge],
                                   _p(labels, u64p), C.byref(nc),
                                   C.byref(st), C.byref(err)), err)
    return Clustering(labels[:n], nc.value), {k: getattr(st, k) for k, _ in ScFilterStats._fields_}


def last_stats():
    st = ScLeidenStats()
    _lib.sc_leiden_last_stats(C.byref(st))
    out = {k: getattr(st, k) for k, _ in ScLeidenStats._fields_}
    out["phase_ms"] = dict(zip(["buckets", "local_move", "refine", "renumber", "aggregate", "quality"],
                               list(st.phase_ms)))
    return out


def modularity(graph, labels, resolution=1.0):
    """modularity (objective.rs:221-244)."""
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    if len(labels) < graph.node_count():
        raise IndexError("labels shorter than the node count")
    return _lib.sc_modularity(graph.handle, _p(labels, u64p), float(resolution))


def quality(graph, labels, objective):
    """Rb::quality / Cpm::quality of Partition::from_membership(labels)."""
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    out = C.c_double(0.0)
    err = ScError()
    _check(_lib.sc_quality(graph.handle, _p(labels, u64p), objective.code, objective.resolution, C.byref(out),
                           C.byref(err)), err)
    return out.value


def renumber(labels):
    """Partition::renumber_into_grouping: (ids compacted ascending on the device, count)."""
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    out = np.zeros(len(labels), dtype=np.uint64)
    k = C.c_uint64(0)
    err = ScError()
    _check(_lib.sc_renumber(_p(labels, u64p), len(labels), _p(out, u64p), C.byref(k), C.byref(err)), err)
    return out, k.value


def disconnected_community_count(graph, labels):
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    out = C.c_uint64(0)
    err = ScError()
    _check(_lib.sc_disconnected_community_count(graph.handle, _p(labels, u64p), C.byref(out), C.byref(err)), err)
    return out.value


def knn(data, k, precision="bf16", return_stats=False):
    """Exact kNN (self excluded): indices (n, k) uint32 and squared distances (n, k) float64."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, d = data.shape
    idx = np.zeros((n, k), dtype=np.uint32)
    d2 = np.zeros((n, k), dtype=np.float64)
    opts = ScKnnOpts(0 if precision == "bf16" else 1, 0)
    st = ScKnnStats()
    err = ScError()
    _check(_lib.sc_knn(_p(data, f32p), n, d, k, C.byref(opts), _p(idx, u32p), _p(d2, f64p), C.byref(st),
                       C.byref(err)), err)
    if return_stats:
        return idx, d2, {k2: getattr(st, k2) for k2, _ in ScKnnStats._fields_}
    return idx, d2


def knn_graph(data, k, precision="bf16", mode=0, return_stats=False):
    """kNN -> Gaussian connectivity -> symmetrization -> CSRNetwork, device-resident."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, d = data.shape
    opts = ScKnnOpts(0 if precision == "bf16" else 1, int(mode))
    h = C.c_void_p()
    st = ScKnnStats()
    err = ScError()
    _check(_lib.sc_knn_graph(_p(data, f32p), n, d, k, C.byref(opts), C.byref(h), C.byref(st), C.byref(err)), err)
    g = CSRNetwork(h)
    if return_stats:
        return g, {k2: getattr(st, k2) for k2, _ in ScKnnStats._fields_}
    return g


class DeviceArray:
    """A host array copied once into device memory (hipMalloc + a synchronous hipMemcpy through
    the HIP runtime the library itself loads), for the entry points that take device-resident
    input. `ptr` is the device address; free() releases it."""

    _hip = None

    def __init__(self, host):
        if DeviceArray._hip is None:
            hip = C.CDLL("libamdhip64.so")
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            hip.hipFree.argtypes = [C.c_void_p]
            DeviceArray._hip = hip
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        p = C.c_void_p()
        if DeviceArray._hip.hipMalloc(C.byref(p), max(self.nbytes, 1)) != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed")
        self.ptr = p.value
        if self.nbytes and DeviceArray._hip.hipMemcpy(p, host.ctypes.data, self.nbytes, 1) != 0:  # host to device
            self.free()
            raise RuntimeError("hipMemcpy to the device failed")

    def free(self):
        if self.ptr:
            DeviceArray._hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def knn_graph_device(d_data, n, d, k, precision="bf16", mode=0, return_stats=False):
    """knn_graph for points already resident in device memory: d_data is the device address of
    n x d row-major f32 (for example a torch tensor's data_ptr(), complete before the call)."""
    opts = ScKnnOpts(0 if precision == "bf16" else 1, int(mode))
    h = C.c_void_p()
    st = ScKnnStats()
    err = ScError()
    _check(_lib.sc_knn_graph_device(C.c_void_p(int(d_data)), int(n), int(d), int(k), C.byref(opts), C.byref(h),
                                    C.byref(st), C.byref(err)), err)
    g = CSRNetwork(h)
    if return_stats:
        return g, {k2: getattr(st, k2) for k2, _ in ScKnnStats._fields_}
    return g


def write_csr_export(directory, node_ptrs, neighbors, weights):
    """Writes a CSR matrix as the export tool does (tools/export_h5ad.py:106-108): indptr.bin
    (u64), indices.bin (u32) and data.bin (f32), raw little-endian, in `directory`."""
    os.makedirs(directory, exist_ok=True)
    np.ascontiguousarray(node_ptrs).astype("<u8").tofile(os.path.join(directory, "indptr.bin"))
    np.ascontiguousarray(neighbors).astype("<u4").tofile(os.path.join(directory, "indices.bin"))
    np.ascontiguousarray(weights).astype("<f4").tofile(os.path.join(directory, "data.bin"))


def knn_gaussian_csr(idx, d2, mode=0):
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    d2 = np.ascontiguousarray(d2, dtype=np.float64)
    n, k = idx.shape
    ptrs = np.zeros(n + 1, dtype=np.uint64)
    nbrs = np.zeros(2 * n * k + 1, dtype=np.uint32)
    wts = np.zeros(2 * n * k + 1, dtype=np.float32)
    nnz = C.c_uint64(0)
    err = ScError()
    _check(_lib.sc_knn_gaussian_csr(_p(idx, u32p), _p(d2, f64p), n, k, int(mode), _p(ptrs, u64p), _p(nbrs, u32p),
                                    _p(wts, f32p), C.byref(nnz), C.byref(err)), err)
    return ptrs, nbrs[:nnz.value], wts[:nnz.value]
