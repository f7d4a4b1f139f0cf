"""The C-ABI library loads and exports every symbol include/scgpu.h declares (no GPU calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "scgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sc_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for required in ["sc_graph_from_csr_parts", "sc_leiden", "sc_modularity", "sc_leiden_config_default",
                     "sc_knn_graph", "sc_graph_aggregate", "sc_quality"]:
        assert required in syms


def test_library_exports_every_declared_symbol():
    import scgpu  # raises if libscgpu.so is missing: there is no CPU fallback

    missing = [s for s in declared_symbols() if not hasattr(scgpu.lib(), s)]
    assert not missing, missing
    # and the ctypes mirror binds exactly the declared set
    assert sorted(scgpu.SIGNATURES) == declared_symbols()


def test_version_string_without_gpu():
    import scgpu

    assert "gfx950" in scgpu.version()


def test_default_config_mirrors_reference_default():
    """LeidenConfig::default (leiden/mod.rs:92-106); a pure host call."""
    import scgpu

    c = scgpu.lib().sc_leiden_config_default()
    assert (c.objective, c.resolution, c.has_seed, c.seed) == (0, 1.0, 1, 42)
    assert (c.n_iterations, c.max_levels, c.tolerance) == (2, 64, 1e-9)
    assert (c.refine, c.refine_randomness, c.patience, c.has_max_community_weight) == (1, 0.01, 2, 0)


def test_entry_count_guard_before_device_work():
    """Graphs of 2^32 - 1 or more stored entries are refused on the host (entry positions and
    scans over entries are 32-bit on the device); no array is read and no GPU is touched."""
    import ctypes as C

    import numpy as np
    import scgpu

    L = scgpu.lib()
    ptrs = np.zeros(2, np.uint64)
    nbrs = np.zeros(1, np.uint32)
    wts = np.zeros(1, np.float32)
    for nnz in (1 << 32, (1 << 32) - 1):
        h = C.c_void_p()
        err = scgpu.ScError()
        st = L.sc_graph_from_csr_parts(ptrs.ctypes.data_as(scgpu.u64p), 1, nbrs.ctypes.data_as(scgpu.u32p),
                                       wts.ctypes.data_as(scgpu.f32p), nnz, None, C.byref(h), C.byref(err))
        assert st == 101 and err.code == 101 and not h.value
        assert err.got == nnz and err.expected == (1 << 32) - 2
        assert b"2^32 - 2" in err.msg
    # kNN graphs hold 2 n k entries before merging: refused before the search
    x = np.zeros((1, 1), np.float32)
    h = C.c_void_p()
    err = scgpu.ScError()
    st = L.sc_knn_graph(x.ctypes.data_as(scgpu.f32p), 1 << 30, 1, 4, None, C.byref(h), None, C.byref(err))
    assert st == 101 and err.got == 1 << 33
    # and more points than the device's node limit (2^31 - 1): TooManyNodes
    err = scgpu.ScError()
    st = L.sc_knn_graph(x.ctypes.data_as(scgpu.f32p), 1 << 31, 1, 1, None, C.byref(h), None, C.byref(err))
    assert st == 8 and err.n_nodes == 1 << 31


def test_from_csr_matrix_checks_row_offsets_on_the_host():
    """A matrix whose row offsets a CsrMatrix could not hold is InvalidCsr before any device work."""
    import ctypes as C

    import numpy as np
    import scgpu

    ptrs = np.array([0, 2, 1], np.uint64)
    cols = np.array([1, 0], np.uint64)
    vals = np.array([1.0, 1.0])
    h, err = C.c_void_p(), scgpu.ScError()
    st = scgpu.lib().sc_graph_from_csr_matrix(2, ptrs.ctypes.data_as(scgpu.u64p), cols.ctypes.data_as(scgpu.u64p),
                                              vals.ctypes.data_as(scgpu.f64p), 2, None, C.byref(h), C.byref(err))
    assert st == 9 and not h.value
