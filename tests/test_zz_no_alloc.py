"""The allocation-free quality entry point (sc_quality_noalloc) against sc_quality through the C
ABI: bit-identical results on two calls in a row (the second on a smaller graph, so stale
per-community totals of the first call would show), device allocator untouched, and the label
limit reported as an error."""
import ctypes as C

import numpy as np
import pytest

import scgpu
import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(g, labels, objective, resolution):
    d = scgpu.DeviceArray(np.ascontiguousarray(labels, dtype=np.uint64))
    try:
        out = C.c_double()
        err = scgpu.ScError()
        st = scgpu.lib().sc_quality_noalloc(g.handle, C.c_void_p(d.ptr), objective, resolution, C.byref(out),
                                            C.byref(err))
        return st, out.value, err
    finally:
        d.free()


def _reference(g, labels, objective, resolution):
    obj = scgpu.ObjectiveKind.Rb(resolution) if objective == 0 else scgpu.ObjectiveKind.Cpm(resolution)
    return scgpu.quality(g, labels, obj)


def test_noalloc_quality_equals_sc_quality_twice_then_smaller():
    x, _ = W.gaussian_mixture(20000, 16, 12, 9.0, seed=4)
    big = scgpu.knn_graph(x, 8)
    big_labels = scgpu.leiden(big, scgpu.LeidenConfig()).labels()
    small = scgpu.CSRNetwork.from_edges(34, W.karate_edges())
    small_labels = scgpu.leiden(small, scgpu.LeidenConfig()).labels()
    rng = np.random.default_rng(0)
    cases = [
        (big, big_labels),                                            # first call, many communities
        (big, rng.integers(0, 5000, big.node_count()).astype(np.uint64)),
        (small, small_labels),                                        # smaller input after a larger one
        (small, np.arange(34, dtype=np.uint64)),
    ]
    live0 = scgpu.pool_bytes()
    for g, labels in cases:
        for objective, resolution in [(0, 1.0), (0, 0.3), (1, 0.05)]:
            st, q, err = _noalloc(g, labels, objective, resolution)
            assert st == 0, err.msg
            ref = _reference(g, labels, objective, resolution)
            assert q == ref, (g.node_count(), objective, resolution, q, ref)   # bit for bit
    # the library's allocator was not used by the allocation-free calls (only by sc_quality,
    # whose blocks return to the cache)
    assert scgpu.pool_bytes()[0] == live0[0]


def test_noalloc_rejects_labels_beyond_its_limit_and_recovers():
    g = scgpu.CSRNetwork.from_edges(34, W.karate_edges())
    labels = np.arange(34, dtype=np.uint64)
    labels[5] = 1 << 20
    st, _, err = _noalloc(g, labels, 0, 1.0)
    assert st == 101 and err.got == 1 << 20 and b"2^20" in err.msg
    labels[5] = (1 << 20) - 1                       # the largest label allowed
    st, q, err = _noalloc(g, labels, 1, 0.05)
    assert st == 0 and q == _reference(g, labels, 1, 0.05)
    ok = scgpu.leiden(g, scgpu.LeidenConfig()).labels()
    st, q, _ = _noalloc(g, ok, 0, 1.0)
    assert st == 0 and q == _reference(g, ok, 0, 1.0)
    # empty graph and null arguments
    e = scgpu.CSRNetwork.from_edges(0, [])
    out, err = C.c_double(7.0), scgpu.ScError()
    assert scgpu.lib().sc_quality_noalloc(e.handle, None, 0, 1.0, C.byref(out), C.byref(err)) == 0 and out.value == 0.0
    assert scgpu.lib().sc_quality_noalloc(g.handle, None, 0, 1.0, C.byref(out), C.byref(err)) == 101
    assert scgpu.lib().sc_quality_noalloc(g.handle, None, 7, 1.0, C.byref(out), C.byref(err)) == 101
