"""GPU parity: ingest, aggregation, modularity/quality and the connectivity check through the
C ABI, against the CPU oracle (bit-exact for integer/structure work, f32 weights bit-exact,
f64 quality values within 1e-12 relative)."""
import numpy as np
import pytest

import pyoracle as O
import scgpu
import workloads as W
from helpers import load_fixture, reference_test_graphs

pytestmark = pytest.mark.gpu


def both_from_edges(n, edges):
    return scgpu.CSRNetwork.from_edges(n, edges), O.Graph.from_edges(n, edges)


def assert_same_graph(g, og):
    gp, gn, gw = g.to_csr_parts()
    op, on, ow = og.csr()
    assert np.array_equal(gp, op)
    assert np.array_equal(gn, on)
    assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32))  # f32 bit-exact
    assert np.array_equal(g.strengths(), og.strengths())          # row-ordered f64 sums
    assert np.array_equal(g.node_weights(), og.node_weights())
    assert g.edge_count() == og.edge_count()
    assert g.has_self_loops() == og.has_self_loops()
    assert g.total_weight() == pytest.approx(og.total_weight(), rel=1e-12, abs=0)


def test_reference_kats_through_the_abi():
    g = scgpu.CSRNetwork.from_edges(3, [(0, 1, 1.0), (1, 2, 1.0), (0, 2, 1.0), (0, 0, 3.0)])
    assert list(g.strengths()) == [8.0, 2.0, 2.0] and g.total_weight() == 6.0
    g = scgpu.CSRNetwork.from_edges(2, [(0, 1, 1.0), (0, 1, 2.0)])
    assert g.total_weight() == 3.0 and g.edge_count() == 1
    g = scgpu.CSRNetwork.from_csr_parts([0, 3, 4], [1, 1, 0, 0], [1.0, 2.0, 0.0, 3.0])
    p, n, w = g.to_csr_parts()
    assert list(p) == [0, 1, 2] and list(n) == [1, 0] and list(w) == [3.0, 3.0]


@pytest.mark.parametrize("args,variant", [
    (([0, 5], [0], [1.0]), "InvalidCsr"),
    (([0, 1], [9], [1.0]), "NodeIndexOutOfRange"),
    (([0, 2, 1], [0, 1], [1.0, 1.0]), "InvalidCsr"),
    (([0, 1], [0], [-1.0]), "NegativeWeight"),
    (([0, 1], [0], [float("nan")]), "NonFiniteWeight"),
    (([0, 1, 1], [1], [1.0]), "AsymmetricGraph"),
    (([0, 1, 2], [1, 0], [1.0, 5.0]), "AsymmetricGraph"),
])
def test_malformed_input_is_rejected(args, variant):
    with pytest.raises(scgpu.ClusteringError) as e:
        scgpu.CSRNetwork.from_csr_parts(*args)
    assert e.value.variant == variant
    if variant == "NodeIndexOutOfRange":
        assert (e.value.node, e.value.n_nodes) == (9, 1)
    if variant == "AsymmetricGraph":
        assert "symmetrised" in e.value.message


def test_validate_symmetry():
    g = scgpu.CSRNetwork.from_csr_parts([0, 2, 3, 4], [1, 2, 0, 0], [1.0, 3.0, 3.0, 1.0])
    with pytest.raises(scgpu.ClusteringError):
        g.validate_symmetry()
    scgpu.CSRNetwork.from_edges(3, [(0, 1, 1.0), (1, 2, 1.0), (0, 0, 2.0)]).validate_symmetry()


def test_ingest_matches_oracle_on_reference_graphs():
    for name, n, edges in reference_test_graphs():
        g, og = both_from_edges(n, edges)
        assert_same_graph(g, og)


def test_ingest_unsorted_duplicates_zeros_random():
    rng = np.random.default_rng(11)
    for trial in range(10):
        n = int(rng.integers(1, 400))
        m = int(rng.integers(0, 4 * n))
        a = rng.integers(0, n, size=m)
        b = rng.integers(0, n, size=m)
        w = np.where(rng.random(m) < 0.1, 0.0, rng.random(m) * 3).astype(np.float32)
        # symmetric entry list with duplicates and zeros, rows shuffled
        rows = np.concatenate([a, b])
        cols = np.concatenate([b, a])
        vals = np.concatenate([w, w])
        perm = rng.permutation(len(rows))
        rows, cols, vals = rows[perm], cols[perm], vals[perm]
        order = np.argsort(rows, kind="stable")
        rows, cols, vals = rows[order], cols[order], vals[order]
        ptrs = np.zeros(n + 1, np.uint64)
        np.add.at(ptrs, rows + 1, 1)
        ptrs = np.cumsum(ptrs).astype(np.uint64)
        try:
            og = O.Graph.from_csr_parts(ptrs, cols, vals)
        except O.OracleError as e:
            with pytest.raises(scgpu.ClusteringError) as ge:
                scgpu.CSRNetwork.from_csr_parts(ptrs, cols, vals)
            assert ge.value.code == e.code
            continue
        g = scgpu.CSRNetwork.from_csr_parts(ptrs, cols, vals)
        assert_same_graph(g, og)


def test_aggregate_bit_exact_vs_oracle():
    rng = np.random.default_rng(3)
    cases = list(reference_test_graphs())
    for i in range(8):
        n = int(rng.integers(20, 300))
        edges = [(int(a), int(b), float(0.05 + 4 * rng.random())) for a, b in
                 zip(rng.integers(0, n, 3 * n), rng.integers(0, n, 3 * n))]
        cases.append((f"random{i}", n, edges))
    for name, n, edges in cases:
        if n == 0:
            continue
        g, og = both_from_edges(n, edges)
        for k in (1, max(1, n // 7), max(1, n // 2), n):
            groups = rng.integers(0, k, size=n)
            _, groups = np.unique(groups, return_inverse=True)
            assert_same_graph(g.aggregate(groups), og.aggregate(groups))


def _csr_from_directed(n, rows, cols, vals):
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ptrs = np.zeros(n + 1, np.uint64)
    np.add.at(ptrs, rows + 1, 1)
    return np.cumsum(ptrs).astype(np.uint64), cols.astype(np.uint32), vals.astype(np.float32)


def test_aggregate_long_pairs_and_wide_exponents():
    # few large groups: pairs with thousands of fine entries (the wave-parallel sums), weights
    # spread over eleven decades so that many pair sums fail the exactness bound and take the
    # sequential chain; both must equal the reference's order bit for bit
    rng = np.random.default_rng(21)
    n = 3000
    m = 40000
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    w = (10.0 ** rng.uniform(-8, 3, m)).astype(np.float32)
    loops = rng.integers(0, n, 50)
    rows = np.concatenate([a, b, loops])
    cols = np.concatenate([b, a, loops])
    vals = np.concatenate([w, w, np.full(50, 0.75, np.float32)])
    ptrs, nbrs, wts = _csr_from_directed(n, rows, cols, vals)
    g = scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts)
    og = O.Graph.from_csr_parts(ptrs, nbrs, wts)
    for k in (1, 3, 17, 400):
        groups = rng.integers(0, k, size=n)
        _, groups = np.unique(groups, return_inverse=True)
        assert_same_graph(g.aggregate(groups), og.aggregate(groups))


def test_aggregate_balanced_asymmetric_weights():
    # the two directions of a pair carry different weights (a permutation, so the degree sum
    # still balances and ingest accepts it): coarse pair weights come from the lower group's
    # side only (csr_network.rs:548-581), which a symmetric shortcut would get wrong
    rng = np.random.default_rng(8)
    n = 500
    a = rng.integers(0, n, 3000)
    b = rng.integers(0, n, 3000)
    keep = a < b
    a, b = a[keep], b[keep]
    pairs = np.unique(np.stack([a, b], 1), axis=0)
    a, b = pairs[:, 0], pairs[:, 1]
    w_up = rng.integers(1, 9, len(a)).astype(np.float32)
    w_dn = rng.permutation(w_up)
    ptrs, nbrs, wts = _csr_from_directed(n, np.concatenate([a, b]), np.concatenate([b, a]),
                                         np.concatenate([w_up, w_dn]))
    g = scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts)
    og = O.Graph.from_csr_parts(ptrs, nbrs, wts)
    assert not og.validate_symmetry()
    for k in (2, 9, 60, 250):
        groups = rng.integers(0, k, size=n)
        _, groups = np.unique(groups, return_inverse=True)
        assert_same_graph(g.aggregate(groups), og.aggregate(groups))


def test_modularity_and_quality_on_fixture_memberships():
    graphs, cases = load_fixture()
    built = {k: (scgpu.CSRNetwork.from_edges(n, e), O.Graph.from_edges(n, e)) for k, (n, e) in graphs.items()}
    for c in cases:
        g, og = built[c["graph"]]
        memb = np.asarray(c["reference_membership"], dtype=np.uint64)
        m_gpu = scgpu.modularity(g, memb, 1.0)
        assert m_gpu == pytest.approx(O.modularity(og, memb, 1.0), rel=1e-12, abs=1e-15)
        assert abs(m_gpu - c["reference_modularity"]) <= 1e-7
        obj = scgpu.ObjectiveKind(c["objective"], c["resolution"])
        q_gpu = scgpu.quality(g, memb, obj)
        assert q_gpu == pytest.approx(O.quality(og, memb, c["objective"], c["resolution"]), rel=1e-12, abs=1e-12)
        assert abs(2 * q_gpu - c["reference_quality"]) <= 1e-6 * max(abs(c["reference_quality"]), 1.0)


def test_disconnected_count_matches_oracle():
    rng = np.random.default_rng(8)
    for name, n, edges in reference_test_graphs():
        if n == 0:
            continue
        g, og = both_from_edges(n, edges)
        for k in (1, 2, 3, max(1, n // 4)):
            labels = rng.integers(0, k, size=n).astype(np.uint64)
            assert scgpu.disconnected_community_count(g, labels) == O.disconnected_community_count(og, labels)
    ptrs, nbrs, wts, truth = W.planted_partition(n=3000, blocks=10, p_in=0.01, p_out=0.0005, seed=2)
    g = scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts)
    og = O.Graph.from_csr_parts(ptrs, nbrs, wts)
    for k in (2, 10, 100):
        labels = rng.integers(0, k, size=3000).astype(np.uint64)
        assert scgpu.disconnected_community_count(g, labels) == O.disconnected_community_count(og, labels)


def test_csr_export_reader_and_real_data_driver(tmp_path):
    # a reference-compatible kNN connectivities matrix (median sigma, lower-row emission, as a
    # scanpy export would hold) written in the export layout, read back through the library
    import json
    import time

    import real_data as RD

    x, truth = W.gaussian_mixture(3000, 20, 6, 12.0, seed=11)
    oi, od = O.knn_bruteforce(x, 10)
    p, nb, w = O.knn_gaussian(oi, od, mode=1)
    scgpu.write_csr_export(str(tmp_path), p, nb, w)
    g = scgpu.CSRNetwork.from_csr_export(str(tmp_path))
    g0 = scgpu.CSRNetwork.from_csr_parts(p, nb, w)
    for a, b in zip(g.to_csr_parts(), g0.to_csr_parts()):
        assert np.array_equal(a, b)
    assert g.total_weight() == g0.total_weight()
    # the driver: the stand-in for scanpy's labels is the CPU oracle's Leiden per resolution
    og = O.Graph.from_csr_parts(p, nb, w)
    ref = {"resolutions": {}, "annotation": {"name": "component", "labels": [int(t) for t in truth],
                                              "categories": [str(i) for i in range(6)]}}
    theirs = {}
    for key in ["0.5", "1.0", "2.0"]:
        t0 = time.perf_counter()
        lab, nc = O.leiden(og, resolution=float(key))
        theirs[key] = lab
        ref["resolutions"][key] = {"labels": [int(v) for v in lab], "n_clusters": int(nc),
                                   "seconds": time.perf_counter() - t0}
    (tmp_path / "reference.json").write_text(json.dumps(ref))
    (tmp_path / "meta.json").write_text(json.dumps({"dataset": "synthetic mixture", "n_cells": 3000,
                                                    "n_entries": int(len(nb)), "n_neighbors": 10,
                                                    "scanpy": "none"}))
    lines = []
    rows = RD.run(str(tmp_path), seed=42, out=lines.append)
    assert [r["resolution"] for r in rows] == [0.5, 1.0, 2.0]
    for r, key in zip(rows, ["0.5", "1.0", "2.0"]):
        # modularity of the stand-in labels on the device graph equals the oracle's (1e-9 rel)
        assert r["scanpy_q"] == pytest.approx(O.modularity(og, theirs[key], r["resolution"]), rel=1e-9)
        assert r["our_q"] >= r["scanpy_q"] * (1 - 0.005)
        assert 0.0 <= r["nmi"] <= 1.0 + 1e-12 and r["ours_k"] > 0
    assert rows[1]["ari"] > 0.9  # six well-separated components at resolution 1
    assert lines[2] == "exhaustive symmetry check: OK" and lines[-1].startswith("best agreement")


def test_graph_freed_on_another_thread_returns_its_memory():
    """A handle built on one thread and freed on another: its blocks leave the building thread's
    allocator (live bytes back to where they were) and are not cached there (they go back to the
    device), so nothing leaks into a pool that may never run again."""
    import threading

    x, _ = W.gaussian_mixture(3000, 8, 5, 9.0, seed=1)
    box = {}
    built, freed = threading.Event(), threading.Event()

    def worker():
        box["before"] = scgpu.pool_bytes()
        box["g"] = scgpu.knn_graph(x, 6)
        box["with_graph"] = scgpu.pool_bytes()
        built.set()
        freed.wait(60)
        box["after"] = scgpu.pool_bytes()

    t = threading.Thread(target=worker)
    t.start()
    assert built.wait(120)
    g = box.pop("g")
    assert g.node_count() == 3000
    del g                                  # sc_graph_free on the main thread
    freed.set()
    t.join(60)
    (live0, _), (live1, cached1), (live2, cached2) = box["before"], box["with_graph"], box["after"]
    assert live1 > live0                    # the graph's arrays were live in the worker's pool
    assert live2 == live0                   # released from the worker's pool ...
    assert cached2 == cached1               # ... and not cached there
    # the main thread's own pool is untouched by the foreign release
    m_live, m_cached = scgpu.pool_bytes()
    g2 = scgpu.knn_graph(x, 6)
    del g2
    assert scgpu.pool_bytes()[0] == m_live


def test_from_csr_matrix_reads_the_upper_triangle():
    """from_csr_matrix (csr_network.rs:360-391): entries with row <= col, zeros skipped, the
    diagonal a self-loop, then from_edges_with_node_weights -- bit-identical to from_edges of
    that list and to the oracle's; to_csr_matrix (:610-618) returns the stored adjacency widened."""
    import scipy.sparse as sp

    rng = np.random.default_rng(21)
    for trial in range(6):
        n = int(rng.integers(1, 300))
        dense = sp.random(n, n, density=min(1.0, 6.0 / n), random_state=trial, format="coo")
        vals = np.where(rng.random(dense.nnz) < 0.1, 0.0, rng.random(dense.nnz) * 3)
        upper = sp.coo_matrix((vals, (dense.row, dense.col)), shape=(n, n))
        # a symmetric matrix, and (trial odd) an unsymmetric one: only the upper part counts
        m = (upper + upper.T).tocsr() if trial % 2 == 0 else upper.tocsr()
        m.sort_indices()
        nw = rng.random(n) + 0.5 if trial == 3 else None
        g = scgpu.CSRNetwork.from_csr_matrix(m, node_weights=nw)
        coo = m.tocoo()
        keep = (coo.row <= coo.col) & (coo.data != 0)
        order = np.lexsort((coo.col[keep], coo.row[keep]))  # triplet order: rows, then columns
        edges = [(int(a), int(b), float(w)) for a, b, w in
                 zip(coo.row[keep][order], coo.col[keep][order], coo.data[keep][order])]
        ref = scgpu.CSRNetwork.from_edges(n, edges, node_weights=nw)
        for x, y in zip(g.to_csr_parts(), ref.to_csr_parts()):
            assert np.array_equal(x, y)
        assert np.array_equal(g.strengths(), ref.strengths())
        assert_same_graph(g, O.Graph.from_edges(n, edges, node_weights=nw))
        p, c, v = g.to_csr_matrix()
        p2, c2, v2 = g.to_csr_parts()
        assert np.array_equal(p, p2) and np.array_equal(c, c2.astype(np.uint64))
        assert np.array_equal(v, v2.astype(np.float64))


@pytest.mark.parametrize("ptrs,cols,vals,variant", [
    ([0, 1, 2], [1, 5], [1.0, 1.0], "NodeIndexOutOfRange"),
    ([0, 1, 2], [1, 0], [-1.0, -1.0], "NegativeWeight"),
    ([0, 2, 1], [1, 0], [1.0, 1.0], "InvalidCsr"),
    ([1, 1, 2], [1, 0], [1.0, 1.0], "InvalidCsr"),
])
def test_from_csr_matrix_rejects_malformed_input(ptrs, cols, vals, variant):
    with pytest.raises(scgpu.ClusteringError) as e:
        scgpu.CSRNetwork.from_csr_matrix((ptrs, cols, vals))
    assert e.value.variant == variant
    with pytest.raises(scgpu.ClusteringError) as e:
        scgpu.CSRNetwork.from_csr_matrix(([0, 1, 2], [1, 0], [1.0, 1.0]), node_weights=[1.0])
    assert e.value.variant == "NodeWeightLengthMismatch"
