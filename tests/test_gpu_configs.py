"""The benchmark configurations on the HIP path, checked against the oracle at their full sizes.

- K: 100,000 x 50 points (30-component mixture), k = 15: every kNN row (indices and FP64 squared
  distances) bit-exact against the kd-tree restatement of the reference's exact path
  (neighborhood/mod.rs:233-314), the BF16 and FP32 MFMA paths identical, the whole Gaussian CSR
  against the oracle's, Leiden quality within 0.5% of the oracle at the same seed and no
  disconnected community by the oracle's count.
- L: 4,000,000 x 30 points (200 components), k = 7: 1,000 sampled rows against the kd-tree (the
  reference would use HNSW here, neighborhood/mod.rs:608-624, whose recall cannot be measured
  offline), Rb 1.0 and CPM 0.05 Leiden against the oracle (quality within 0.5%, connectivity).
- F: 8,000,000 x 30 points (400 components), k = 7: sampled rows against the kd-tree, and the
  properties that need no full CPU run: connectivity by the oracle's count, quality invariant
  under aggregation, quality values equal to the oracle's f64 recomputation within 1e-9
  relative, and learned-filter Leiden within 0.1% of plain modularity.
- B: 64 LFR-style graphs x 8 Rb resolutions: every batched job equal to its standalone call, and
  16 jobs against the oracle (quality within 0.5%, connectivity).
"""
import os
import threading

import numpy as np
import pytest

import pyoracle as O
import scgpu
import workloads as W

pytestmark = pytest.mark.gpu


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def _oracle_graph(g):
    return O.Graph.from_csr_parts(*g.to_csr_parts())


def _oracle_leiden_concurrently(og, runs):
    """Runs several single-threaded oracle leiden() calls side by side (ctypes releases the GIL);
    runs: list of keyword dicts. Returns their labels in order."""
    out = [None] * len(runs)

    def work(i):
        out[i] = O.leiden(og, **runs[i])[0]

    ts = [threading.Thread(target=work, args=(i,)) for i in range(len(runs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return out


def _check_rows(x, k, rows, idx, d2):
    ki, kd = O.knn_kdtree(x, k, rows=rows.astype(np.uint64), threads=_threads())
    r = rows.astype(np.int64)
    assert np.array_equal(idx[r], ki)
    assert np.array_equal(d2[r].view(np.uint64), kd.view(np.uint64))   # FP64 bit-exact


def _check_leiden(g, og, cfg, labels_oracle, label):
    kind, res = cfg.objective.kind, cfg.objective.resolution
    c = scgpu.leiden(g, cfg)
    qg, qo = O.quality(og, c.labels(), kind, res), O.quality(og, labels_oracle, kind, res)
    ari = O.ari(c.labels(), labels_oracle)
    print(f"{label}: GPU {kind} quality {qg:.6f} ({c.n_clusters()}), oracle {qo:.6f}, ARI {ari:.4f}")
    assert qg >= qo - 0.005 * abs(qo), (label, qg, qo)
    assert O.disconnected_community_count(og, c.labels()) == 0
    # the GPU's own quality value equals the f64 recomputation on the CPU
    assert scgpu.quality(g, c.labels(), cfg.objective) == pytest.approx(qg, rel=1e-9, abs=1e-12)
    return c


# ---- K ------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config_k():
    x, _ = W.gaussian_mixture(100_000, 50, 30, 16.0, sd=1.0, alpha=2.0, seed=42)
    return x, scgpu.knn_graph(x, 15)


def test_config_k_knn_every_row_and_gaussian_csr(config_k):
    x, g = config_k
    idx, d2 = scgpu.knn(x, 15)
    ki, kd = O.knn_kdtree(x, 15, threads=_threads())
    assert np.array_equal(idx, ki)
    assert np.array_equal(d2.view(np.uint64), kd.view(np.uint64))
    fi, fd = scgpu.knn(x, 15, precision="fp32")
    assert np.array_equal(fi, idx) and np.array_equal(fd.view(np.uint64), d2.view(np.uint64))
    op, on, ow = O.knn_gaussian(ki, kd, mode=0)
    p, nb, w = g.to_csr_parts()
    assert np.array_equal(p, op) and np.array_equal(nb, on)
    assert np.allclose(w, ow, rtol=1e-6, atol=0)
    g.validate_symmetry()


def test_config_k_leiden_vs_oracle(config_k):
    _, g = config_k
    og = _oracle_graph(g)
    cfg = scgpu.LeidenConfig(seed=42)
    (ol,) = _oracle_leiden_concurrently(og, [dict(seed=42)])
    _check_leiden(g, og, cfg, ol, "K Rb 1.0")


# ---- L ------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config_l():
    x, _ = W.gaussian_mixture(4_000_000, 30, 200, 9.0, sd=1.0, alpha=1.0, seed=42)
    return x, scgpu.knn_graph(x, 7)


def test_config_l_knn_sampled_rows(config_l):
    x, g = config_l
    idx, d2 = scgpu.knn(x, 7)
    rows = np.random.default_rng(1).choice(len(x), 1000, replace=False)
    _check_rows(x, 7, rows, idx, d2)
    # every sampled row's neighbours are its row of the symmetric graph
    p, nb, _ = g.to_csr_parts()
    for r in rows[:200]:
        assert set(idx[r].tolist()) <= set(nb[p[r]:p[r + 1]].tolist())
    assert 4_000_000 * 7 // 2 <= g.edge_count() <= 4_000_000 * 7


def test_config_l_leiden_rb_and_cpm_vs_oracle(config_l):
    _, g = config_l
    og = _oracle_graph(g)
    rb = scgpu.LeidenConfig(seed=42)
    cpm = scgpu.LeidenConfig(seed=42, objective=scgpu.ObjectiveKind.Cpm(0.05))
    ol_rb, ol_cpm = _oracle_leiden_concurrently(og, [dict(seed=42), dict(seed=42, objective="cpm", resolution=0.05)])
    _check_leiden(g, og, rb, ol_rb, "L Rb 1.0")
    _check_leiden(g, og, cpm, ol_cpm, "L CPM 0.05")


# ---- F ------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config_f():
    x, _ = W.gaussian_mixture(8_000_000, 30, 400, 9.0, sd=1.0, alpha=1.0, seed=42)
    return x, scgpu.knn_graph(x, 7)


def test_config_f_knn_sampled_rows(config_f):
    x, g = config_f
    idx, d2 = scgpu.knn(x, 7)
    rows = np.random.default_rng(2).choice(len(x), 300, replace=False)
    _check_rows(x, 7, rows, idx, d2)
    assert 8_000_000 * 7 // 2 <= g.edge_count() <= 8_000_000 * 7


def test_config_f_leiden_properties(config_f):
    _, g = config_f
    og = _oracle_graph(g)
    c = scgpu.leiden(g, scgpu.LeidenConfig(seed=42))
    labels = c.labels()
    assert O.disconnected_community_count(og, labels) == 0
    q = scgpu.modularity(g, labels, 1.0)
    assert q == pytest.approx(O.modularity(og, labels, 1.0), rel=1e-9, abs=0)
    coarse = g.aggregate(labels)
    cl = np.arange(coarse.node_count(), dtype=np.uint64)
    # the coarse graph stores its summed weights as f32 (csr_network.rs:566-581), so the
    # invariance holds to f32 rounding: the reference's bound, 1e-6 relative (invariants.rs)
    for obj in (scgpu.ObjectiveKind.Rb(1.0), scgpu.ObjectiveKind.Rb(0.25), scgpu.ObjectiveKind.Cpm(0.05)):
        a, b = scgpu.quality(g, labels, obj), scgpu.quality(coarse, cl, obj)
        assert abs(a - b) <= 1e-6 * max(abs(a), 1.0), (obj, a, b)
    print(f"F: {g.node_count()} nodes, {g.edge_count()} edges, {c.n_clusters()} clusters, Q {q:.10f}")


def test_config_f_learned_filter_keeps_modularity(config_f):
    _, g = config_f
    cfg = scgpu.LeidenConfig(seed=42)
    Xs, ys = [], []
    for sd in (101, 102):
        xt, _ = W.gaussian_mixture(300_000, 30, 200, 9.0, sd=1.0, alpha=1.0, seed=sd)
        X, y = scgpu.leiden_trace(scgpu.knn_graph(xt, 7), cfg)
        Xs.append(X)
        ys.append(y)
    w, _ = scgpu.mlp_train(np.concatenate(Xs), np.concatenate(ys), epochs=6, batch=8192, lr=3e-3, seed=7)
    q = scgpu.modularity(g, scgpu.leiden(g, cfg).labels(), 1.0)
    for judge in ("full", "skipped"):
        cf, st = scgpu.leiden_filtered(g, w, 0.05, cfg, judge=judge)
        qf = scgpu.modularity(g, cf.labels(), 1.0)
        print(f"F filter judge={judge}: Q {qf:.10f} vs plain {q:.10f}, skipped {st['skipped']}, "
              f"skip-error rate {st['skip_error_rate']:.5f}")
        assert qf >= q * (1 - 0.001), (judge, qf, q)


# ---- B ------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config_b():
    sizes = W.batch_graph_sizes(64, 2000, 200000, seed=42)
    parts = [W.lfr_like(int(n), 1000 + i) for i, n in enumerate(sizes)]
    graphs = [scgpu.CSRNetwork.from_csr_parts(*p[:3]) for p in parts]
    return parts, graphs, np.linspace(0.25, 2.0, 8)


def test_config_b_batch_equals_standalone(config_b):
    _, graphs, res = config_b
    cfg = scgpu.LeidenConfig(seed=42)
    jobs = scgpu.leiden_batch(graphs, res, cfg)
    assert len(jobs) == 512
    for gi, g in enumerate(graphs):
        for ri, r in enumerate(res):
            single = scgpu.leiden(g, scgpu.LeidenConfig(seed=42, objective=scgpu.ObjectiveKind.Rb(float(r))))
            assert np.array_equal(jobs[gi * len(res) + ri].labels(), single.labels()), (gi, r)


def test_config_b_jobs_vs_oracle(config_b):
    parts, graphs, res = config_b
    cfg = scgpu.LeidenConfig(seed=42)
    picks = [(gi, gi % len(res)) for gi in range(16)]            # 16 graphs, every resolution twice
    jobs = scgpu.leiden_batch([graphs[gi] for gi, _ in picks], res, cfg)
    ogs = [O.Graph.from_csr_parts(*parts[gi][:3]) for gi, _ in picks]
    for j, (gi, ri) in enumerate(picks):
        labels = jobs[j * len(res) + ri].labels()
        r = float(res[ri])
        ol, _ = O.leiden(ogs[j], seed=42, resolution=r)
        qg, qo = O.quality(ogs[j], labels, "rb", r), O.quality(ogs[j], ol, "rb", r)
        assert qg >= qo - 0.005 * abs(qo), (gi, r, qg, qo)
        assert O.disconnected_community_count(ogs[j], labels) == 0, (gi, r)
