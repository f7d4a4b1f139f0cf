"""The reference's property and known-answer suites (tests/properties.rs, tests/known_answers.rs)
run on the HIP path. Random graphs and configs follow random_graph / random_config
(properties.rs:21-95) in distribution, drawn with numpy's generator instead of ChaCha8; every
invariant the reference asserts is asserted here on the GPU's results, with the oracle as the
CPU recomputation of qualities and of the connectivity count."""
from dataclasses import replace

import numpy as np
import pytest

import pyoracle as O
import scgpu
import workloads as W

pytestmark = pytest.mark.gpu


def random_graph(seed):
    """properties.rs:21-61 in distribution: n in 1..=120, density u^2 * 0.6 + 0.01, planted
    blocks (2..=6) half the time with 4x density inside, weights 0.05 + 4u when weighted, and
    self-loops on 20% of nodes for 30% of graphs."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 121))
    density = rng.random() ** 2 * 0.6 + 0.01
    weighted = bool(rng.random() < 0.5)
    self_loops = bool(rng.random() < 0.3)
    blocks = min(int(rng.integers(2, 7)), n) if rng.random() < 0.5 else 1
    edges = []
    for a in range(n):
        for b in range(a + 1, n):
            same = blocks > 1 and (a * blocks // n) == (b * blocks // n)
            p = density * 4.0 if same else density
            if rng.random() < min(p, 1.0):
                edges.append((a, b, 0.05 + 4.0 * rng.random() if weighted else 1.0))
    if self_loops:
        for v in range(n):
            if rng.random() < 0.2:
                edges.append((v, v, 0.5 + 2.0 * rng.random()))
    name = f"seed={seed} n={n} density={density:.3f} blocks={blocks} weighted={weighted} loops={self_loops}"
    return name, n, edges


def random_config(seed, salt):
    """properties.rs:63-87 in distribution."""
    rng = np.random.default_rng(seed ^ salt)
    if rng.random() < 0.5:
        obj = scgpu.ObjectiveKind.Rb(float(rng.choice([0.1, 0.5, 1.0, 2.0, 5.0])))
    else:
        obj = scgpu.ObjectiveKind.Cpm(float(rng.choice([0.01, 0.05, 0.2, 1.0])))
    cap = float(rng.integers(2, 41)) if rng.random() < 0.2 else None
    return scgpu.LeidenConfig(objective=obj, seed=seed, n_iterations=int(rng.integers(1, 5)),
                              refine=bool(rng.random() < 0.5),
                              refine_randomness=float(rng.choice([0.0, 0.01, 0.1])),
                              max_community_weight=cap)


def test_invariants_hold_on_random_inputs():
    """properties.rs:97-150 over 400 random graph/config pairs."""
    for seed in range(400):
        name, n, edges = random_graph(seed)
        g = scgpu.CSRNetwork.from_edges(n, edges)
        og = O.Graph.from_edges(n, edges)
        cfg = random_config(seed, 0x9E3779B9)
        c = scgpu.leiden(g, cfg)
        labels = c.labels()
        ctx = f"{name} {cfg.objective}"
        assert len(labels) == n, ctx
        assert all(l < c.n_clusters() for l in labels.tolist()), ctx
        assert (c.cluster_sizes() > 0).all(), ctx
        kind, res = cfg.objective.kind, cfg.objective.resolution
        start = O.quality(og, np.arange(n), kind, res)
        end = O.quality(og, labels, kind, res)
        assert end >= start - 1e-6 * max(abs(start), 1.0), ctx
        if cfg.refine:
            assert O.disconnected_community_count(og, labels) == 0, ctx
        if cfg.max_community_weight is not None:
            assert (c.cluster_sizes() <= cfg.max_community_weight).all(), ctx
        assert scgpu.leiden(g, cfg) == c, f"{ctx}: not reproducible"


def test_aggregation_preserves_quality_on_random_inputs():
    """properties.rs:153-192: GPU leiden, renumber, GPU aggregate; strength sum, symmetry and
    Rb/CPM quality survive the collapse."""
    for seed in range(250):
        name, n, edges = random_graph(seed)
        if n == 0:
            continue
        g = scgpu.CSRNetwork.from_edges(n, edges)
        labels = scgpu.leiden(g, scgpu.LeidenConfig()).labels()
        groups, k = scgpu.renumber(labels)
        coarse = g.aggregate(groups)
        assert coarse.node_count() == k
        fine_deg, coarse_deg = g.strengths().sum(), coarse.strengths().sum()
        assert abs(fine_deg - coarse_deg) < 1e-6 * max(abs(fine_deg), 1.0), name
        coarse.validate_symmetry()
        cl = np.arange(k, dtype=np.uint64)
        for obj in (scgpu.ObjectiveKind.Rb(0.25), scgpu.ObjectiveKind.Rb(1.0), scgpu.ObjectiveKind.Rb(3.0),
                    scgpu.ObjectiveKind.Cpm(0.05)):
            a, b = scgpu.quality(g, groups, obj), scgpu.quality(coarse, cl, obj)
            assert abs(a - b) < 1e-6 * max(abs(a), 1.0), (name, obj, a, b)


def test_both_constructors_produce_identical_results():
    """properties.rs:249-298: from_edges and from_csr_parts(to_csr_parts, node_weights) give the
    same graph bit for bit and the same clustering."""
    for seed in range(300):
        name, n, edges = random_graph(seed)
        via_edges = scgpu.CSRNetwork.from_edges(n, edges)
        ptrs, nbrs, wts = via_edges.to_csr_parts()
        via_csr = scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts, via_edges.node_weights())
        assert via_edges.node_count() == via_csr.node_count(), name
        assert via_edges.edge_count() == via_csr.edge_count(), name
        assert via_edges.total_weight() == via_csr.total_weight(), name
        assert np.array_equal(via_edges.strengths(), via_csr.strengths()), name
        for a, b in zip(via_edges.to_csr_parts(), via_csr.to_csr_parts()):
            assert np.array_equal(a, b), name
        via_csr.validate_symmetry()
        cfg = random_config(seed, 0x5555)
        assert scgpu.leiden(via_edges, cfg) == scgpu.leiden(via_csr, cfg), name


def test_resolution_is_monotone_on_random_inputs():
    """properties.rs:301-327: the trend of the cluster count over resolutions holds, and it
    never drops by more than 2 between adjacent resolutions."""
    for seed in range(120):
        name, n, edges = random_graph(seed)
        g = scgpu.CSRNetwork.from_edges(n, edges)
        if n < 10 or g.edge_count() == 0:
            continue
        counts = [scgpu.leiden(g, scgpu.LeidenConfig.with_resolution(r)).n_clusters()
                  for r in (0.05, 0.25, 1.0, 4.0, 16.0, 64.0)]
        assert counts[0] <= counts[-1], (name, counts)
        for a, b in zip(counts, counts[1:]):
            assert b + 2 >= a, (name, counts)


def test_rb_is_scale_invariant_on_random_inputs():
    """properties.rs:330-359: weights x 64 (exact in binary) leave the Rb partition identical
    and modularity unchanged to 1e-6."""
    for seed in range(120):
        name, n, edges = random_graph(seed)
        g = scgpu.CSRNetwork.from_edges(n, edges)
        if g.edge_count() == 0:
            continue
        ptrs, nbrs, wts = g.to_csr_parts()
        scaled = scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts * np.float32(64.0), g.node_weights())
        a = scgpu.leiden(g, scgpu.LeidenConfig())
        b = scgpu.leiden(scaled, scgpu.LeidenConfig())
        assert np.array_equal(a.labels(), b.labels()), name
        qa, qb = scgpu.modularity(g, a.labels(), 1.0), scgpu.modularity(scaled, b.labels(), 1.0)
        assert abs(qa - qb) < 1e-6, (name, qa, qb)


def test_rb_is_invariant_to_uniform_edge_scaling():
    """known_answers.rs:156-171: every weight x 17.5 leaves the labels identical."""
    edges, _ = W.sbm(30, 4, 0.30, 0.03, 2)
    g = scgpu.CSRNetwork.from_edges(120, edges)
    scaled = scgpu.CSRNetwork.from_edges(120, [(a, b, w * 17.5) for a, b, w in edges])
    a = scgpu.leiden(g, scgpu.LeidenConfig())
    b = scgpu.leiden(scaled, scgpu.LeidenConfig())
    assert np.array_equal(a.labels(), b.labels())


def test_results_are_stable_under_node_relabelling():
    """known_answers.rs:173-198: reversing the node numbering changes labels, not the grouping."""
    edges, _ = W.sbm(25, 4, 0.35, 0.02, 6)
    n = 100
    a = scgpu.leiden(scgpu.CSRNetwork.from_edges(n, edges), scgpu.LeidenConfig())
    b = scgpu.leiden(scgpu.CSRNetwork.from_edges(n, [(n - 1 - x, n - 1 - y, w) for x, y, w in edges]),
                     scgpu.LeidenConfig())
    assert O.nmi(a.labels(), b.labels()[::-1].copy()) > 0.95


def test_found_partition_scores_at_least_the_planted_one():
    """known_answers.rs:71-84."""
    edges, truth = W.sbm(60, 6, 0.25, 0.01, 11)
    g = scgpu.CSRNetwork.from_edges(360, edges)
    c = scgpu.leiden(g, scgpu.LeidenConfig())
    found = scgpu.modularity(g, c.labels(), 1.0)
    planted = scgpu.modularity(g, truth.astype(np.uint64), 1.0)
    assert found >= planted - 1e-9, (found, planted)


def test_karate_splits_in_two_at_low_resolution():
    """known_answers.rs:32-42."""
    g = scgpu.CSRNetwork.from_edges(34, W.karate_edges())
    assert 2 <= scgpu.leiden(g, scgpu.LeidenConfig.with_resolution(0.5)).n_clusters() <= 3


def test_karate_reaches_four_communities():
    """known_answers.rs:12-29: 4 communities at Q = 0.4198 +- 1e-3."""
    g = scgpu.CSRNetwork.from_edges(34, W.karate_edges())
    c = scgpu.leiden(g, scgpu.LeidenConfig())
    assert c.n_clusters() == 4
    assert abs(scgpu.modularity(g, c.labels(), 1.0) - 0.4198) < 1e-3


def test_degenerate_graphs_behave():
    """known_answers.rs:120-153: isolated nodes and self-loop-only graphs stay fully split."""
    assert scgpu.leiden(scgpu.CSRNetwork.from_edges(6, []), scgpu.LeidenConfig()).n_clusters() == 6
    loops = scgpu.CSRNetwork.from_edges(3, [(0, 0, 1.0), (1, 1, 1.0), (2, 2, 1.0)])
    assert scgpu.leiden(loops, scgpu.LeidenConfig()).n_clusters() == 3


def test_renumber_is_bit_exact_against_the_oracle():
    """partition.rs:265-306 / grouping.rs:154-172, 226-254 on the device (the compaction every
    level of leiden() runs) against the oracle's restatement, including the reference's KAT."""
    out, k = scgpu.renumber(np.array([5, 5, 5, 2, 2, 2], np.uint64))
    assert list(out) == [1, 1, 1, 0, 0, 0] and k == 2
    rng = np.random.default_rng(12)
    for n, hi in [(1, 1), (7, 3), (1000, 50), (100_000, 10 ** 6), (1_000_000, 5000), (300_000, 10 ** 8)]:
        ids = rng.integers(0, hi, n, dtype=np.uint64) if hi > 1 else np.zeros(n, np.uint64)
        got, gk = scgpu.renumber(ids)
        want, wk = O.normalize_groups(ids)
        assert gk == wk and np.array_equal(got, want), (n, hi)
    got, gk = scgpu.renumber(np.zeros(0, np.uint64))
    assert gk == 0 and len(got) == 0
    with pytest.raises(scgpu.ClusteringError) as e:
        scgpu.renumber(np.array([0, 2 ** 32 - 1], np.uint64))
    assert e.value.variant == "InvalidArgument"


def hub_graph(seed=5):
    """Three hubs of degree 700, 1500 and 3200 (the 513..1024, 1025..2048 and > 2048 degree
    classes: the two largest LDS table layouts and the global-memory table), each joined to
    its own leaves, which form 6-cliques; hubs joined to each other and a few random leaf
    links across hubs. Weights ~ U(0.5, 1.5)."""
    rng = np.random.default_rng(seed)
    edges = []
    nxt = 3
    edges += [(0, 1, 1.0), (1, 2, 1.0), (0, 2, 1.0)]
    leaves = []
    for h, deg in enumerate((700, 1500, 3200)):
        mine = list(range(nxt, nxt + deg))
        nxt += deg
        leaves.append(mine)
        for v in mine:
            edges.append((h, v, float(rng.uniform(0.5, 1.5))))
        for c0 in range(0, deg - deg % 6, 6):
            for i in range(6):
                for j in range(i + 1, 6):
                    edges.append((mine[c0 + i], mine[c0 + j], float(rng.uniform(0.5, 1.5))))
    for _ in range(400):
        a, b = rng.choice(3, 2, replace=False)
        edges.append((int(rng.choice(leaves[a])), int(rng.choice(leaves[b])), 1.0))
    return nxt, edges


def test_hub_vertices_use_the_large_degree_classes():
    """Hubs of degree 700, 1500 and 3200 exercise the decide step's largest LDS layouts and its
    global-memory table. Quality within 0.5% of the oracle, no disconnected community, and the
    persistent, multi-launch and graph-replayed sweeps agree label for label."""
    n, edges = hub_graph()
    g = scgpu.CSRNetwork.from_edges(n, edges)
    og = O.Graph.from_edges(n, edges)
    degs = np.diff(g.to_csr_parts()[0].astype(np.int64))
    assert degs.max() > 2048 and ((degs > 1024) & (degs <= 2048)).any() and ((degs > 512) & (degs <= 1024)).any()
    cfgs = [scgpu.LeidenConfig(), scgpu.LeidenConfig(seed=3, objective=scgpu.ObjectiveKind.Cpm(0.1)),
            scgpu.LeidenConfig(seed=9, max_community_weight=60.0), scgpu.LeidenConfig(refine=False)]
    try:
        for cfg in cfgs:
            scgpu.set_sweep_mode("persistent")
            c = scgpu.leiden(g, cfg)
            scgpu.set_sweep_mode("multi")
            m = scgpu.leiden(g, cfg)
            scgpu.set_sweep_mode("graph")
            r = scgpu.leiden(g, cfg)
            assert np.array_equal(c.labels(), m.labels()) and np.array_equal(c.labels(), r.labels()), cfg
            kind, res = cfg.objective.kind, cfg.objective.resolution
            ol, _ = O.leiden(og, objective=kind, resolution=res, seed=cfg.seed, refine=int(cfg.refine),
                             max_community_weight=cfg.max_community_weight)
            qg, qo = O.quality(og, c.labels(), kind, res), O.quality(og, ol, kind, res)
            assert qg >= qo - 0.005 * abs(qo), (cfg, qg, qo)
            if cfg.refine:
                assert O.disconnected_community_count(og, c.labels()) == 0
            if cfg.max_community_weight is not None:
                assert (c.cluster_sizes() <= cfg.max_community_weight).all()
    finally:
        scgpu.set_sweep_mode("persistent")
