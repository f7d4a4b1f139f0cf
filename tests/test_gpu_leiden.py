"""GPU Leiden: the reference's invariants (tests/invariants.rs), known answers
(tests/known_answers.rs) and optimizer statistics (tests/reference.rs), plus quality parity with
the CPU oracle. Parallel local moving cannot reproduce the sequential visit order, so full runs
are judged by quality (within 0.5% of the oracle at the same seed), invariants and ARI."""
import numpy as np
import pytest

import pyoracle as O
import scgpu
import workloads as W
from helpers import load_fixture, reference_test_graphs, ring_of_cliques

pytestmark = pytest.mark.gpu

OBJECTIVES = [scgpu.ObjectiveKind.Rb(0.5), scgpu.ObjectiveKind.Rb(1.0), scgpu.ObjectiveKind.Rb(3.0),
              scgpu.ObjectiveKind.Cpm(0.05)]


def test_karate_reaches_known_optimum():
    g = scgpu.CSRNetwork.from_edges(34, W.karate_edges())
    c = scgpu.leiden(g, scgpu.LeidenConfig())
    q = scgpu.modularity(g, c.labels(), 1.0)
    assert q >= 0.4198 - 1e-3 - 0.002, q   # within 0.5% of the optimum the reference reaches
    assert 3 <= c.n_clusters() <= 5


def test_invariants_on_reference_graphs():
    for name, n, edges in reference_test_graphs():
        g = scgpu.CSRNetwork.from_edges(n, edges)
        og = O.Graph.from_edges(n, edges)
        for seed in range(3):
            for obj in OBJECTIVES:
                c = scgpu.leiden(g, scgpu.LeidenConfig(objective=obj, seed=seed))
                labels = c.labels()
                assert len(labels) == n
                if n:
                    assert sorted(set(labels.tolist())) == list(range(c.n_clusters())), name
                assert O.disconnected_community_count(og, labels) == 0, (name, seed, obj)
        if n:
            first = scgpu.leiden(g, scgpu.LeidenConfig(seed=7))
            for _ in range(3):
                assert scgpu.leiden(g, scgpu.LeidenConfig(seed=7)) == first, name
            for gamma in (0.5, 1.0, 2.0):
                start = O.quality(og, np.arange(n), "rb", gamma)
                c = scgpu.leiden(g, scgpu.LeidenConfig.with_resolution(gamma))
                assert O.quality(og, c.labels(), "rb", gamma) >= start - 1e-9, name


def test_quality_invariant_under_aggregation():
    for name, n, edges in reference_test_graphs():
        if n == 0:
            continue
        g = scgpu.CSRNetwork.from_edges(n, edges)
        c = scgpu.leiden(g, scgpu.LeidenConfig())
        coarse = g.aggregate(c.labels())
        cl = np.arange(coarse.node_count(), dtype=np.uint64)
        for obj in (scgpu.ObjectiveKind.Rb(0.25), scgpu.ObjectiveKind.Rb(1.0), scgpu.ObjectiveKind.Rb(4.0),
                    scgpu.ObjectiveKind.Cpm(0.05)):
            fine = scgpu.quality(g, c.labels(), obj)
            agg = scgpu.quality(coarse, cl, obj)
            assert abs(fine - agg) < 1e-6 * max(abs(fine), 1.0), name


def test_components_never_merged():
    edges = [(b * 6 + i, b * 6 + j, 1.0) for b in range(3) for i in range(6) for j in range(i + 1, 6)]
    g = scgpu.CSRNetwork.from_edges(18, edges)
    for seed in range(5):
        labels = scgpu.leiden(g, scgpu.LeidenConfig(seed=seed)).labels()
        for a in range(18):
            for b in range(18):
                if a // 6 != b // 6:
                    assert labels[a] != labels[b]


def test_cluster_count_monotone_in_resolution():
    edges, _ = W.sbm(40, 5, 0.30, 0.02, 4)
    g = scgpu.CSRNetwork.from_edges(200, edges)
    counts = [scgpu.leiden(g, scgpu.LeidenConfig.with_resolution(r)).n_clusters()
              for r in (0.1, 0.5, 1.0, 2.0, 8.0, 32.0)]
    assert all(a <= b for a, b in zip(counts, counts[1:])), counts
    assert counts[0] < counts[-1]


def test_planted_partitions_recovered():
    for n_per, blocks, p_in, p_out, seed in [(60, 6, 0.25, 0.01, 11), (40, 4, 0.35, 0.02, 5), (80, 3, 0.20, 0.01, 9)]:
        edges, truth = W.sbm(n_per, blocks, p_in, p_out, seed)
        g = scgpu.CSRNetwork.from_edges(n_per * blocks, edges)
        c = scgpu.leiden(g, scgpu.LeidenConfig())
        assert O.nmi(c.labels(), truth) > 0.95
        assert c.n_clusters() == blocks


def test_cpm_ring_of_cliques():
    n, edges = ring_of_cliques()
    g = scgpu.CSRNetwork.from_edges(n, edges)
    c = scgpu.leiden(g, scgpu.LeidenConfig(objective=scgpu.ObjectiveKind.Cpm(0.1)))
    assert c.n_clusters() == 20


def test_max_community_weight_respected():
    edges, _ = W.sbm(40, 5, 0.3, 0.02, 4)
    g = scgpu.CSRNetwork.from_edges(200, edges)
    c = scgpu.leiden(g, scgpu.LeidenConfig(max_community_weight=25.0))
    assert c.cluster_sizes().max() <= 25


def test_louvain_mode():
    """leiden/mod.rs:483-494 (refine: false on karate reaches Q > 0.41), and the same run of
    the oracle: quality within 0.5% of it on karate and on a planted partition."""
    g = scgpu.CSRNetwork.from_edges(34, W.karate_edges())
    c = scgpu.leiden(g, scgpu.LeidenConfig(refine=False))
    q = scgpu.modularity(g, c.labels(), 1.0)
    assert q > 0.41, q
    og = O.Graph.from_edges(34, W.karate_edges())
    ol, _ = O.leiden(og, refine=0)
    assert O.modularity(og, ol) > 0.41
    assert q >= O.modularity(og, ol) * (1 - 0.005)
    ptrs, nbrs, wts, _ = W.planted_partition(n=3000, blocks=10, p_in=0.03, p_out=0.001, seed=3)
    g = scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts)
    og = O.Graph.from_csr_parts(ptrs, nbrs, wts)
    # CPM at 0.01 (between p_out and p_in): the oracle itself lands within 0.02% across seeds
    # here; at 0.02 its own results spread by 2.7% over seeds 0..5, too flat for a 0.5% bar
    for obj, kind, res in [(scgpu.ObjectiveKind.Rb(1.0), "rb", 1.0), (scgpu.ObjectiveKind.Cpm(0.01), "cpm", 0.01)]:
        c = scgpu.leiden(g, scgpu.LeidenConfig(refine=False, objective=obj))
        ol, _ = O.leiden(og, refine=0, objective=kind, resolution=res)
        qg, qo = O.quality(og, c.labels(), kind, res), O.quality(og, ol, kind, res)
        assert qg >= qo - 0.005 * abs(qo), (kind, qg, qo)


def test_degenerate_graphs_and_bad_configs():
    for n, edges, expect in [(0, [], 0), (1, [], 1), (5, [], 5), (2, [(0, 1, 1.0)], 1)]:
        g = scgpu.CSRNetwork.from_edges(n, edges)
        c = scgpu.leiden(g, scgpu.LeidenConfig())
        assert len(c) == n and c.n_clusters() == expect
    g = scgpu.CSRNetwork.from_edges(34, W.karate_edges())
    for bad in [scgpu.LeidenConfig(n_iterations=0), scgpu.LeidenConfig(max_levels=0),
                scgpu.LeidenConfig.with_resolution(-1.0), scgpu.LeidenConfig(max_community_weight=0.0)]:
        with pytest.raises(scgpu.ClusteringError) as e:
            scgpu.leiden(g, bad)
        assert e.value.variant == "InvalidConfig"


def test_fixture_quality_vs_leidenalg():
    """tests/reference.rs:248-280 bound (mean >= -0.5%), plus no case grossly short."""
    graphs, cases = load_fixture()
    built = {k: (scgpu.CSRNetwork.from_edges(n, e), O.Graph.from_edges(n, e)) for k, (n, e) in graphs.items()}
    rels = []
    for c in cases:
        g, og = built[c["graph"]]
        obj = scgpu.ObjectiveKind(c["objective"], c["resolution"])
        runs = [scgpu.leiden(g, scgpu.LeidenConfig(objective=obj, seed=42 + s, n_iterations=5)) for s in range(2)]
        quals = [2 * O.quality(og, r.labels(), c["objective"], c["resolution"]) for r in runs]
        ref = c["reference_quality"]
        rels.append((quals[0] - ref) / max(abs(ref), 1e-9))
        assert max(quals) >= ref - 0.08 * max(abs(ref), 1.0), (c["name"], quals, ref)
    print(f"GPU vs leidenalg mean rel {100 * np.mean(rels):+.4f}%")
    assert np.mean(rels) > -0.005


def test_planted_partition_small_config_vs_oracle():
    """Config S (10k nodes, 20 blocks, p=0.01/0.0005, Rb 1.0, seed 42): quality within 0.5%."""
    ptrs, nbrs, wts, truth = W.planted_partition(seed=42)
    g = scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts)
    og = O.Graph.from_csr_parts(ptrs, nbrs, wts)
    c = scgpu.leiden(g, scgpu.LeidenConfig(seed=42))
    ol, onc = O.leiden(og, seed=42)
    q_gpu = O.quality(og, c.labels(), "rb", 1.0)
    q_cpu = O.quality(og, ol, "rb", 1.0)
    ari = O.ari(c.labels(), ol)
    print(f"config S: gpu Q={q_gpu / og.total_weight():.6f} ({c.n_clusters()}), "
          f"oracle Q={q_cpu / og.total_weight():.6f} ({onc}), ARI={ari:.4f}, stats={scgpu.last_stats()}")
    # mixing is ~0.49 here, so many near-optimal partitions exist and two good optimizers can
    # disagree (ARI is stated, not bounded); quality and connectivity are the bar
    assert q_gpu >= q_cpu - 0.005 * abs(q_cpu)
    assert O.disconnected_community_count(og, c.labels()) == 0


def test_persistent_sweeps_match_multi_launch():
    """The cooperative one-launch-per-phase sweeps, the one-launch-per-kernel sweeps and the
    graph-replayed iterations run the same synchronous algorithm, so labels must be identical.
    40k points give ~600 workgroups, many per XCD, which exercises the grid barrier's cross-XCD
    visibility; each graph is replayed once per iteration of its level."""
    x, _ = W.gaussian_mixture(40_000, 24, 20, 9.0, sd=1.0, alpha=2.0, seed=3)
    g = scgpu.knn_graph(x, 10)
    # the same graph with self-loops on every fifth vertex: the refinement's first step and the
    # singleton terms read a vertex's own loop weight, so the three paths must agree there too
    import scipy.sparse as sp
    ptrs, nbrs, wts = g.to_csr_parts()
    a = sp.csr_matrix((wts, nbrs, ptrs.astype(np.int64)), shape=(len(ptrs) - 1,) * 2)
    loops = np.zeros(len(ptrs) - 1, dtype=np.float32)
    loops[::5] = 0.75
    a = (a + sp.diags(loops)).tocsr()
    a.sort_indices()
    gl = scgpu.CSRNetwork.from_csr_parts(a.indptr, a.indices, a.data.astype(np.float32))
    runs = [(g, scgpu.LeidenConfig(seed=42)), (g, scgpu.LeidenConfig(seed=7, objective=scgpu.ObjectiveKind.Cpm(0.05))),
            (g, scgpu.LeidenConfig(seed=11, max_community_weight=400.0)), (gl, scgpu.LeidenConfig(seed=5)),
            (gl, scgpu.LeidenConfig(seed=9, objective=scgpu.ObjectiveKind.Cpm(0.05)))]
    try:
        for g, cfg in runs:
            scgpu.set_sweep_mode("persistent")
            p = [scgpu.leiden(g, cfg).labels() for _ in range(3)]
            scgpu.set_sweep_mode("multi")
            m = scgpu.leiden(g, cfg).labels()
            scgpu.set_sweep_mode("graph")
            gr = [scgpu.leiden(g, cfg).labels() for _ in range(2)]
            for r in p[1:]:
                assert np.array_equal(p[0], r), "persistent sweeps are not deterministic"
            assert np.array_equal(p[0], m), f"persistent and multi-launch labels differ for {cfg}"
            for r in gr:
                assert np.array_equal(m, r), f"graph-replayed and multi-launch labels differ for {cfg}"
    finally:
        scgpu.set_sweep_mode("persistent")
    with pytest.raises(ValueError):
        scgpu.set_sweep_mode("bogus")
