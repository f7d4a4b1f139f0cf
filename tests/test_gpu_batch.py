"""Batched Leiden (many graphs x resolutions in one run): every job's labels must equal the
standalone leiden() call on that graph at that resolution, which the batch is defined to be
equivalent to (a loop of leiden(graph, config.with_resolution(r)), leiden/mod.rs:168-204)."""
from dataclasses import replace

import numpy as np
import pytest

import scgpu
import workloads as W
from helpers import ring_of_cliques

pytestmark = pytest.mark.gpu


def _graphs():
    gs = [scgpu.CSRNetwork.from_edges(34, W.karate_edges())]
    edges, _ = W.sbm(40, 5, 0.30, 0.02, 4)
    gs.append(scgpu.CSRNetwork.from_edges(200, edges))
    n, e = ring_of_cliques()
    gs.append(scgpu.CSRNetwork.from_edges(n, e))
    ptrs, nbrs, wts, _ = W.planted_partition(n=3000, blocks=6, p_in=0.02, p_out=0.001, seed=5)
    gs.append(scgpu.CSRNetwork.from_csr_parts(ptrs, nbrs, wts))
    x, _ = W.gaussian_mixture(6000, 16, 8, 9.0, seed=2)
    gs.append(scgpu.knn_graph(x, 8))
    return gs


def _check(gs, resolutions, cfg):
    jobs = scgpu.leiden_batch(gs, resolutions, cfg)
    assert len(jobs) == len(gs) * len(resolutions)
    for gi, g in enumerate(gs):
        for ri, r in enumerate(resolutions):
            single = scgpu.leiden(g, replace(cfg, objective=scgpu.ObjectiveKind(cfg.objective.kind, r)))
            got = jobs[gi * len(resolutions) + ri]
            assert got.n_clusters() == single.n_clusters(), (gi, r)
            assert np.array_equal(got.labels(), single.labels()), (gi, r)


def test_batch_rb_equals_standalone():
    _check(_graphs(), [0.25, 1.0, 3.0], scgpu.LeidenConfig(seed=42))


def test_batch_cpm_and_seed_equal_standalone():
    _check(_graphs(), [0.02, 0.1], scgpu.LeidenConfig(seed=7, objective=scgpu.ObjectiveKind.Cpm(0.05)))


def test_batch_louvain_and_capped_equal_standalone():
    gs = _graphs()[:4]
    _check(gs, [0.5, 1.0], scgpu.LeidenConfig(refine=False))
    _check(gs, [1.0], scgpu.LeidenConfig(max_community_weight=30.0))


def test_batch_degenerate_and_errors():
    gs = [scgpu.CSRNetwork.from_edges(0, []), scgpu.CSRNetwork.from_edges(1, []),
          scgpu.CSRNetwork.from_edges(2, [(0, 1, 1.0)])]
    jobs = scgpu.leiden_batch(gs, [1.0, 2.0])
    assert [len(j) for j in jobs] == [0, 0, 1, 1, 2, 2]
    # one edge: joined at resolution 1 (gain 1 - 3/4 > 0), apart at 2 (1 - 6/4 < 0)
    assert [j.n_clusters() for j in jobs] == [0, 0, 1, 1, 1, 2]
    _check(gs, [1.0, 2.0], scgpu.LeidenConfig())
    with pytest.raises(scgpu.ClusteringError) as e:
        scgpu.leiden_batch(gs, [-1.0])
    assert e.value.variant == "InvalidConfig"
