"""Shared fixtures: the reference's test graph set (tests/common/mod.rs:57-130) as edge lists,
plus the leidenalg fixture loader (tests/reference.rs:31-83)."""
import json
import os

import numpy as np

import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference_test_graphs():
    """(name, n, edges) for the reference's test_graphs() set."""
    out = [
        ("empty", 0, []),
        ("single_node", 1, []),
        ("no_edges", 8, []),
        ("one_edge", 2, [(0, 1, 1.0)]),
        ("self_loops_only", 4, [(0, 0, 1.0), (1, 1, 2.0), (2, 2, 0.5)]),
        ("with_self_loops", 5, [(0, 1, 1.0), (1, 2, 1.0), (0, 0, 3.0), (3, 4, 1.0), (3, 3, 2.0)]),
        ("star", 9, [(0, i, 1.0) for i in range(1, 9)]),
        ("path", 12, [(i, i + 1, 1.0) for i in range(11)]),
        ("karate", 34, W.karate_edges()),
    ]
    rng = np.random.default_rng(3)
    out.append(("karate_weighted", 34, [(a, b, 0.25 + 3.5 * rng.random()) for a, b, _ in W.karate_edges()]))
    edges, _ = W.sbm(25, 4, 0.35, 0.03, 1)
    out.append(("sbm", 100, edges))
    cl = []
    for blk in range(3):
        for i in range(5):
            for j in range(i + 1, 5):
                cl.append((blk * 5 + i, blk * 5 + j, 1.0))
    out.append(("disconnected_cliques", 15, cl))
    return out


def ring_of_cliques(n_cliques=20, size=5):
    edges = []
    for c in range(n_cliques):
        base = c * size
        for i in range(size):
            for j in range(i + 1, size):
                edges.append((base + i, base + j, 1.0))
        edges.append((base, ((c + 1) % n_cliques) * size, 1.0))
    return n_cliques * size, edges


def load_fixture():
    with open(os.path.join(GOLDEN, "leidenalg_reference.json")) as f:
        doc = json.load(f)
    graphs = {k: (v["n_nodes"], [tuple(e) for e in v["edges"]]) for k, v in doc["graphs"].items()}
    return graphs, doc["cases"]
