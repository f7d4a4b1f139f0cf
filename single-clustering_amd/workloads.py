"""Synthetic inputs for the benchmark configurations and the tests (numpy, seeded).

- planted_partition: n nodes in equal blocks, each unordered pair joined with p_in inside a
  block and p_out across (unit f32 weights) -> full symmetric CSR.
- gaussian_mixture: points from an isotropic Gaussian mixture (centres ~ N(0, centre_var I),
  per-cluster sd, proportions ~ Dirichlet(alpha)), rows shuffled.
- sbm / karate: the small graphs of the reference's tests (tests/common/mod.rs:11-130), drawn
  with numpy's generator instead of ChaCha8 (same distribution, different stream).
"""
import numpy as np

KARATE_EDGES = [
    (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (0, 6), (0, 7), (0, 8), (0, 10), (0, 11),
    (0, 12), (0, 13), (0, 17), (0, 19), (0, 21), (0, 31), (1, 2), (1, 3), (1, 7), (1, 13),
    (1, 17), (1, 19), (1, 21), (1, 30), (2, 3), (2, 7), (2, 8), (2, 9), (2, 13), (2, 27),
    (2, 28), (2, 32), (3, 7), (3, 12), (3, 13), (4, 6), (4, 10), (5, 6), (5, 10), (5, 16),
    (6, 16), (8, 30), (8, 32), (8, 33), (9, 33), (13, 33), (14, 32), (14, 33), (15, 32),
    (15, 33), (18, 32), (18, 33), (19, 33), (20, 32), (20, 33), (22, 32), (22, 33),
    (23, 25), (23, 27), (23, 29), (23, 32), (23, 33), (24, 25), (24, 27), (24, 31),
    (25, 31), (26, 29), (26, 33), (27, 33), (28, 31), (28, 33), (29, 32), (29, 33),
    (30, 32), (30, 33), (31, 32), (31, 33), (32, 33),
]
KARATE_OPTIMUM = [0, 0, 0, 0, 1, 1, 1, 0, 2, 2, 1, 0, 0, 0, 2, 2, 1, 0, 2, 0, 2, 0, 2, 3, 3, 3, 2,
                  3, 3, 2, 2, 3, 2, 2]


def karate_edges():
    return [(a, b, 1.0) for a, b in KARATE_EDGES]


def edges_to_csr(n, a, b, w):
    """Full symmetric CSR (rows sorted) from undirected edges a[i] < b[i] (or loops a == b)."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    w = np.asarray(w, dtype=np.float32)
    loop = a == b
    rows = np.concatenate([a, b[~loop]])
    cols = np.concatenate([b, a[~loop]])
    vals = np.concatenate([w, w[~loop]])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ptrs = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(ptrs, rows + 1, 1)
    ptrs = np.cumsum(ptrs).astype(np.uint64)
    return ptrs, cols.astype(np.uint32), vals.astype(np.float32)


def _sample_pairs_within(rng, size, p):
    """Bernoulli(p) over all unordered pairs of `size` items: returns (i, j) with i < j."""
    total = size * (size - 1) // 2
    m = rng.binomial(total, p)
    if m == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    idx = np.unique(rng.integers(0, total, size=int(m * 1.02) + 16))
    while len(idx) < m:
        idx = np.unique(np.concatenate([idx, rng.integers(0, total, size=m - len(idx) + 16)]))
    idx = rng.permutation(idx)[:m]
    # decode triangular index: pairs ordered by j then i < j
    j = np.floor((1 + np.sqrt(1 + 8 * idx.astype(np.float64))) / 2).astype(np.int64)
    j = np.where(j * (j - 1) // 2 > idx, j - 1, j)
    j = np.where((j + 1) * j // 2 <= idx, j + 1, j)
    i = idx - j * (j - 1) // 2
    return i, j


def planted_partition(n=10000, blocks=20, p_in=0.01, p_out=0.0005, seed=42):
    """Planted partition graph; returns (ptrs, nbrs, wts, truth)."""
    rng = np.random.default_rng(seed)
    size = n // blocks
    truth = np.repeat(np.arange(blocks), size)
    A, B = [], []
    for blk in range(blocks):
        i, j = _sample_pairs_within(rng, size, p_in)
        A.append(i + blk * size)
        B.append(j + blk * size)
    # across blocks: sample pairs of all nodes, keep those in different blocks
    total_across = n * (n - 1) // 2 - blocks * (size * (size - 1) // 2)
    m = rng.binomial(total_across, p_out)
    got_a, got_b = [], []
    have = 0
    seen = set()
    while have < m:
        k = (m - have) * 2 + 64
        x = rng.integers(0, n, size=k)
        y = rng.integers(0, n, size=k)
        for u, v in zip(x, y):
            if u == v or truth[u] == truth[v]:
                continue
            key = (min(u, v), max(u, v))
            if key in seen:
                continue
            seen.add(key)
            got_a.append(key[0])
            got_b.append(key[1])
            have += 1
            if have == m:
                break
    A.append(np.asarray(got_a, dtype=np.int64))
    B.append(np.asarray(got_b, dtype=np.int64))
    a = np.concatenate(A)
    b = np.concatenate(B)
    ptrs, nbrs, wts = edges_to_csr(n, a, b, np.ones(len(a), np.float32))
    return ptrs, nbrs, wts, truth


def gaussian_mixture(n, d, n_comp, centre_var, sd=1.0, alpha=2.0, seed=0):
    """Points (n, d) f32 and their component labels, rows in random order."""
    rng = np.random.default_rng(seed)
    props = rng.dirichlet(np.full(n_comp, alpha))
    comp = rng.choice(n_comp, size=n, p=props)
    centres = rng.normal(0.0, np.sqrt(centre_var), size=(n_comp, d))
    x = centres[comp] + rng.normal(0.0, sd, size=(n, d))
    return x.astype(np.float32), comp


def sbm(n_per, blocks, p_in, p_out, seed):
    """Edge list and truth of the reference's sbm() (tests/common/mod.rs:34-54), numpy stream."""
    rng = np.random.default_rng(seed)
    n = n_per * blocks
    truth = np.arange(n) // n_per
    iu, ju = np.triu_indices(n, 1)
    p = np.where(truth[iu] == truth[ju], p_in, p_out)
    keep = rng.random(len(iu)) < p
    return [(int(a), int(b), 1.0) for a, b in zip(iu[keep], ju[keep])], truth


def edges_csr_from_list(n, edges):
    if not edges:
        return np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.float32)
    a = np.array([e[0] for e in edges], dtype=np.int64)
    b = np.array([e[1] for e in edges], dtype=np.int64)
    w = np.array([e[2] for e in edges], dtype=np.float64)
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    return edges_to_csr(n, lo, hi, w)


def _power_law_ints(rng, size, exponent, lo, hi):
    """Integers in [lo, hi] with P(x) proportional to x^-exponent (inverse-CDF sampling)."""
    u = rng.random(size)
    a = 1.0 - exponent
    x = (lo ** a + u * (hi ** a - lo ** a)) ** (1.0 / a)
    return np.clip(np.floor(x).astype(np.int64), lo, hi)


def lfr_like(n, seed, mean_degree=15.0, tau_degree=2.5, tau_size=1.5, mu=0.3):
    """Planted partition with power-law degrees (exponent tau_degree, mean ~mean_degree) and
    power-law community sizes (exponent tau_size); a fraction mu of every vertex's stubs leaves
    its community; stubs are paired configuration-model style (self-loops and repeats dropped);
    weights ~ U(0.5, 1.5) f32. Returns (ptrs u64, nbrs u32, wts f32, truth)."""
    rng = np.random.default_rng(seed)
    dmax = int(min(n - 1, max(50, 4 * np.sqrt(n))))
    dmin = 1
    for cand in range(1, 40):  # smallest minimum degree whose expected mean reaches the target
        d = np.arange(cand, dmax + 1, dtype=np.float64)
        p = d ** -tau_degree
        if (d * p).sum() / p.sum() >= mean_degree:
            dmin = cand
            break
    deg = _power_law_ints(rng, n, tau_degree, dmin, dmax)
    cmin, cmax = 20, max(100, n // 10)
    sizes = []
    while sum(sizes) < n:
        sizes.extend(_power_law_ints(rng, 64, tau_size, cmin, cmax).tolist())
    sizes = np.array(sizes, dtype=np.int64)
    sizes = sizes[np.cumsum(sizes) - sizes < n]
    sizes[-1] = n - (sizes.sum() - sizes[-1])
    truth = rng.permutation(np.repeat(np.arange(len(sizes)), sizes))
    k_in = np.minimum(np.rint((1.0 - mu) * deg).astype(np.int64), np.maximum(sizes[truth] - 1, 0))
    k_out = deg - k_in
    # intra-community stubs: shuffled within their community, paired consecutively
    stubs = np.repeat(np.arange(n), k_in)
    comm = truth[stubs]
    order = np.lexsort((rng.random(len(stubs)), comm))
    stubs, comm = stubs[order], comm[order]
    counts = np.bincount(comm, minlength=len(sizes))
    ends = np.cumsum(counts)
    drop = ends[counts % 2 == 1] - 1
    keep = np.ones(len(stubs), dtype=bool)
    keep[drop] = False
    stubs = stubs[keep]
    a_in, b_in = stubs[0::2], stubs[1::2]
    out = rng.permutation(np.repeat(np.arange(n), k_out))
    if len(out) % 2:
        out = out[:-1]
    a = np.concatenate([a_in, out[0::2]])
    b = np.concatenate([b_in, out[1::2]])
    m = a != b
    lo, hi = np.minimum(a[m], b[m]), np.maximum(a[m], b[m])
    key = np.unique(lo.astype(np.int64) * n + hi)
    lo, hi = key // n, key % n
    w = rng.uniform(0.5, 1.5, len(lo)).astype(np.float32)
    rows = np.concatenate([lo, hi])
    cols = np.concatenate([hi, lo])
    ww = np.concatenate([w, w])
    o = np.lexsort((cols, rows))
    rows, cols, ww = rows[o], cols[o], ww[o]
    ptrs = np.zeros(n + 1, dtype=np.uint64)
    ptrs[1:] = np.cumsum(np.bincount(rows, minlength=n))
    return ptrs, cols.astype(np.uint32), ww, truth


def batch_graph_sizes(count=64, lo=2000, hi=200000, seed=42):
    """Graph sizes log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
