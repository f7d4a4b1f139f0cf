"""GPU kNN: exact index sets and FP64 distances against the brute-force FP64 oracle, for the
split-BF16 MFMA path and the FP32 MFMA path; Gaussian connectivity against the oracle."""
import numpy as np
import pytest

import pyoracle as O
import scgpu
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("n,d,k", [(700, 50, 15), (2000, 30, 7), (3000, 50, 15), (129, 5, 4)])
def test_knn_exact_vs_oracle(precision, n, d, k):
    x, _ = W.gaussian_mixture(n, d, 10, 16.0, seed=n + d)
    idx, d2, st = scgpu.knn(x, k, precision=precision, return_stats=True)
    rows = np.arange(n) if n <= 1000 else np.random.default_rng(0).choice(n, 400, replace=False)
    oi, od = O.knn_bruteforce(x, k, rows=rows)
    assert np.array_equal(idx[rows], oi)
    assert np.array_equal(d2[rows].view(np.uint64), od.view(np.uint64))  # FP64 bit-exact
    print(f"{precision} n={n} d={d} k={k}: fallback rows {st['rows_fallback']}, tile {st['tile_kernel_ms']:.3f} ms")


@pytest.fixture
def knn_prune(monkeypatch):
    def set_mode(v):
        monkeypatch.setenv("SCGPU_KNN_PRUNE", v)
    return set_mode


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("n,d,k,comps,spread", [
    (1500, 50, 15, 10, 16.0),   # a few row blocks, clusters far apart: most column blocks skipped
    (5000, 30, 7, 40, 9.0),     # clusters closer together than their radius
    (4000, 8, 10, 1, 1.0),      # one Gaussian blob: hardly anything skipped
    (3000, 60, 20, 12, 9.0),    # k > 16 (32-entry bound network), 57-64 coordinates in registers
    (2500, 100, 25, 8, 12.0),   # more than 64 coordinates (the generic assignment and bound kernels)
])
def test_knn_pruned_exact_vs_oracle(knn_prune, precision, n, d, k, comps, spread):
    # block pruning (k-means grouped layout, bounding spheres) keeps the brute-force result:
    # index sets and FP64 distances bit-exact, and equal to the unpruned search
    x, _ = W.gaussian_mixture(n, d, comps, spread, seed=n + 7 * d)
    knn_prune("1")
    idx, d2, st = scgpu.knn(x, k, precision=precision, return_stats=True)
    knn_prune("0")
    idx0, d20, st0 = scgpu.knn(x, k, precision=precision, return_stats=True)
    assert np.array_equal(idx, idx0) and np.array_equal(d2.view(np.uint64), d20.view(np.uint64))
    rows = np.random.default_rng(1).choice(n, 300, replace=False)
    oi, od = O.knn_bruteforce(x, k, rows=rows)
    assert np.array_equal(idx[rows], oi)
    assert np.array_equal(d2[rows].view(np.uint64), od.view(np.uint64))
    print(f"pruned {precision} n={n} d={d}: pairs {st['pairs_computed']:.3g} of {st0['pairs_computed']:.3g}, "
          f"fallback rows {st['rows_fallback']}")


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("d,offset", [(12, 400.0), (30, 150.0), (50, 100.0)])
def test_knn_pruned_far_from_origin(knn_prune, precision, d, offset):
    # clusters a few units wide at a distance of hundreds from the origin: |x|^2 - 2 x.c + |c|^2
    # cancels to a small fraction of its terms, so the f32 and split-BF16 bounds of the block
    # filter and the tile (scaled by the norms) must hold where they are loosest
    x, _ = W.gaussian_mixture(3000, d, 20, 9.0, seed=d)
    x = (x + np.float32(offset)).astype(np.float32)
    knn_prune("1")
    idx, d2, st = scgpu.knn(x, 9, precision=precision, return_stats=True)
    knn_prune("0")
    idx0, d20 = scgpu.knn(x, 9, precision=precision)
    assert np.array_equal(idx, idx0) and np.array_equal(d2.view(np.uint64), d20.view(np.uint64))
    rows = np.random.default_rng(2).choice(3000, 200, replace=False)
    oi, od = O.knn_bruteforce(x, 9, rows=rows)
    assert np.array_equal(idx[rows], oi) and np.array_equal(d2[rows].view(np.uint64), od.view(np.uint64))
    print(f"{precision} offset {offset} d={d}: pairs {st['pairs_computed']:.3g}, fallback rows {st['rows_fallback']}")


def test_knn_pruned_ties_duplicates_default_size():
    # above the default pruning size: duplicated points and exact ties across groups
    rng = np.random.default_rng(5)
    base, _ = W.gaussian_mixture(20000, 16, 25, 16.0, seed=9)
    x = base.copy()
    x[1000:1300] = x[5000:5300]                      # exact duplicates far apart in input order
    x[7000:7400] = np.round(x[7000:7400])            # rounded rows: many equal distances
    idx, d2, st = scgpu.knn(x, 9, return_stats=True)
    assert st["pairs_computed"] < 0.75 * 20224 ** 2  # the pruned path ran and skipped work
    rows = np.concatenate([np.arange(1000, 1010), np.arange(5000, 5010), np.arange(7000, 7020),
                           rng.choice(20000, 150, replace=False)])
    oi, od = O.knn_bruteforce(x, 9, rows=rows)
    assert np.array_equal(idx[rows], oi) and np.array_equal(d2[rows], od)


def test_knn_ties_and_duplicates():
    rng = np.random.default_rng(1)
    x = rng.integers(0, 3, size=(500, 4)).astype(np.float32)  # many exact ties
    idx, d2 = scgpu.knn(x, 6)
    oi, od = O.knn_bruteforce(x, 6)
    assert np.array_equal(idx, oi) and np.array_equal(d2, od)


@pytest.mark.parametrize("mode", [0, 1])
def test_gaussian_graph_vs_oracle(mode):
    x, _ = W.gaussian_mixture(1500, 20, 8, 9.0, seed=3)
    oi, od = O.knn_bruteforce(x, 7)
    op, on, ow = O.knn_gaussian(oi, od, mode=mode)
    gp, gn, gw = scgpu.knn_gaussian_csr(oi, od, mode=mode)
    assert np.array_equal(gp, op) and np.array_equal(gn, on)
    assert np.allclose(gw, ow, rtol=1e-6, atol=0)
    g = scgpu.knn_graph(x, 7, mode=mode)
    p, n_, w = g.to_csr_parts()
    assert np.array_equal(p, op) and np.array_equal(n_, on) and np.allclose(w, ow, rtol=1e-6, atol=0)
    g.validate_symmetry()


@pytest.mark.parametrize("n,d,k", [(3000, 30, 9), (20000, 16, 7)])
def test_knn_graph_device_input_equals_host_input(n, d, k):
    # points resident in device memory give the same graph as points handed over from the host
    # (the second size runs the block-pruned search)
    x, _ = W.gaussian_mixture(n, d, 8, 9.0, seed=n)
    xd = scgpu.DeviceArray(x.astype(np.float32))
    g1, st1 = scgpu.knn_graph_device(xd.ptr, n, d, k, return_stats=True)
    g0, st0 = scgpu.knn_graph(x, k, return_stats=True)
    xd.free()
    for a, b in zip(g1.to_csr_parts(), g0.to_csr_parts()):
        assert np.array_equal(a, b)
    assert st1["pairs_computed"] == st0["pairs_computed"]


def test_gaussian_mode1_by_hand_on_the_gpu():
    """The hand-worked six-point line of the oracle's mode-1 test, end to end on the GPU (kNN,
    sigma, emission, CSR): the same pairs and weights."""
    from test_oracle import LINE_PAIRS_MODE1, LINE_POINTS, LINE_SIGMA_MEDIAN, _gauss

    g = scgpu.knn_graph(LINE_POINTS, 4, mode=1)
    p, nb, w = g.to_csr_parts()
    got = {(i, int(nb[e])): float(w[e]) for i in range(6) for e in range(p[i], p[i + 1])}
    assert set(got) == LINE_PAIRS_MODE1 | {(b, a) for a, b in LINE_PAIRS_MODE1}
    x = LINE_POINTS[:, 0].astype(np.float64)
    for (i, j), wt in got.items():
        ref = np.float32(_gauss(LINE_SIGMA_MEDIAN[i], LINE_SIGMA_MEDIAN[j], (x[i] - x[j]) ** 2))
        assert abs(wt - ref) <= 1e-6 * ref, (i, j, wt, ref)


@pytest.mark.parametrize("n,d,k", [(3000, 50, 15), (2500, 30, 7), (1200, 100, 20)])
def test_knn_fp64_fallback_scan_vs_oracle(monkeypatch, n, d, k):
    # every row through the exact FP64 scan that serves rows the certificate rejects: its index
    # sets and distances must equal the brute force bit for bit, ties to the lower index
    x, _ = W.gaussian_mixture(n, d, 8, 9.0, seed=3 * n + d)
    x[5] = x[17]  # an exact duplicate: a tie at distance 0 for two rows
    monkeypatch.setenv("SCGPU_KNN_FALLBACK_ALL", "1")
    idx, d2, st = scgpu.knn(x, k, return_stats=True)
    assert st["rows_fallback"] == n
    oi, od = O.knn_bruteforce(x, k)
    assert np.array_equal(idx, oi)
    assert np.array_equal(d2.view(np.uint64), od.view(np.uint64))
