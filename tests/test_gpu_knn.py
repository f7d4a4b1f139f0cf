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
