"""The learned active-vertex filter: HIP forward/backward against the float64 oracle,
training, trace collection, and filtered Leiden (no counterpart in the reference)."""
import numpy as np
import pytest

import mlp_oracle as M
import scgpu
import workloads as W

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((n, 8)).astype(np.float32)
    y = (rng.random(n) < 0.3).astype(np.uint8)
    return X, y


def test_hip_gradient_matches_oracle():
    w = scgpu.mlp_init(5)
    X, y = _data(3000, 1)
    for pw in (1.0, 4.0):
        loss, g = scgpu.mlp_grad(w, X, y, pw)
        g_ref = M.grad(w.astype(np.float64), X, y, pw)
        assert abs(loss - M.loss(w, X, y, pw)) <= 1e-5 * max(1.0, abs(loss))
        assert np.max(np.abs(g - g_ref)) <= 2e-5 * max(1.0, np.max(np.abs(g_ref))), np.max(np.abs(g - g_ref))


def test_predict_matches_oracle_fp32_and_bf16():
    w = scgpu.mlp_init(9)
    X, _ = _data(5000, 2)
    z32 = scgpu.mlp_predict(w, X, bf16=False)
    z_ref, _ = M.forward(w, X)
    assert np.max(np.abs(z32 - z_ref)) <= 1e-4 * max(1.0, np.max(np.abs(z_ref)))
    # BF16-rounded weights (round to nearest even), FP32 accumulation
    u = w.view(np.uint32).astype(np.uint64)
    wb = (((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)).view(np.float32)
    zb = scgpu.mlp_predict(w, X, bf16=True)
    zb_ref, _ = M.forward(wb, X)
    assert np.max(np.abs(zb - zb_ref)) <= 1e-4 * max(1.0, np.max(np.abs(zb_ref)))


def test_training_learns_a_separable_rule():
    rng = np.random.default_rng(4)
    X = rng.random((20000, 8)).astype(np.float32)
    y = (X[:, 0] + 0.5 * X[:, 3] > 0.8).astype(np.uint8)
    w0 = scgpu.mlp_init(1)
    l0, _ = scgpu.mlp_grad(w0, X, y)
    w, hist = scgpu.mlp_train(X, y, epochs=15, batch=1024, lr=3e-3, seed=1, w=w0)
    assert hist[-1] < 0.5 * l0 and hist[-1] < hist[0]
    acc = np.mean((scgpu.mlp_predict(w, X) > 0) == (y == 1))
    assert acc > 0.95, acc
    w2, hist2 = scgpu.mlp_train(X, y, epochs=15, batch=1024, lr=3e-3, seed=1, w=w0)
    assert np.array_equal(w, w2) and np.array_equal(hist, hist2)  # deterministic


def _knn(n, seed):
    x, _ = W.gaussian_mixture(n, 16, 10, 9.0, seed=seed)
    return scgpu.knn_graph(x, 10)


def test_trace_and_filtered_leiden():
    g_train, g = _knn(6000, 11), _knn(8000, 12)
    X, y = scgpu.leiden_trace(g_train, scgpu.LeidenConfig(seed=3))
    assert len(X) > 1000 and X.shape[1] == 8 and 0 < y.mean() < 1
    assert np.all(np.isfinite(X))
    w, _ = scgpu.mlp_train(X, y, epochs=20, batch=2048, lr=3e-3, seed=2)
    plain = scgpu.leiden(g)
    q_plain = scgpu.modularity(g, plain.labels())
    # a filter that never skips reproduces plain Leiden exactly, unless the judge is a full
    # sweep (which also re-decides vertices plain local moving left inactive)
    never = w.copy()
    never[-1] = 1e4
    for judge in ("none", "skipped"):
        c, st = scgpu.leiden_filtered(g, never, 0.5, judge=judge)
        assert np.array_equal(c.labels(), plain.labels()) and st["skipped"] == 0
    c, st = scgpu.leiden_filtered(g, never, 0.5, judge="full")
    assert st["skipped"] == 0 and st["skipped_at_judge"] == 0 and st["judged"] > 0
    assert scgpu.modularity(g, c.labels()) >= q_plain - 1e-3 * abs(q_plain)
    for judge in ("full", "skipped"):
        c, st = scgpu.leiden_filtered(g, w, 0.05, judge=judge)
        q = scgpu.modularity(g, c.labels())
        print(f"filter ({judge} judge): skipped {st['skipped']}, judged {st['judged']}, "
              f"skipped at judge {st['skipped_at_judge']}, moved {st['skipped_moved']} "
              f"(rate {st['skip_error_rate']:.4f}); modularity {q:.6f} vs plain {q_plain:.6f}")
        assert q >= q_plain - 0.01 * abs(q_plain)
        assert 0.0 <= st["skip_error_rate"] <= 1.0 and st["skipped_moved"] <= st["skipped_at_judge"]
        assert np.array_equal(scgpu.leiden_filtered(g, w, 0.05, judge=judge)[0].labels(), c.labels())
    # without the judge nothing re-decides the skipped vertices: the result rests on the model
    # alone (a sanity bound here), and is still deterministic
    cf, fst = scgpu.leiden_filtered(g, w, 0.05, judge="none")
    qf = scgpu.modularity(g, cf.labels())
    print(f"filter (no judge): skipped {fst['skipped']}, modularity {qf:.6f}")
    assert fst["judged"] == 0 and fst["skipped"] > 0
    assert qf >= q_plain - 0.1 * abs(q_plain)
    assert np.array_equal(scgpu.leiden_filtered(g, w, 0.05, judge="none")[0].labels(), cf.labels())
    with pytest.raises(ValueError):
        scgpu.leiden_filtered(g, w, 0.05, judge="sometimes")


def test_trained_filter_finds_non_movers_far_better_than_chance():
    """What the trained network gains over the method without it: on the decisions of a graph it
    was not trained on, the vertices it would skip (move probability below the threshold) move
    far less often than local moving's decisions do overall, so a filtered run skips a real share
    of decisions at equal modularity (within 0.1%), with the full plain judge at every level."""
    g_train, g = _knn(20000, 21), _knn(20000, 22)
    cfg = scgpu.LeidenConfig(seed=3)
    X, y = scgpu.leiden_trace(g_train, cfg)
    w, _ = scgpu.mlp_train(X, y, epochs=12, batch=4096, lr=3e-3, seed=2)
    Xh, yh = scgpu.leiden_trace(g, cfg)  # held-out decisions
    p = 1.0 / (1.0 + np.exp(-scgpu.mlp_predict(w, Xh).astype(np.float64)))
    thr = 0.05
    skip = p < thr
    base = float(yh.mean())
    wrong = float(yh[skip].mean()) if skip.any() else 0.0
    print(f"held-out decisions {len(yh)}: move rate {base:.4f}; the filter would skip {skip.mean():.3f} "
          f"of them, of which {wrong:.4f} move")
    assert skip.mean() >= 0.2, skip.mean()
    assert wrong <= 0.1 * base, (wrong, base)
    plain = scgpu.leiden(g, cfg)
    q_plain = scgpu.modularity(g, plain.labels())
    c, st = scgpu.leiden_filtered(g, w, thr, cfg, judge="full")
    q = scgpu.modularity(g, c.labels())
    print(f"filtered: skipped {st['skipped']}, judge decisions {st['judged']}, skip-error rate "
          f"{st['skip_error_rate']:.4f}; modularity {q:.8f} vs plain {q_plain:.8f}")
    assert st["skipped"] > 0.05 * len(yh)
    assert abs(q - q_plain) <= 1e-3 * abs(q_plain)


def test_trace_is_reproducible():
    g = _knn(5000, 13)
    X1, y1 = scgpu.leiden_trace(g, scgpu.LeidenConfig(seed=5))
    X2, y2 = scgpu.leiden_trace(g, scgpu.LeidenConfig(seed=5))
    assert np.array_equal(X1, X2) and np.array_equal(y1, y2)


def test_bf16_inference_wave_layout():
    """The MFMA inference's packed layout, by impulses: with one nonzero path through the network
    (feature f -> hidden i -> hidden j -> output, or bias b1[i] -> j with bias b2[j]) the logit has
    a closed form, which fails if a fragment's row, column or k order is transposed. Weights are
    exact in BF16."""
    W1, B1, W2, B2, W3, B3 = 0, 512, 576, 4672, 4736, 4800
    rng = np.random.default_rng(8)
    X = rng.random((300, 8)).astype(np.float32)
    x64 = X.astype(np.float64)
    for case in range(12):
        f, i, j = int(rng.integers(8)), int(rng.integers(64)), int(rng.integers(64))
        w = np.zeros(4801, dtype=np.float32)
        w[W3 + j], w[B3] = 0.5, 0.25
        if case % 2 == 0:
            w[W1 + f * 64 + i], w[W2 + i * 64 + j] = 1.5, 0.75
            want = 0.25 + 0.5 * (1.5 * 0.75) * x64[:, f]
        else:
            w[B1 + i], w[W2 + i * 64 + j], w[B2 + j] = 0.5, 1.25, -0.25
            w[W1 + f * 64 + i] = -1.0   # h1[i] = relu(0.5 - x_f)
            want = 0.25 + 0.5 * np.maximum(1.25 * np.maximum(0.5 - x64[:, f], 0) - 0.25, 0)
        z = scgpu.mlp_predict(w, X, bf16=True)
        assert np.max(np.abs(z - want)) <= 1e-5, (case, f, i, j, np.max(np.abs(z - want)))
