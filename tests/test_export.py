"""Scanpy CSR export reader and the real-data comparison's partition metrics (host-only parts;
the device ingest behind the reader is covered by the GPU graph tests)."""
import os

import numpy as np
import pytest
import sklearn.metrics as skm

import real_data as RD
import scgpu


def _write(d, ptr_bytes=None, idx=None, data=None):
    os.makedirs(d, exist_ok=True)
    if ptr_bytes is not None:
        open(os.path.join(d, "indptr.bin"), "wb").write(ptr_bytes)
    if idx is not None:
        np.asarray(idx, dtype="<u4").tofile(os.path.join(d, "indices.bin"))
    if data is not None:
        np.asarray(data, dtype="<f4").tofile(os.path.join(d, "data.bin"))


def _code(directory):
    with pytest.raises(scgpu.ClusteringError) as ei:
        scgpu.CSRNetwork.from_csr_export(str(directory))
    return ei.value.code, str(ei.value)


def test_export_reader_file_errors_before_device_work(tmp_path):
    # a missing file is an invalid argument; lengths that do not fit together are InvalidCsr,
    # all reported on the host (these calls never reach the GPU)
    code, msg = _code(tmp_path / "nowhere")
    assert code == 101 and "indptr.bin" in msg
    ok_ptrs = np.array([0, 1, 2], dtype="<u8").tobytes()
    _write(tmp_path / "a", ok_ptrs, [1, 0], [1.0, 1.0])
    os.remove(tmp_path / "a" / "data.bin")
    code, msg = _code(tmp_path / "a")
    assert code == 101 and "data.bin" in msg
    _write(tmp_path / "b", ok_ptrs[:-1], [1, 0], [1.0, 1.0])          # 23 bytes of u64
    code, msg = _code(tmp_path / "b")
    assert code == 9 and "malformed CSR input" in msg and "8-byte" in msg
    _write(tmp_path / "c", ok_ptrs, [1, 0], [1.0, 1.0, 2.0])          # 2 ids, 3 weights
    code, msg = _code(tmp_path / "c")
    assert code == 9 and "indices.bin has 2 entries but data.bin has 3" in msg
    _write(tmp_path / "d", b"", [], [])                               # no row offsets at all
    code, msg = _code(tmp_path / "d")
    assert code == 9 and "empty" in msg
    _write(tmp_path / "e", ok_ptrs, [1, 0], [1.0, 1.0])
    open(tmp_path / "e" / "indices.bin", "wb").write(b"\x01\x00\x00")  # 3 bytes of u32
    code, msg = _code(tmp_path / "e")
    assert code == 9 and "4-byte" in msg


def test_write_csr_export_layout(tmp_path):
    # the export tool's byte layout: u64 / u32 / f32 little-endian, nothing else in the files
    ptrs = np.array([0, 2, 3, 4], dtype=np.uint64)
    idx = np.array([1, 2, 0, 0], dtype=np.uint32)
    w = np.array([0.5, 1.5, 0.5, 1.5], dtype=np.float32)
    scgpu.write_csr_export(str(tmp_path), ptrs, idx, w)
    raw = open(tmp_path / "indptr.bin", "rb").read()
    assert raw == b"".join(int(p).to_bytes(8, "little") for p in ptrs)
    assert open(tmp_path / "indices.bin", "rb").read() == b"".join(int(i).to_bytes(4, "little") for i in idx)
    assert np.array_equal(np.fromfile(tmp_path / "data.bin", dtype="<f4"), w)


def test_partition_metrics_match_sklearn_and_closed_forms():
    rng = np.random.default_rng(3)
    for n, ka, kb in [(50, 3, 4), (400, 12, 7), (1000, 30, 30)]:
        a = rng.integers(0, ka, n)
        b = rng.integers(0, kb, n)
        assert RD.adjusted_rand_index(a, b) == pytest.approx(skm.adjusted_rand_score(a, b), abs=1e-12)
        assert RD.nmi(a, b) == pytest.approx(
            skm.normalized_mutual_info_score(a, b, average_method="arithmetic"), abs=1e-12)
    a = rng.integers(0, 9, 300)
    perm = rng.permutation(9)
    assert RD.adjusted_rand_index(a, perm[a] + 100) == pytest.approx(1.0)  # relabelling
    assert RD.nmi(a, perm[a]) == pytest.approx(1.0)
    assert RD.adjusted_rand_index([], []) == 1.0 and RD.nmi([], []) == 1.0
    # a single cluster against itself: degenerate denominator (real_data.rs:71-73)
    assert RD.adjusted_rand_index([0] * 5, [0] * 5) == 1.0
    assert RD.nmi([0] * 5, [0] * 5) == 1.0
    # 2x2 known answer: {0,0,1,1} vs {0,1,0,1} is -0.5
    assert RD.adjusted_rand_index([0, 0, 1, 1], [0, 1, 0, 1]) == pytest.approx(-0.5)
