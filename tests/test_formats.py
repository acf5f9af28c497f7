"""The examples' on-disk formats (examples/cluster.rs:11-14, 22-23;
examples/outliers.rs:70-71): raw little-endian f32 rows / centroids and u32
labels, no header."""
import numpy as np

import emeans_hip as E


def test_f32_rows_are_raw_little_endian(tmp_path):
    p = tmp_path / "x.f32"
    x = np.array([[1.0, -2.0, 0.5], [3.0, 0.0, -0.25]], np.float32)
    E.write_f32(p, x)
    raw = p.read_bytes()
    assert len(raw) == 24
    assert raw[:4] == b"\x00\x00\x80\x3f" and raw[4:8] == b"\x00\x00\x00\xc0"  # 1.0, -2.0
    back = E.read_f32(p, 3)
    assert back.shape == (2, 3) and np.array_equal(back, x)
    # a big-endian array is written little-endian
    E.write_f32(p, x.astype(">f4"))
    assert p.read_bytes() == raw


def test_u32_labels_round_trip(tmp_path):
    p = tmp_path / "out.labels.u32"
    labels = np.array([0, 1, 70000, 2**32 - 1], np.uint32)
    E.write_u32(p, labels)
    assert p.read_bytes()[:8] == b"\x00\x00\x00\x00\x01\x00\x00\x00"
    assert np.array_equal(E.read_u32(p), labels)


def test_read_f32_rejects_ragged_files(tmp_path):
    p = tmp_path / "bad.f32"
    E.write_f32(p, np.zeros(7, np.float32))
    try:
        E.read_f32(p, 3)
    except ValueError:
        return
    raise AssertionError("a file of 7 floats is not a multiple of d = 3")
