"""The report's nearest-centroid step on the GPU (src/report.rs:169-183):
indices and distances bit-identical to the CPU oracle's restatement, on the
reference test's case, every tile / stage / remainder shape, ties, empty
clusters and NaN centroids, through the host and the device entry points."""
import numpy as np
import pytest

import emeans_hip as E
import oracle as O

pytestmark = pytest.mark.gpu


def _same(c, sizes=None):
    g = E.report_nearest(c, sizes)
    o = O.report_nearest(c, sizes)
    assert np.array_equal(g[0], o[0]), np.flatnonzero(g[0] != o[0])[:10]
    assert g[1].tobytes() == o[1].tobytes()
    return g


def test_reference_case(hip):
    # src/report.rs:283-312
    c = np.array([[0.0], [1.0], [50.0]], np.float32)
    nearest, dist = _same(c, np.array([101, 100, 3], np.uint32))
    assert nearest.tolist() == [1, 0, 1] and dist.tolist() == [1.0, 1.0, 49.0]


def test_order_known_answers(hip):
    # the two hand-derived f32 values the oracle's own test pins (16777218)
    a = np.zeros((2, 32), np.float32)
    a[0, [0, 2, 16, 18]] = [4096, 1, 1, 1]
    b = np.zeros((2, 34), np.float32)
    b[0, [0, 1, 32, 33]] = [4096, 1, 1, 1]
    for c in (a, b):
        nearest, dist = E.report_nearest(c)
        assert nearest.tolist() == [1, 0]
        assert dist[0] == np.float32(np.sqrt(np.float32(16777218.0)))


# d: remainder only (5, 15), one chunk (16), one full stage (128), a stage
# and a partial one plus a remainder (200), several stages (300); k: one
# tile, partial tiles, many tiles
@pytest.mark.parametrize("k,d", [(2, 5), (40, 15), (33, 16), (100, 37), (257, 128), (1000, 200), (700, 300)])
def test_matches_cpu(hip, k, d):
    print(f"report nearest k={k} d={d}", flush=True)
    rng = np.random.default_rng(k * 1000 + d)
    c = rng.normal(size=(k, d)).astype(np.float32)
    _same(c)
    sizes = rng.integers(0, 3, size=k).astype(np.uint32)
    _same(c, sizes)


def test_ties_duplicates_and_integers(hip):
    rng = np.random.default_rng(7)
    c = rng.integers(-2, 3, size=(300, 20)).astype(np.float32)
    c[10] = c[250]
    c[11] = c[12]
    sizes = rng.integers(0, 4, size=300).astype(np.uint32)
    _same(c)
    _same(c, sizes)


def test_empty_and_nan(hip):
    c = np.arange(12, dtype=np.float32).reshape(4, 3)
    n, d = _same(c, np.zeros(4, np.uint32))
    assert n.tolist() == [0, 1, 2, 3] and np.all(np.isinf(d))
    n, d = _same(np.ones((1, 4), np.float32))
    assert n.tolist() == [0] and np.isinf(d[0])
    c[2, 1] = np.nan
    n, d = _same(c)
    assert n.tolist() == [1, 0, 2, 1] and np.isinf(d[2])
    big = np.random.default_rng(3).normal(size=(90, 40)).astype(np.float32)
    big[[5, 60]] = np.nan
    _same(big)


def test_device_entry_point(hip):
    import torch

    rng = np.random.default_rng(11)
    k, d = 3000, 128
    c = rng.normal(size=(k, d)).astype(np.float32)
    sizes = rng.integers(0, 5, size=k).astype(np.uint32)
    cd = torch.from_numpy(c).cuda()
    sd = torch.from_numpy(sizes.view(np.int32)).cuda()
    keys = torch.empty(k, dtype=torch.int64, device="cuda")
    nearest = torch.empty(k, dtype=torch.int32, device="cuda")
    dist = torch.empty(k, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    E.report_nearest_device(cd, sd, keys, nearest, dist)
    torch.cuda.synchronize()
    o = O.report_nearest(c, sizes)
    assert np.array_equal(nearest.cpu().numpy().view(np.uint32), o[0])
    assert dist.cpu().numpy().tobytes() == o[1].tobytes()
    # a second call on the same workspace with fewer centroids
    E.report_nearest_device(cd[:1000], None, keys[:1000], nearest[:1000], dist[:1000])
    torch.cuda.synchronize()
    o = O.report_nearest(c[:1000])
    assert np.array_equal(nearest[:1000].cpu().numpy().view(np.uint32), o[0])
    assert dist[:1000].cpu().numpy().tobytes() == o[1].tobytes()


def test_rejects_bad_shapes(hip):
    with pytest.raises(ValueError):
        E.report_nearest(np.zeros(5, np.float32))
    with pytest.raises(E.EmeansError):
        E.report_nearest(np.zeros((0, 4), np.float32))


# ---- report::report around the two device steps ------------------------------
def _reference_report_case():
    labels, sq = [], []
    for i in range(200):
        off = np.float32(0.001) * np.float32(i % 100)
        labels.append(i % 2)
        sq.append(off * off)
    labels.append(0)
    sq.append(np.float32(0.45) * np.float32(0.45))
    for _ in range(3):
        labels.append(2)
        sq.append(np.float32(1e-4))
    return np.array(labels, np.uint32), np.array(sq, np.float32), np.array([[0.0], [1.0], [50.0]], np.float32)


def test_report_reference_test(hip):
    # src/report.rs:283-312, assertion by assertion
    import io

    labels, sq, c = _reference_report_case()
    r = E.report(labels, sq, c)
    assert r.clusters["size"].tolist() == [101, 100, 3]
    assert r.top_outliers(1)[0][0] == 200
    assert r.local_z[200] > 10.0
    assert len(r.families) == 2
    assert r.families[0].size == 201
    assert r.clusters["family"][0] == r.clusters["family"][1]
    assert r.rare_prototypes(0.25, 2.0) == [(2, 3, np.float32(49.0))]
    buf = io.StringIO()
    r.write_clusters_tsv(buf)
    assert buf.getvalue().count("\n") == 4


def _compare_report(r, o):
    for f in r.clusters.dtype.names:
        assert r.clusters[f].tobytes() == o["clusters"][f].tobytes(), f
    assert r.local_z.tobytes() == o["local_z"].tobytes()
    assert r.global_z.tobytes() == o["global_z"].tobytes()
    for a, b in (("median_dist", "median_dist"), ("median_rms", "median_rms"),
                 ("median_nearest_dist", "median_nearest_dist")):
        assert np.float32(getattr(r, a)).tobytes() == np.float32(o[b]).tobytes(), a
    assert len(r.families) == len(o["families"])
    for f, g in zip(r.families, o["families"]):
        assert (f.leader, f.members.tolist(), f.size, f.share) == (g["leader"], g["members"], g["size"], g["share"])


@pytest.mark.parametrize("radius", [0.5, 1.25, 3.0])
@pytest.mark.parametrize("n,k,d", [(3000, 40, 8), (20000, 150, 37)])
def test_report_matches_cpu(hip, n, k, d, radius):
    print(f"report n={n} k={k} d={d} radius={radius}", flush=True)
    rng = np.random.default_rng(n + k + d)
    c = rng.normal(size=(k, d)).astype(np.float32)
    # uneven cluster sizes, some clusters empty, a NaN and a negative distance
    w = rng.pareto(1.5, size=k)
    w[rng.choice(k, size=k // 10, replace=False)] = 0
    labels = rng.choice(k, size=n, p=w / w.sum()).astype(np.uint32)
    sq = (rng.gamma(2.0, 1.0, size=n) * rng.uniform(0.5, 2.0, size=k)[labels]).astype(np.float32)
    sq[0], sq[1] = np.nan, -0.5
    cfg = E.ReportConfig(family_radius=radius, min_local_size=30)
    r = E.report(labels, sq, c, cfg)
    o = O.report(labels, sq, c, family_radius=radius, min_local_size=30)
    _compare_report(r, o)


def test_report_rejects_bad_labels(hip):
    with pytest.raises(E.EmeansError):
        E.report(np.array([0, 3], np.uint32), np.ones(2, np.float32), np.zeros((3, 2), np.float32))
