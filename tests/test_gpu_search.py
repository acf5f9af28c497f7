"""GPU parity of the search kernels against the CPU oracle.

FP32: the reference's own bounds (src/cpu.rs:298-306 vs FP64 brute force;
src/cuda.rs:244-248 vs the CPU search). I8: bit-exact labels and distances
vs the CpuI8Assigner restatement (src/cpu_i8.rs:294-324). BF16/FP8 + FP32
rescore: label disagreement and distance error vs FP64 brute force, with the
tolerances written in each test.
"""
import functools

import numpy as np
import pytest

import emeans_hip as E
import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 16), (37, 70, 33), (130, 257, 100), (64, 1030, 384), (9, 6, 1027)]
I8_SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 64), (37, 70, 33), (130, 257, 100), (20, 1030, 512), (9, 6, 1027)]


def _bound_ok(x, c, labels, dist, bias=None):
    """src/cpu.rs:298-306: chosen centroid within 1e-4(1+best) of the FP64
    best, and the returned distance within the same bound."""
    bl, bd = O.brute_f64(x, c, bias)
    got = O.dist_f64(x, c, labels)
    if bias is not None:
        got = got + bias[labels].astype(np.float64)
    tol = 1e-4 * (1.0 + np.abs(bd))
    return bool(np.all(got <= bd + tol)), bool(np.all(np.abs(dist.astype(np.float64) - bd) <= tol))


@pytest.mark.parametrize("n,k,d", SHAPES)
def test_fp32_matches_brute_force(hip, n, k, d):
    print(f"fp32 n={n} k={k} d={d}", flush=True)
    x = O.lcg_f32(n * d, 1).reshape(n, d)
    c = O.lcg_f32(k * d, 2).reshape(k, d)
    a = E.HipAssigner(E.FP32)
    a.set_centroids(c)
    labels, dist = a.assign(x)
    assert labels.max() < k
    lab_ok, dist_ok = _bound_ok(x, c, labels, dist)
    assert lab_ok and dist_ok


def test_fp32_matches_cpu(hip):
    # src/cuda.rs:221-249
    n, k, d = 3000, 1031, 100
    data = O.lcg_f32(n * d + k * d, 1)  # one generator: rows, then centroids
    x = data[: n * d].reshape(n, d)
    c = data[n * d:].reshape(k, d)
    cl, cd = O.cpu_assign(x, c)
    a = E.HipAssigner(E.FP32)
    a.set_centroids(c)
    gl, gd = a.assign(x)
    same = int((cl == gl).sum())
    assert same >= 0.999 * n, f"only {same} of {n} labels agree"
    assert np.all(np.abs(cd - gd) <= 1e-3 * (1.0 + np.abs(cd)))


@pytest.mark.parametrize("rows", ["i8", "u8"])
def test_fp32_int_rows_equal_f32_rows(hip, rows):
    rng = np.random.default_rng(5)
    xi = (rng.integers(-128, 128, size=(300, 40)) if rows == "i8" else rng.integers(0, 256, size=(300, 40)))
    xi = xi.astype(np.int8 if rows == "i8" else np.uint8)
    c = rng.normal(scale=40, size=(70, 40)).astype(np.float32)
    a = E.HipAssigner(E.FP32)
    a.set_centroids(c)
    l1, d1 = a.assign(xi)
    l2, d2 = a.assign(xi.astype(np.float32))
    assert np.array_equal(l1, l2) and np.array_equal(d1, d2)


def test_fp32_bias(hip):
    rng = np.random.default_rng(6)
    x = rng.normal(size=(500, 24)).astype(np.float32)
    c = rng.normal(size=(90, 24)).astype(np.float32)
    bias = rng.uniform(0, 3, size=90).astype(np.float32)
    a = E.HipAssigner(E.FP32)
    a.set_centroids(c)
    a.set_bias(bias)
    labels, dist = a.assign(x)
    lab_ok, dist_ok = _bound_ok(x, c, labels, dist, bias)
    assert lab_ok and dist_ok
    # set_centroids resets the bias (src/lib.rs:78-81)
    a.set_centroids(c)
    l2, d2 = a.assign(x)
    assert _bound_ok(x, c, l2, d2) == (True, True)


@pytest.mark.parametrize("n,k,d", I8_SHAPES)
def test_i8_bit_exact(hip, n, k, d):
    print(f"i8 n={n} k={k} d={d}", flush=True)
    x = O.lcg_i8(n * d, 1).reshape(n, d)
    ci = O.lcg_i8(k * d, 2).reshape(k, d).astype(np.float32)
    el, ed = O.brute_i64(x, ci.astype(np.int8))
    a = E.HipAssigner(E.I8)
    a.set_centroids(ci)
    labels, dist = a.assign(x)
    assert np.array_equal(labels, el)
    assert np.array_equal(dist, ed.astype(np.float32))


def test_i8_fractional_centroids_and_bias_match_cpu(hip):
    rng = np.random.default_rng(7)
    x = rng.integers(-128, 128, size=(2000, 96)).astype(np.int8)
    c = rng.normal(scale=60, size=(333, 96)).astype(np.float32)
    c[:5] = np.round(c[:5]) + 0.5  # exact halves round away from zero
    c[5, :3] = [300.0, -300.0, np.float32(127.49)]
    bias = rng.uniform(-50, 50, size=333).astype(np.float32)
    bias[:4] = [0.5, -0.5, 2.5, -2.5]
    a = E.HipAssigner(E.I8)
    for b in (None, bias):
        a.set_centroids(c)
        if b is not None:
            a.set_bias(b)
        gl, gd = a.assign(x)
        cl, cd = O.i8_assign(x, c, bias=b)
        assert np.array_equal(gl, cl) and np.array_equal(gd, cd)


@pytest.mark.parametrize("d", [64, 512])
def test_i8_large_bias_takes_the_i64_path(hip, d):
    """Biases beyond the packed int32 key range switch the search to the i64
    kernel; results must stay bit-exact either way."""
    rng = np.random.default_rng(9)
    x = rng.integers(-128, 128, size=(700, d)).astype(np.int8)
    c = rng.normal(scale=40, size=(150, d)).astype(np.float32)
    for scale in (1.0, 3e7):
        bias = (rng.uniform(0, 1, size=150) * scale).astype(np.float32)
        bias[3] = -bias[3]
        a = E.HipAssigner(E.I8)
        a.set_centroids(c)
        a.set_bias(bias)
        gl, gd = a.assign(x)
        cl, cd = O.i8_assign(x, c, bias=bias)
        assert np.array_equal(gl, cl) and np.array_equal(gd, cd), f"d={d} bias scale {scale}"


@pytest.mark.parametrize("d", [96, 512])
def test_i16_large_bias_takes_the_i64_path(hip, d):
    """I16 offsets past 2^46 (sum c16^2 + round(bias 4^shift)) leave the
    register kernel's exact f64 keys for the LDS kernel's i64 scores; both
    stay bit-exact, with and without a split of the centroid range."""
    rng = np.random.default_rng(10)
    x = rng.integers(-128, 128, size=(700, d)).astype(np.int8)
    c = rng.normal(scale=40, size=(300, d)).astype(np.float32)
    for scale in (1.0, 1e10):
        bias = (rng.uniform(0, 1, size=300) * scale).astype(np.float32)
        bias[3] = -bias[3]
        cl, cd = O.i8_assign(x, c, bits=O.BITS_I16, bias=bias)
        for split in (0, 4):
            a = E.HipAssigner(E.I16)
            a.set_split(split)
            a.set_centroids(c)
            a.set_bias(bias)
            gl, gd = a.assign(x)
            assert np.array_equal(gl, cl) and np.array_equal(gd, cd), f"d={d} bias scale {scale} split {split}"


@pytest.mark.parametrize("n,k,d", I8_SHAPES)
def test_i16_bit_exact(hip, n, k, d):
    """src/cpu_i8.rs:298-324 with CentroidBits::I16: integer centroids give
    the exact brute-force labels and distances."""
    print(f"i16 n={n} k={k} d={d}", flush=True)
    x = O.lcg_i8(n * d, 1).reshape(n, d)
    ci = O.lcg_i8(k * d, 2).reshape(k, d).astype(np.float32)
    el, ed = O.brute_i64(x, ci.astype(np.int8))
    a = E.HipAssigner(E.I16)
    a.set_centroids(ci)
    labels, dist = a.assign(x)
    assert np.array_equal(labels, el)
    assert np.array_equal(dist, ed.astype(np.float32))


def test_i16_resolves_fractional_centroids(hip):
    # src/cpu_i8.rs:326-336
    a = E.HipAssigner(E.I16)
    a.set_centroids(np.array([[0.26, 0.0], [-0.26, 0.0]], np.float32))
    labels, dist = a.assign(np.array([[1, 0]], np.int8))
    assert labels[0] == 0
    assert abs(dist[0] - 0.74 * 0.74) < 0.01


@pytest.mark.parametrize("d", [7, 96, 300])
def test_i16_fractional_centroids_and_bias_match_cpu(hip, d):
    rng = np.random.default_rng(12)
    x = rng.integers(-128, 128, size=(1500, d)).astype(np.int8)
    c = rng.normal(scale=50, size=(261, d)).astype(np.float32)
    c[:3] = np.round(c[:3] * 4) / 4 + 1 / 256  # near rounding boundaries of the 1/2^s grid
    bias = rng.uniform(-40, 40, size=261).astype(np.float32)
    a = E.HipAssigner(E.I16)
    for b in (None, bias):
        a.set_centroids(c)
        if b is not None:
            a.set_bias(b)
        gl, gd = a.assign(x)
        cl, cd = O.i8_assign(x, c, bits=O.BITS_I16, bias=b)
        assert np.array_equal(gl, cl) and np.array_equal(gd, cd), f"d={d} bias={b is not None}"


@pytest.mark.parametrize("d", [512, 1024])
def test_i16_dot_wraps_like_the_cpu(hip, d):
    """Centroids past the i8 range: the i32 dot product wraps on the CPU
    (vpdpwssd) and must wrap the same way on the GPU (d = 512: the
    register-resident kernel, where rows of -128 against centroids clamped
    to -32768 sum to exactly 2^31; d = 1024: the LDS kernel)."""
    rng = np.random.default_rng(11)
    n, k = 40, 9
    x = rng.integers(100, 128, size=(n, d)).astype(np.int8)
    x[:5] = -128
    c = np.full((k, d), 600.0, np.float32)
    c[1::2] *= -1
    if d > 512:
        c[:, :5] = rng.normal(size=(k, 5))
    a = E.HipAssigner(E.I16)
    a.set_centroids(c)
    gl, gd = a.assign(x)
    cl, cd = O.i8_assign(x, c, bits=O.BITS_I16)
    assert np.array_equal(gl, cl) and np.array_equal(gd, cd)


def test_i8_rejects_float_rows(hip):
    a = E.HipAssigner(E.I8)
    a.set_centroids(np.zeros((3, 4), np.float32))
    with pytest.raises(E.EmeansError) as ei:
        a.assign(np.zeros((2, 4), np.float32))
    assert ei.value.code == E.E_UNSUPPORTED
    b = E.HipAssigner(E.I16)
    b.set_centroids(np.zeros((3, 4), np.float32))
    with pytest.raises(E.EmeansError) as ei:
        b.assign(np.zeros((2, 4), np.uint8))
    assert ei.value.code == E.E_UNSUPPORTED


def _mixture(n, d, k_blobs, seed, scale=4.0):
    rng = np.random.default_rng(seed)
    centres = rng.normal(scale=scale, size=(k_blobs, d)).astype(np.float32)
    x = centres[rng.integers(0, k_blobs, n)] + rng.normal(size=(n, d)).astype(np.float32)
    return x.astype(np.float32)


CAND_N = 100_000
CAND_CASES = [(k, d, prec, m) for k in (64, 128, 256, 1024, 16384) for d in (64, 128, 256)
              for prec, m in (("bf16", 4), ("fp8", 8))]


@functools.lru_cache(maxsize=1)
def _judged_mixture(k, d):
    """100,000 mixture rows (2048 blobs, or k blobs at k = 16384 as in the
    2M x 128 workload), centroids = the rows at permutation positions 0..k
    (src/lib.rs:375-379), and the FP64 brute-force best distances of the
    judged rows: all of them, or a fixed 20,000-row sample at k = 16384."""
    x = _mixture(CAND_N, d, max(2048, k), 100 + d)
    c = x[O.permutation(1234, CAND_N)[:k]].copy()
    rows = np.arange(CAND_N) if k <= 1024 else np.sort(O.permutation(77, CAND_N)[:20000])
    _, bd = O.brute_f64(x[rows], c)
    return x, c, rows, bd


@pytest.mark.parametrize("k,d,prec,m", CAND_CASES)
def test_rescored_search_vs_fp64(hip, k, d, prec, m):
    """BF16 top-4 and FP8 top-8 + FP32 rescore against FP64 brute force:
    true label errors (a label whose FP64 distance exceeds the best by more
    than 1e-6(1+best)) at most 1e-4 of the judged rows, and the returned
    distance within 1e-4(1+d) of FP64 at the returned label (the FP32
    rescore). The candidate list is the exact m smallest keys, so a row is
    lost only when the true nearest is not among the m best low-precision
    scores."""
    print(f"{prec} top-{m} k={k} d={d}", flush=True)
    x, c, rows, bd = _judged_mixture(k, d)
    a = E.HipAssigner(prec, rescore_m=m)
    a.set_centroids(c)
    labels, dist = a.assign(x)
    got = O.dist_f64(x[rows], c, labels[rows])
    wrong = int((got > bd + 1e-6 * (1 + bd)).sum())
    assert wrong <= 1e-4 * len(rows), f"{prec} top-{m} k={k} d={d}: {wrong} of {len(rows)} true label errors"
    assert np.all(np.abs(dist[rows] - got) <= 1e-4 * (1 + got))


def _near_duplicates(k, d, m, where, seed):
    """k centroids far from `base`, except m of them equal to base up to
    offsets of at most 1.4e-3 in dimension 0, which BF16 (spacing 2^-7 at
    1.0) and e4m3 (2^-3) round away: their low-precision scores tie exactly.
    The truly nearest (offset 0) is the one with the highest index, so only a
    list holding all m of them lets the FP32 rescore find it."""
    rng = np.random.default_rng(seed)
    base = rng.normal(size=d).astype(np.float32)
    base[0] = 1.0
    c = (base + rng.normal(scale=3.0, size=(k, d))).astype(np.float32)
    lane_half0 = [8 * (r >> 2) + (r & 3) for r in range(16)]  # positions one lane sees in a 32-block
    if where == "block":  # one lane, one 32-centroid block of the first tile
        idx = lane_half0[:m]
    elif where == "spread":  # one lane, both blocks of the first 64-centroid tile
        idx = sorted([p for r in range(8) for p in (lane_half0[r], 32 + lane_half0[r])][:m])
    else:  # "late": one lane and block of tile 37
        idx = [64 * 37 + p for p in lane_half0[:m]]
    for j, i in enumerate(idx):
        c[i] = base
        c[i, 0] = np.float32(1.0) + np.float32((m - 1 - j) * 2e-4)
    x = np.concatenate([np.repeat(base[None], 8, 0), rng.normal(size=(56, d)).astype(np.float32)])
    return x, c, idx


@pytest.mark.parametrize("prec,m", [("bf16", 4), ("bf16", 8), ("fp8", 8), ("fp8", 4)])
@pytest.mark.parametrize("k,d,where,split", [(64, 64, "block", 0), (64, 128, "spread", 0), (300, 256, "spread", 0),
                                             (4096, 128, "late", 0), (4096, 128, "late", 3), (200, 512, "block", 0),
                                             (100, 300, "block", 0)])
def test_candidates_keep_tied_centroids(hip, prec, m, k, d, where, split):
    """m centroids whose BF16 / FP8 scores tie exactly, all inside one lane's
    share of one tile: the list must hold every one of them (the m smallest
    keys), and the FP32 rescore then returns the truly nearest, the last.
    (BF16 d = 300 and d = 512 rows are wider than 512 bytes and take the LDS
    kernel; FP8 d = 300 and 512 the register kernel's 32-centroid tiles.)"""
    print(f"{prec} top-{m} k={k} d={d} {where} split={split}", flush=True)
    x, c, idx = _near_duplicates(k, d, m, where, seed=k + d + m)
    a = E.HipAssigner(prec, rescore_m=m)
    a.set_split(split)
    a.set_centroids(c)
    labels, dist = a.assign(x)
    assert np.all(labels[:8] == idx[-1]), f"labels {labels[:8]}, expected {idx[-1]} (tied set {idx})"
    assert np.all(dist[:8] == 0.0)
    _, bd = O.brute_f64(x, c)
    got = O.dist_f64(x, c, labels)
    assert np.all(got <= bd + 1e-6 * (1 + bd))


@pytest.mark.parametrize("mag", [1e-39, 1e-30, 1.0, 1e15])
def test_fp8_rows_of_extreme_magnitude(hip, mag):
    """FP8 rows scaled to the f32 denormal range (largest |value| near
    1e-39, where 448 / amax overflows), to 1e-30 and to 1e15 each get a
    finite power-of-two scale: labels match FP64 brute force and distances
    are finite and exact to the FP32 rescore. (One magnitude per call: the
    positivity offset of the keys is taken per block of 256 rows.)"""
    rng = np.random.default_rng(21)
    d, k = 64, 200
    c = rng.normal(size=(k, d)).astype(np.float32)
    x = (rng.normal(size=(300, d)) * mag).astype(np.float32)
    a = E.HipAssigner("fp8", rescore_m=8)
    a.set_centroids(c)
    labels, dist = a.assign(x)
    assert np.all(np.isfinite(dist))
    bl, bd = O.brute_f64(x, c)
    got = O.dist_f64(x, c, labels)
    assert np.all(got <= bd + 1e-6 * (1 + bd)), np.nonzero(got > bd + 1e-6 * (1 + bd))[0][:10]
    assert np.all(np.abs(dist - got) <= 1e-4 * (1 + got))


@pytest.mark.parametrize("d,split", [(64, 0), (64, 3), (256, 0), (600, 0)])
def test_fp8_small_centroids_beside_large(hip, d, split):
    """FP8 centroids of very different magnitudes: 128 of scale 1e3 and 128
    of scale 1e-3, rows drawn near the small ones. Each centroid has its own
    power-of-two e4m3 scale, so the small ones keep their 3 mantissa bits
    (one scale from the largest |c| would flush every small centroid to zero
    and leave their scores tied at |c|^2 / 2). Labels match FP64 brute force.
    (d = 600 rows are wider than 512 bytes and take the LDS kernel.)"""
    print(f"fp8 mixed-magnitude centroids d={d} split={split}", flush=True)
    rng = np.random.default_rng(d + 5)
    small = (rng.normal(size=(128, d)) * 1e-3).astype(np.float32)
    big = (rng.normal(size=(128, d)) * 1e3).astype(np.float32)
    c = np.empty((256, d), np.float32)
    c[0::2], c[1::2] = small, big
    pick = rng.integers(0, 128, size=2000)
    x = (small[pick] + rng.normal(size=(2000, d)) * 1e-4).astype(np.float32)
    a = E.HipAssigner("fp8", rescore_m=8)
    a.set_split(split)
    a.set_centroids(c)
    labels, dist = a.assign(x)
    bl, bd = O.brute_f64(x, c)
    got = O.dist_f64(x, c, labels)
    wrong = np.nonzero(got > bd + 1e-6 * (1 + bd))[0]
    assert len(wrong) == 0, f"{len(wrong)} true label errors, e.g. rows {wrong[:10]}"
    assert np.array_equal(bl, 2 * pick)
    assert np.all(np.abs(dist - got) <= 1e-4 * (1 + got))


def test_bf16_odd_shapes(hip):
    for n, k, d in SHAPES:
        x = O.lcg_f32(n * d, 1).reshape(n, d)
        c = O.lcg_f32(k * d, 2).reshape(k, d)
        a = E.HipAssigner("bf16", rescore_m=8)
        a.set_centroids(c)
        labels, dist = a.assign(x)
        assert labels.max() < k
        got = O.dist_f64(x, c, labels)
        assert np.all(np.abs(dist - got) <= 1e-4 * (1 + got))


def test_bf16_bias(hip):
    rng = np.random.default_rng(8)
    x = rng.normal(size=(800, 64)).astype(np.float32)
    c = rng.normal(size=(200, 64)).astype(np.float32)
    bias = rng.uniform(0, 5, size=200).astype(np.float32)
    a = E.HipAssigner("bf16", rescore_m=8)
    a.set_centroids(c)
    a.set_bias(bias)
    labels, dist = a.assign(x)
    bl, bd = O.brute_f64(x, c, bias)
    got = O.dist_f64(x, c, labels) + bias[labels]
    assert (got > bd + 1e-6 * (1 + bd)).mean() < 0.01
    assert np.all(np.abs(dist - got) <= 1e-4 * (1 + got))


def test_name_and_state_errors(hip):
    a = E.HipAssigner(E.FP32)
    with pytest.raises(E.EmeansError) as ei:
        a.assign(np.zeros((2, 3), np.float32))
    assert ei.value.code == E.E_STATE
    a.set_centroids(np.zeros((4, 3), np.float32))
    assert "FP32" in a.name()
    with pytest.raises(E.EmeansError):
        a.set_bias(np.zeros(5, np.float32))
    with pytest.raises(E.EmeansError):
        E.HipAssigner(E.I8).set_centroids(np.zeros((1, 40000), np.float32))


@pytest.mark.parametrize("prec,rows", [("fp32", "f32"), ("i8", "i8"), ("i16", "i8"), ("bf16", "f32"), ("fp8", "f32"),
                                       ("fp32", "u8"), ("bf16", "i8")])
def test_tiles_do_not_change_results(hip, prec, rows):
    """Every tile of the search (128 or 256 rows per workgroup, the centroid
    tiles split 1..8 ways, or autotuned; uneven tile counts) returns the same
    labels and distances bit for bit."""
    rng = np.random.default_rng(13)
    n, k, d = 3000, 1300, 64
    if rows in ("i8", "u8"):
        lo, hi = (-128, 128) if rows == "i8" else (0, 256)
        x = rng.integers(lo, hi, size=(n, d)).astype(np.int8 if rows == "i8" else np.uint8)
        c = rng.normal(scale=40, size=(k, d)).astype(np.float32)
    else:
        x = _mixture(n, d, 600, 14)
        c = x[O.permutation(99, n)[:k]]
    bias = rng.uniform(0, 2, size=k).astype(np.float32)
    a = E.HipAssigner(prec, rescore_m=4)
    ref = None
    for fr in (2, 1, 0):
        for ns in (1, 2, 3, 5, 8, 0):
            if fr == 0 and ns != 0:
                continue
            print(f"{prec} rows {rows} tile fr={fr} split={ns}", flush=True)
            a.set_tile(fr)
            a.set_split(ns)
            a.set_centroids(c)
            a.set_bias(bias)
            got = a.assign(x)
            if ref is None:
                ref = got
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), f"tile {fr}/{ns}"
    if prec in ("i8", "i16"):
        cl, cd = O.i8_assign(x, c, bits=O.BITS_I16 if prec == "i16" else O.BITS_I8, bias=bias)
        assert np.array_equal(ref[0], cl) and np.array_equal(ref[1], cd)


def test_tile_choices_persist(hip, tmp_path):
    """The tuner's choice for a launch shape is written to EMH_TUNE_CACHE by
    one process and used by the next without timing again."""
    import json
    import subprocess
    import sys

    cache = tmp_path / "tiles.json"
    prog = (
        "import sys, numpy as np\n"
        "sys.path[:0] = [%r, %r]\n"
        "import torch; torch.cuda.init()\n"
        "import emeans_hip as E\n"
        "rng = np.random.default_rng(1)\n"
        "x = rng.normal(size=(5000, 96)).astype(np.float32)\n"
        "a = E.HipAssigner('bf16', rescore_m=4)\n"
        "a.set_centroids(x[:2000])\n"
        "lab, dist = a.assign(x)\n"
        "print(*a.last_tile(), int(lab.sum()))\n"
    ) % (E.__file__.rsplit("/", 1)[0], E.__file__.rsplit("/", 2)[0])
    env = dict(__import__("os").environ, EMH_TUNE_CACHE=str(cache))
    first = subprocess.run([sys.executable, "-c", prog], env=env, capture_output=True, text=True, timeout=300)
    assert first.returncode == 0, first.stderr[-2000:]
    entries = json.loads(cache.read_text())
    assert len(entries) == 1
    (key, tile), = entries.items()
    assert tile["fr"] in (1, 2) and tile["nslice"] >= 1 and "|" in key
    second = subprocess.run([sys.executable, "-c", prog], env=env, capture_output=True, text=True, timeout=300)
    assert second.returncode == 0, second.stderr[-2000:]
    fr1, ns1, s1 = map(int, first.stdout.split())
    fr2, ns2, s2 = map(int, second.stdout.split())
    assert (fr2, ns2) == (tile["fr"], tile["nslice"]) == (fr1, ns1) and s1 == s2
