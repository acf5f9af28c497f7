"""Pins the CPU oracle against the reference's own tests and known answers.

Each test names the reference test or closed form it restates. These run on
the CPU only; the GPU parity tests then use the oracle as the checker.
"""
import numpy as np
import pytest

import oracle as O

M64 = (1 << 64) - 1


# ---------------------------------------------------------------- RNG ----
def test_splitmix64_published_outputs():
    # splitmix64 seeded with 0 (Vigna's splitmix64.c): published first outputs.
    got = [int(v) for v in O.splitmix64(0, 4)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


def _xoshiro256pp_bigint(seed, count):
    """xoshiro256++ 1.0 as published (prng.di.unimi.it/xoshiro256plusplus.c),
    in Python integers, seeded with splitmix64 as src/lib.rs:231-241 does."""
    z = seed
    s = []
    for _ in range(4):
        z = (z + 0x9E3779B97F4A7C15) & M64
        x = z
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
        s.append(x ^ (x >> 31))

    def rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M64

    out = []
    for _ in range(count):
        result = (rotl((s[0] + s[3]) & M64, 23) + s[0]) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        out.append(result)
    return out


@pytest.mark.parametrize("seed", [0, 1, 1234, 7, (1234 * 1000003 + 1) & M64, 1234 ^ 0x5EED5EED5EED5EED])
def test_xoshiro_matches_published_algorithm(seed):
    r = O.OracleRng(seed)
    assert [r.next_u64() for _ in range(64)] == _xoshiro256pp_bigint(seed, 64)


def test_below_is_lemire_and_unbiased():
    # Lemire's method: result = high word of next*bound, rejecting low < (2^64 - bound) % bound
    for bound in (1, 2, 3, 7, 1000, 2**33 + 5, 2**63 + 11):
        a, b = O.OracleRng(99), O.OracleRng(99)
        for _ in range(50):
            got = a.below(bound)
            while True:
                m = b.next_u64() * bound
                if (m & M64) >= ((-bound) & M64) % bound:
                    break
            assert got == m >> 64 and 0 <= got < bound
    r = O.OracleRng(5)
    counts = np.bincount([r.below(6) for _ in range(60000)], minlength=6)
    assert counts.min() > 9500 and counts.max() < 10500


def test_open01_range_and_mean():
    v = O.OracleRng(3).open01_many(200000)
    assert v.min() > 0.0 and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.005
    r = O.OracleRng(3)
    x = r.next_u64()
    assert O.OracleRng(3).open01() == ((x >> 11) + 0.5) / 2.0**53


def test_permutation_is_fisher_yates():
    for n in (0, 1, 2, 10, 1000):
        p = O.permutation(1234, n)
        assert sorted(p.tolist()) == list(range(n))
    # explicit Fisher-Yates from the top with below(i + 1), src/lib.rs:267-273
    r = O.OracleRng(1234)
    p = list(range(10))
    for i in range(9, 0, -1):
        j = r.below(i + 1)
        p[i], p[j] = p[j], p[i]
    assert O.permutation(1234, 10).tolist() == p
    # position 0 is uniform over the values
    firsts = np.bincount([O.permutation(s, 5)[0] for s in range(5000)], minlength=5)
    assert firsts.min() > 850


# ------------------------------------------------------------- recipe ----
def test_recipe_matches_reference():
    # src/lib.rs:526-540: values printed by faiss 1.16.0
    r = O.resolve_recipe(200_000, 96, 1024)
    assert r.batch_size == 16384
    assert abs(r.alpha0 - 0.728183) < 1e-5
    r = O.resolve_recipe(1_000_000, 8, 100, O.Config(nepochs=2.5))
    assert r.epoch_rows == [1_000_000, 500_000, 1_000_000]
    assert r.batch_size == 4096
    assert abs(r.alpha0 - 0.319215) < 1e-5
    r = O.resolve_recipe(4000, 4, 8)
    assert r.batch_size == 32
    assert abs(r.alpha0 - 0.144604) < 1e-5
    assert O.resolve_recipe(1000, 2, 10).unused_threshold == 0.0


def test_recipe_closed_forms():
    # batch override clamps to [1, n]; alpha0 override wins; fractional epochs
    r = O.resolve_recipe(1000, 32, 10, O.Config(batch_size=5000, alpha0=0.3, unused_threshold=0.25))
    assert r.batch_size == 1000 and r.alpha0 == 0.3 and abs(r.unused_threshold - 0.25) < 1e-7
    r = O.resolve_recipe(1000, 32, 10, O.Config(nepochs=0.5, batch_size=300))
    assert r.epoch_rows == [500] and r.total_steps == 2
    r = O.resolve_recipe(200_000, 96, 1024)
    assert r.epoch_rows == [200_000] * 5 and r.total_steps == 5 * 13
    # the step law is c * k^a * b^b * n^c * t^d clamped to [1e-4, 1]
    c, ak, ab, an, at = 0.8244, -0.4617, 0.9003, -0.4498, -0.1057
    want = c * 1024**ak * 16384**ab * 200_000**an * 5.0**at
    assert abs(r.alpha0 - want) < 1e-12


def test_schedule_endpoints():
    # src/lib.rs:542-546
    assert abs(O.scheduled_alpha(0.5, 0, 100) - 0.5) < 1e-12
    assert abs(O.scheduled_alpha(0.5, 100, 100) - 0.05) < 1e-12
    assert abs(O.scheduled_alpha(0.5, 50, 100) - 0.5 * (0.1 + 0.9 * 0.5)) < 1e-12
    assert abs(O.scheduled_alpha(0.5, 500, 100) - 0.05) < 1e-12


# -------------------------------------------------------------- state ----
def test_full_batch_step_one_is_lloyd():
    # src/lib.rs:548-562
    x = np.array([0.0, 0.0, 0.0, 2.0, 10.0, 0.0, 10.0, 2.0], np.float32).reshape(4, 2)
    cent = np.array([[0.0, 0.0], [10.0, 0.0]], np.float32)
    usage = np.full(2, 1.0 / 65536.0, np.float32)
    sums = np.zeros((2, 2), np.float32)
    for i, row in enumerate(x):
        sums[i // 2] += row
    O.state_apply(cent, usage, [0, 1], [2, 2], sums, 1.0)
    assert cent.tolist() == [[0.0, 1.0], [10.0, 1.0]]
    assert usage.tolist() == [2.0, 2.0]


def test_apply_is_the_mass_position_ema():
    # N' = (1-a) N + a n ; c' = c + (a n / N') (s/n - c) in f32 (src/lib.rs:285-309)
    rng = np.random.default_rng(0)
    cent = rng.normal(size=(3, 4)).astype(np.float32)
    usage = np.array([0.5, 2.0, 1.0], np.float32)
    sums = rng.normal(size=(3, 4)).astype(np.float32)
    counts = np.array([3, 0, 5], np.uint32)
    c0, u0 = cent.copy(), usage.copy()
    a = np.float32(0.25)
    O.state_apply(cent, usage, [2, 0], counts, sums, float(a))
    for j in (0, 2):
        n = np.float32(counts[j])
        mass = np.float32(u0[j] * (np.float32(1) - a)) + np.float32(a * n)
        w = np.float32(np.float32(a * n) / mass)
        want = c0[j] + w * (sums[j] / n - c0[j])
        assert np.array_equal(cent[j], want.astype(np.float32))
        assert usage[j] == mass
    assert np.array_equal(cent[1], c0[1]) and usage[1] == np.float32(u0[1] * np.float32(0.75))


def test_revival_splits_a_donor():
    # src/lib.rs:564-573
    cent = np.array([[1.0, 1.0], [5.0, 5.0], [9.0, 9.0]], np.float32)
    usage = np.array([4.0, 0.0, 4.0], np.float32)
    n = O.state_revive(cent, usage, 0.1, O.OracleRng(1))
    assert n == 1
    assert usage.sum() == 8.0 and usage[1] == 2.0
    src = 0 if usage[0] == 2.0 else 2
    assert abs(cent[1, 0] - cent[src, 0]) < 1e-2
    assert cent[1, 0] != cent[src, 0]
    # split perturbation: even coordinates donor*(1-eps), child*(1+eps); odd reversed
    eps = np.float32(1e-4)
    base = np.float32(1.0 if src == 0 else 9.0)
    assert cent[src, 0] == base * (np.float32(1) - eps) and cent[1, 0] == base * (np.float32(1) + eps)
    assert cent[src, 1] == base * (np.float32(1) + eps) and cent[1, 1] == base * (np.float32(1) - eps)


def test_revival_zero_donor_nudges_child():
    cent = np.zeros((2, 3), np.float32)
    usage = np.array([4.0, 0.0], np.float32)
    assert O.state_revive(cent, usage, 0.1, O.OracleRng(2)) == 1
    assert cent[1, 0] == np.nextafter(np.float32(0), np.float32(1)) and cent[0, 0] == 0.0


# ----------------------------------------------------------- searches ----
SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 16), (37, 70, 33), (130, 257, 100), (64, 1030, 384), (9, 6, 1027)]


def _levels():
    lv = [O.SIMD_PORTABLE]
    if O.lib().or_simd_best() == O.SIMD_AVX512:
        lv.append(O.SIMD_AVX512)
    return lv


@pytest.mark.parametrize("n,k,d", SHAPES)
def test_kernels_match_brute_force(n, k, d):
    # src/cpu.rs:282-309
    x = O.lcg_f32(n * d, 1).reshape(n, d)
    c = O.lcg_f32(k * d, 2).reshape(k, d)
    bl, bd = O.brute_f64(x, c)
    for simd in _levels():
        labels, dist = O.cpu_assign(x, c, simd=simd, nthreads=2)
        got = O.dist_f64(x, c, labels)
        tol = 1e-4 * (1.0 + bd)
        assert np.all(got <= bd + tol)
        assert np.all(np.abs(dist.astype(np.float64) - bd) <= tol)


def test_brute_force_is_exact_on_small_case():
    # the FP64 reference itself: a numpy FP64 search on a tiny case, ties to the lowest index
    rng = np.random.default_rng(4)
    x = rng.integers(-3, 4, size=(40, 5)).astype(np.float32)
    c = rng.integers(-3, 4, size=(17, 5)).astype(np.float32)
    c[5] = c[2]  # duplicate centroid: ties must pick 2
    d2 = ((x[:, None, :].astype(np.float64) - c[None].astype(np.float64)) ** 2).sum(-1)
    bl, bd = O.brute_f64(x, c)
    assert np.array_equal(bl, d2.argmin(1)) and np.array_equal(bd, d2.min(1))


def test_avx512_and_portable_agree_closely():
    if len(_levels()) < 2:
        pytest.skip("no AVX-512 on this host")
    x = O.lcg_f32(300 * 96, 3).reshape(300, 96)
    c = O.lcg_f32(500 * 96, 4).reshape(500, 96)
    l0, d0 = O.cpu_assign(x, c, simd=O.SIMD_PORTABLE)
    l1, d1 = O.cpu_assign(x, c, simd=O.SIMD_AVX512)
    assert (l0 == l1).mean() > 0.99
    assert np.allclose(d0, d1, rtol=1e-4, atol=1e-4)


def test_int_rows_widen_to_identical_results():
    # non-f32 rows are widened per tile: identical to passing f32 (src/cpu.rs:106-127)
    rng = np.random.default_rng(1)
    x8 = rng.integers(-128, 128, size=(77, 40)).astype(np.int8)
    xu = rng.integers(0, 256, size=(77, 40)).astype(np.uint8)
    c = rng.normal(scale=40, size=(33, 40)).astype(np.float32)
    for xi in (x8, xu):
        a = O.cpu_assign(xi, c)
        b = O.cpu_assign(xi.astype(np.float32), c)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_bias_shifts_scores():
    # set_bias: rows minimise |x-c|^2 + b_j and the distance includes the bias
    x = np.array([[0.0, 0.0]], np.float32)
    c = np.array([[1.0, 0.0], [2.0, 0.0]], np.float32)
    l, d = O.cpu_assign(x, c)
    assert l[0] == 0 and d[0] == 1.0
    l, d = O.cpu_assign(x, c, bias=np.array([5.0, 0.0], np.float32))
    assert l[0] == 1 and d[0] == 4.0


I8_SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 64), (37, 70, 33), (130, 257, 100), (20, 1030, 512), (9, 6, 1027)]


@pytest.mark.parametrize("n,k,d", I8_SHAPES)
def test_integer_kernels_are_exact(n, k, d):
    # src/cpu_i8.rs:294-324
    x = O.lcg_i8(n * d, 1).reshape(n, d)
    ci = O.lcg_i8(k * d, 2).reshape(k, d)
    el, ed = O.brute_i64(x, ci)
    for bits in (O.BITS_I8, O.BITS_I16):
        for vnni in [False] + ([True] if O.has_vnni() else []):
            labels, dist = O.i8_assign(x, ci.astype(np.float32), bits=bits, vnni=vnni, nthreads=2)
            assert np.array_equal(labels, el)
            assert np.array_equal(dist, ed.astype(np.float32))


def test_i16_resolves_fractional_centroids():
    # src/cpu_i8.rs:326-336
    x = np.array([[1, 0]], np.int8)
    c = np.array([[0.26, 0.0], [-0.26, 0.0]], np.float32)
    l, d = O.i8_assign(x, c, bits=O.BITS_I16)
    assert l[0] == 0 and abs(d[0] - 0.74 * 0.74) < 0.01


def test_i8_offsets_are_the_vnni_identity():
    # offset = sum c^2 + 256 sum c, and offset - 2 * sum (x+128) c == |x-c|^2 - |x|^2
    rng = np.random.default_rng(2)
    c = rng.normal(scale=50, size=(9, 21)).astype(np.float32)
    c8, _, off, mult, shift = O.i8_prepare(c)
    cr = np.clip(np.where(c >= 0, np.floor(c + 0.5), np.ceil(c - 0.5)), -128, 127).astype(np.int64)
    assert np.array_equal(c8.astype(np.int64), cr) and mult == 2 and shift == 0
    assert np.array_equal(off, (cr * cr + 256 * cr).sum(1))
    x = rng.integers(-128, 128, size=(4, 21)).astype(np.int64)
    dp = (x + 128) @ cr.T
    assert np.array_equal(off[None] - 2 * dp, ((x[:, None] - cr[None]) ** 2).sum(-1) - (x * x).sum(1)[:, None])
    # bias is rounded half away from zero into the offset (src/cpu_i8.rs:112-117)
    b = np.array([0.5, -0.5, 1.49, 2.5, -2.5, 0, 0, 0, 0], np.float32)
    _, _, off2, _, _ = O.i8_prepare(c, bias=b)
    assert (off2 - off).tolist() == [1, -1, 1, 3, -3, 0, 0, 0, 0]


def test_i16_dot_wraps_like_vpdpwssd():
    """Centroids past the i8 range clamp to 32767 and the i32 dot product
    wraps (vpdpwssd lanes and a release-build i32 sum both wrap): labels and
    distances equal an exact i64 brute force whose dot is wrapped to i32."""
    rng = np.random.default_rng(11)
    n, k, d = 40, 9, 1024
    x = rng.integers(100, 128, size=(n, d)).astype(np.int8)
    c = np.full((k, d), 600.0, np.float32)
    c[1::2] *= -1
    c[:, :5] = rng.normal(size=(k, 5))
    _, c16, off, mult, shift = O.i8_prepare(c, bits=O.BITS_I16)
    dp = x.astype(np.int64) @ c16.astype(np.int64).T
    dpw = (dp + 2**31) % 2**32 - 2**31
    assert (dp != dpw).any()
    score = off[None, :] - mult * dpw
    want_l = score.argmin(1)
    norm = (x.astype(np.int64) ** 2).sum(1)
    want_d = np.maximum(norm + score[np.arange(n), want_l] / 4.0**shift, 0.0).astype(np.float32)
    for vnni in sorted({False, O.has_vnni()}):
        lab, dist = O.i8_assign(x, c, bits=O.BITS_I16, vnni=vnni)
        assert np.array_equal(lab, want_l) and np.array_equal(dist, want_d), f"vnni={vnni}"


def test_i16_shift_rule():
    # largest s <= 7 with d * 2^(14+s) < 2^31 (src/cpu_i8.rs:95-107)
    for d, s in ((1, 7), (1023, 7), (1024, 6), (16384, 2), (8192, 3), (32768, 1)):
        _, _, _, mult, shift = O.i8_prepare(np.zeros((1, d), np.float32), bits=O.BITS_I16)
        assert shift == s and mult == 2 << s


# ---------------------------------------------------------- quantizer ----
def test_faiss_calibration_and_round_trip():
    # src/quantize.rs:102-116
    sample = np.array([15.0, -3.0, 5.0, -3.0, 10.0, -2.0, 10.0, -4.0], np.float32).reshape(4, 2)
    q = O.Quantizer.fit(sample)
    assert q.shift.tolist() == [10.0, -3.0]
    assert q.scale == 16.0
    codes = q.encode(sample)
    assert codes.ravel().tolist() == [80, 0, -80, 0, 0, 16, 0, -16]
    back = q.decode(codes.astype(np.float32))
    assert np.array_equal(back, sample)
    assert q.decode_sqdist(256.0) == 1.0
    assert q.encode(np.array([1000.0, -3.0], np.float32)).ravel().tolist() == [127, 0]
    assert q.clipped_fraction(np.array([1000.0, -3.0], np.float32)) == 0.5


def test_quantile_calibration():
    rng = np.random.default_rng(3)
    x = rng.standard_t(3, size=(2000, 8)).astype(np.float32)
    q = O.Quantizer.fit(x, target_peak=127.0, quantile=0.99, power_of_two=False)
    mags = np.abs(x - q.shift).ravel()
    idx = int(round((mags.size - 1) * 0.99))
    assert q.scale == np.float32(127.0) / np.sort(mags)[idx]
    assert np.allclose(q.shift, x.astype(np.float64).mean(0).astype(np.float32))


# ----------------------------------------------------------- training ----
def _blobs_6d():
    """The data of src/lib.rs:577-587, from the crate's own RNG."""
    n, d = 20_000, 6
    r = O.OracleRng(7)
    u = r.open01_many(20 * d)
    centers = (np.float32(8.0) * (u.astype(np.float32) - np.float32(0.5))).astype(np.float32)
    noise_u = r.open01_many(n * d * 4).astype(np.float32).reshape(n, d, 4) - np.float32(0.5)
    noise = np.zeros((n, d), np.float32)
    for t in range(4):
        noise = (noise + noise_u[:, :, t]).astype(np.float32)
    x = (centers.reshape(20, d)[np.arange(n) % 20] + noise).astype(np.float32)
    return x


def test_closing_mean_and_loss():
    # src/lib.rs:575-614
    x = _blobs_6d()
    n, d, k = x.shape[0], x.shape[1], 32
    out = O.train(x, k)
    counts = np.bincount(out.labels, minlength=k)
    assert counts.sum() == n
    sums = np.zeros((k, d))
    np.add.at(sums, out.labels, x.astype(np.float64))
    live = counts > 0
    mean = sums[live] / counts[live, None]
    assert np.all(np.abs(out.centroids[live].astype(np.float64) - mean) < 1e-3)

    def mse(c):
        return O.assign(x, c)[1].astype(np.float64).mean()

    init = x[O.permutation(1234, n)[:k]]
    trained, start = mse(out.centroids), mse(init)
    assert trained < 0.75 * start and trained < 2.0
    assert all(b < a for a, b in zip(out.objective, out.objective[1:]))


def test_train_i8_rows_integer_sums():
    # int8 codes through CpuI8Assigner: labels exact, closing mean = row means
    rng = np.random.default_rng(11)
    centers = rng.integers(-100, 100, size=(12, 16))
    x = (centers[rng.integers(0, 12, 3000)] + rng.integers(-6, 7, size=(3000, 16))).clip(-128, 127).astype(np.int8)
    out = O.train(x, 10, O.Config(nepochs=2), assigner=O.ASSIGNER_I8)
    counts = np.bincount(out.labels, minlength=10)
    sums = np.zeros((10, 16))
    np.add.at(sums, out.labels, x.astype(np.float64))
    live = counts > 0
    assert np.all(np.abs(out.centroids[live] - sums[live] / counts[live, None]) < 1e-3)


def test_balance_stats():
    imb, mx, empty = O.balance_stats(np.array([0, 0, 1, 1], np.uint32), 3)
    assert abs(imb - 3 * 8 / 16) < 1e-12 and abs(mx - 2 * 3 / 4) < 1e-12 and empty == 1


# ---- report's nearest-centroid step (src/report.rs:104-114, 169-183) -------
def test_report_nearest_reference_case():
    # src/report.rs:283-312: 1-D centroids 0, 1, 50 with sizes 101, 100, 3;
    # the test's rare_prototypes == [(2, 3, 49.0)] fixes cluster 2's nearest
    # distance to 49 (isolation 49 / median(1, 1, 49)).
    c = np.array([[0.0], [1.0], [50.0]], np.float32)
    nearest, dist = O.report_nearest(c, np.array([101, 100, 3], np.uint32))
    assert nearest.tolist() == [1, 0, 1]
    assert dist.tolist() == [1.0, 1.0, 49.0]


def test_report_sqdist_order_known_answers():
    # Values derived by hand in f32 (spacing 2 above 2^24, ties to even):
    # d = 32, v0 = 4096, v2 = v16 = v18 = 1: lane 0 = 2^24 + 1 -> 2^24, lane
    # 2 = 1 + 1 = 2, so the fold gives 2^24 + 2. A single left fold gives
    # 2^24 (each +1 rounds away), lanes over contiguous blocks give 2^24, the
    # exact sum 2^24 + 3 rounds to 2^24 + 4.
    a = np.zeros(32, np.float32)
    a[[0, 2, 16, 18]] = [4096, 1, 1, 1]
    assert O.report_sqdist(a, np.zeros(32, np.float32)) == np.float32(16777218.0)
    # d = 34, v0 = 4096, v1 = 1, remainder v32 = v33 = 1: lanes fold to 2^24,
    # the remainder folds to 2 on its own, and 2^24 + 2 is exact. Adding the
    # remainder one element at a time would give 2^24, exactly 2^24 + 4.
    b = np.zeros(34, np.float32)
    b[[0, 1, 32, 33]] = [4096, 1, 1, 1]
    assert O.report_sqdist(b, np.zeros(34, np.float32)) == np.float32(16777218.0)
    # symmetric bit for bit
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=(2, 77)).astype(np.float32)
    assert O.report_sqdist(x, y).tobytes() == O.report_sqdist(y, x).tobytes()


@pytest.mark.parametrize("k,d", [(50, 3), (200, 37), (300, 64)])
def test_report_nearest_integer_brute_force(k, d):
    # small integers: every f32 operation is exact, so the distances are the
    # exact integer ones, and ties (frequent here) must go to the lowest index
    rng = np.random.default_rng(k + d)
    c = rng.integers(-3, 4, size=(k, d)).astype(np.float32)
    c[5] = c[9]  # an exact duplicate: distance 0
    sizes = rng.integers(0, 3, size=k).astype(np.uint32)
    nearest, dist = O.report_nearest(c, sizes, nthreads=2)
    ci = c.astype(np.int64)
    sq = ((ci[:, None, :] - ci[None, :, :]) ** 2).sum(-1)
    sq[np.arange(k), np.arange(k)] = 1 << 62
    sq[:, sizes == 0] = 1 << 62
    best = sq.argmin(1)
    has = sq[np.arange(k), best] < (1 << 62)
    assert np.array_equal(nearest[has], best[has])
    assert np.array_equal(dist[has], np.sqrt(sq[np.arange(k), best][has].astype(np.float32)))
    assert np.array_equal(nearest[~has], np.arange(k)[~has]) and np.all(np.isinf(dist[~has]))


def test_report_nearest_edge_cases():
    one = np.ones((1, 4), np.float32)
    n, dd = O.report_nearest(one)
    assert n.tolist() == [0] and np.isinf(dd[0])
    c = np.arange(12, dtype=np.float32).reshape(4, 3)
    n, dd = O.report_nearest(c, np.zeros(4, np.uint32))  # no live cluster
    assert n.tolist() == [0, 1, 2, 3] and np.all(np.isinf(dd))
    c[2, 1] = np.nan  # its distances are NaN: never chosen, and it finds none
    n, dd = O.report_nearest(c)
    assert n.tolist() == [1, 0, 2, 1] and np.isinf(dd[2])
    r = O.report_nearest(c, rows=(1, 3))
    assert r[0].tolist() == [0, 2]


def _reference_report_case():
    # the inputs of src/report.rs:285-299, in f32 as the test builds them
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


def test_report_reference_case():
    # src/report.rs:283-312
    labels, sq, c = _reference_report_case()
    r = O.report(labels, sq, c)
    cl = r["clusters"]
    assert cl["size"].tolist() == [101, 100, 3]
    score = np.fmax(r["local_z"], r["global_z"])
    assert int(np.argmax(score)) == 200
    assert r["local_z"][200] > 10.0
    assert len(r["families"]) == 2 and r["families"][0]["size"] == 201
    assert cl["family"][0] == cl["family"][1]
    assert cl["isolation"].tolist() == [1.0, 1.0, 49.0]


def test_report_quantiles_and_spread_closed_forms():
    # quantile index = round half away from zero of (len - 1) q
    # (src/report.rs:83-88): len 6 -> index 3 for the median (2.5 -> 3)
    v = np.array([1, 2, 3, 4, 5, 6], np.float32)
    assert O._quantile(v, 0.5) == 4.0 and O._quantile(v, 0.9) == 6.0 and O._quantile(v, 0.0) == 1.0
    # MAD of 1..7 around the median 4: |dev| = 3,2,1,0,1,2,3 -> sorted median 2
    assert O._median_mad(np.arange(1, 8, dtype=np.float32)) == (4.0, 2.0)
    # one cluster of distances 3 and 4 (squared 9, 16): rms = sqrt(12.5)
    r = O.report(np.zeros(2, np.uint32), np.array([9.0, 16.0], np.float32), np.zeros((1, 2), np.float32))
    c = r["clusters"]
    assert c["rms"][0] == np.float32(np.sqrt(12.5)) and c["max"][0] == 4.0 and c["median"][0] == 4.0
    # no other live centroid: nearest is itself at +inf; the median nearest
    # distance is then +inf too, and isolation = inf / inf = NaN
    assert c["nearest_centroid"][0] == 0 and np.isinf(c["nearest_dist"][0]) and np.isnan(c["isolation"][0])
    # the only cluster is its own family holding every row
    assert len(r["families"]) == 1 and r["families"][0]["share"] == 1.0


def _balance_ladder(lam):
    """1-D, k = 3, one batch per epoch, 2 epochs, alpha0 = 1, no revival.
    Initial centroids (permutation positions 0..2, src/lib.rs:375-379) are
    0, 10 and 100; the rows are a core {0, 0, 0, 0, 1, -1} around 0, two
    rungs 4.0 and 4.9 between 0 and 10, one more 10 and one more 100."""
    n = 11
    perm = O.permutation(1234, n)
    vals = [0.0, 10.0, 100.0, 0.0, 0.0, 0.0, 1.0, -1.0, 4.0, 4.9, 10.0]
    x = np.empty((n, 1), np.float32)
    for pos, v in zip(perm, vals):
        x[pos, 0] = v
    cfg = O.Config(nepochs=2.0, batch_size=n, alpha0=1.0, balance=lam, seed=1234)
    return x, perm, O.train(x, 3, cfg)


def test_balance_penalty_hand_derived():
    """src/lib.rs:411-430 on a case worked by hand. Step 0 has no penalty
    (s = 0): every row below 5 joins centroid 0, so after the replace step
    (alpha = 1) c = (8.9/8, 10, 100) = (1.1125, 10, 100), N = (8, 2, 1) and
    s = (4 * 0 + 1 + 1 + 16 + 24.01) / 11 = 3.81909.
    Step 1 biases b_j = lambda s / mean(N) * N_j with mean(N) = 11/3: the
    rung 4.9 leaves cluster 0 once 6 * scale > 5.1^2 - (4.9 - 1.1125)^2
    = 11.66 (scale > 1.944); the rung 4.0 once 6 * scale > 6^2 - 2.8875^2
    = 27.66 (scale > 4.610). lambda = 2.88 gives scale = 3.000: exactly the
    4.9 rung moves. The epoch-1 objective is the unpenalized distance sum
    (the bias is taken back off, src/lib.rs:425-430) and the closing mean
    (src/lib.rs:481-489) averages the last pass's members."""
    x, perm, r = _balance_ladder(2.88)
    lab = {float(x[i, 0]): int(r.labels[i]) for i in range(11)}
    assert lab[4.0] == 0 and abs(float(x[perm[8], 0]) - 4.0) < 1e-6
    assert int(r.labels[perm[9]]) == 1, "the 4.9 rung must move to centroid 10 under the penalty"
    assert int(r.labels[perm[8]]) == 0, "the 4.0 rung must stay (scale below 4.61)"
    c0 = np.float32(4.0) * (np.float32(1) / np.float32(7))
    c1 = np.float32(24.9) * (np.float32(1) / np.float32(3))
    assert abs(r.centroids[0, 0] - c0) <= 1e-6 and abs(r.centroids[1, 0] - c1) <= 2e-6
    assert r.centroids[2, 0] == np.float32(100.0)
    # epoch 0: distances to the initial centroids; epoch 1: to (1.1125, 10, 100)
    assert abs(r.objective[0] - 42.01) <= 1e-4
    c = 1.1125
    e1 = 4 * c * c + (1 - c) ** 2 + (1 + c) ** 2 + (4 - c) ** 2 + 5.1 ** 2
    assert abs(r.objective[1] - e1) <= 1e-4 * e1
    # without the penalty nothing moves: centroid 0 keeps both rungs
    _, _, r0 = _balance_ladder(0.0)
    assert int(r0.labels[perm[9]]) == 0 and int(r0.labels[perm[8]]) == 0
    assert abs(r0.centroids[0, 0] - np.float32(8.9) * (np.float32(1) / np.float32(8))) <= 1e-6
    # a scale past 4.61 moves both rungs (lambda = 5: scale = 5.21)
    _, _, r5 = _balance_ladder(5.0)
    assert int(r5.labels[perm[9]]) == 1 and int(r5.labels[perm[8]]) == 1


def _smooth(n, d, seed):
    rng = np.random.default_rng(seed)
    scales = np.linspace(1.0, 0.2, d).astype(np.float32)
    return (rng.normal(size=(n, d)) * scales).astype(np.float32)


def test_balance_closing_mean_invariant_and_imbalance():
    """lambda = 0.1 keeps the closing-mean invariant (src/lib.rs:575-590:
    each centroid is the mean of its last-pass members); on smooth data the
    imbalance factor k sum n_j^2 / n^2 of the penalized labels falls as
    lambda rises (README.md:111-122)."""
    x = _smooth(20000, 24, 3)
    k = 64
    imb = []
    for lam in (0.0, 0.1, 1.0):
        r = O.train(x, k, O.Config(nepochs=3.0, balance=lam))
        if lam == 0.1:
            for j in range(k):
                members = x[r.labels == j]
                if len(members):
                    assert np.allclose(r.centroids[j], members.astype(np.float64).mean(0), atol=1e-3)
        imb.append(O.balance_stats(r.labels, k)[0])
    assert imb[0] > imb[1] > imb[2], f"imbalance {imb}"
