"""ctypes wrapper of the CPU oracle (oracle/emeans_oracle.c).

TEST INFRASTRUCTURE ONLY: the plain-C restatement of the reference e-means
crate used as the checker by tests/, by ``__graft_entry__.smoke()`` and as
the CPU baseline of bench.py. The product (e-means_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ROWS = {np.dtype(np.float32): 0, np.dtype(np.int8): 1, np.dtype(np.uint8): 2}
SIMD_PORTABLE, SIMD_AVX512 = 0, 2
BITS_I8, BITS_I16 = 0, 1
ASSIGNER_F32, ASSIGNER_I8, ASSIGNER_I16 = 0, 1, 2
LAW_AUTO, LAW_THROUGHPUT, LAW_QUALITY = 0, 1, 2
MAX_D = 32768


class Rng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


class Config(C.Structure):
    """Mirror of ``Config`` (src/lib.rs:103-139); ``Config()`` is the default."""

    _fields_ = [
        ("nepochs", C.c_double),
        ("batch_law", C.c_int32),
        ("has_batch_size", C.c_int32),
        ("batch_size", C.c_uint64),
        ("has_alpha0", C.c_int32),
        ("has_unused_threshold", C.c_int32),
        ("alpha0", C.c_double),
        ("unused_threshold", C.c_float),
        ("balance", C.c_float),
        ("seed", C.c_uint64),
    ]

    def __init__(self, nepochs=5.0, batch_law=LAW_AUTO, batch_size=None, alpha0=None,
                 unused_threshold=None, balance=0.0, seed=1234):
        super().__init__()
        self.nepochs = nepochs
        self.batch_law = batch_law
        self.has_batch_size = batch_size is not None
        self.batch_size = batch_size or 0
        self.has_alpha0 = alpha0 is not None
        self.alpha0 = alpha0 or 0.0
        self.has_unused_threshold = unused_threshold is not None
        self.unused_threshold = unused_threshold or 0.0
        self.balance = balance
        self.seed = seed


class Recipe(C.Structure):
    _fields_ = [
        ("batch_size", C.c_uint64),
        ("alpha0", C.c_double),
        ("total_steps", C.c_uint64),
        ("unused_threshold", C.c_float),
        ("n_epochs", C.c_uint32),
        ("epoch_rows_", C.c_uint64 * 1024),
    ]

    @property
    def epoch_rows(self):
        return [int(self.epoch_rows_[i]) for i in range(self.n_epochs)]


class TrainOut(C.Structure):
    _fields_ = [
        ("n_splits", C.c_uint64),
        ("n_objective", C.c_uint32),
        ("objective_", C.c_double * 1024),
        ("recipe", Recipe),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, sz, u64, i32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_int
        L.or_rng_init.argtypes = [C.POINTER(Rng), u64]
        L.or_rng_next.argtypes = [C.POINTER(Rng)]
        L.or_rng_next.restype = u64
        L.or_rng_below.argtypes = [C.POINTER(Rng), u64]
        L.or_rng_below.restype = u64
        L.or_rng_open01.argtypes = [C.POINTER(Rng)]
        L.or_rng_open01.restype = C.c_double
        L.or_rng_permutation.argtypes = [C.POINTER(Rng), C.c_uint32, vp]
        L.or_rng_fill_open01.argtypes = [C.POINTER(Rng), sz, vp]
        L.or_splitmix64_stream.argtypes = [u64, sz, vp]
        L.or_resolve_recipe.argtypes = [u64, u64, u64, C.POINTER(Config), C.POINTER(Recipe)]
        L.or_scheduled_alpha.argtypes = [C.c_double, u64, u64]
        L.or_scheduled_alpha.restype = C.c_double
        L.or_state_apply.argtypes = [sz, sz, vp, vp, vp, sz, vp, vp, C.c_float]
        L.or_state_revive.argtypes = [sz, sz, vp, vp, C.c_float, C.POINTER(Rng)]
        L.or_state_revive.restype = sz
        L.or_simd_best.restype = i32
        L.or_sqnorm.argtypes = [vp, sz]
        L.or_sqnorm.restype = C.c_float
        L.or_cpu_norms.argtypes = [vp, sz, sz, vp, vp]
        L.or_cpu_assign.argtypes = [i32, i32, vp, i32, sz, sz, vp, vp, sz, vp, vp]
        L.or_i8_has_vnni.restype = i32
        L.or_i8_prepare.argtypes = [i32, vp, sz, sz, vp, vp, vp, vp, vp, vp]
        L.or_i8_search.argtypes = [i32, i32, i32, vp, sz, sz, vp, sz, vp, vp, vp]
        L.or_quant_fit.argtypes = [vp, sz, sz, C.c_float, C.c_double, i32, vp, vp]
        L.or_quant_encode.argtypes = [vp, sz, sz, vp, C.c_float, vp]
        L.or_quant_fit_mt.argtypes = [i32, vp, sz, sz, C.c_float, C.c_double, i32, vp, vp]
        L.or_quant_encode_mt.argtypes = [i32, vp, sz, sz, vp, C.c_float, vp]
        L.or_pool_pins_threads.restype = i32
        L.or_quant_decode.argtypes = [vp, sz, sz, vp, C.c_float, vp]
        L.or_quant_clipped_fraction.argtypes = [vp, sz, sz, vp, C.c_float]
        L.or_quant_clipped_fraction.restype = C.c_double
        L.or_report_sqdist.argtypes = [vp, vp, sz]
        L.or_report_sqdist.restype = C.c_float
        L.or_report_nearest.argtypes = [i32, vp, sz, sz, vp, sz, sz, vp, vp]
        L.or_brute_f64.argtypes = [i32, vp, i32, sz, sz, vp, sz, vp, vp, vp]
        L.or_dist_f64.argtypes = [i32, vp, i32, sz, sz, vp, vp, vp]
        L.or_brute_i64.argtypes = [vp, sz, sz, vp, sz, vp, vp]
        L.or_train.argtypes = [vp, i32, sz, sz, sz, C.POINTER(Config), i32, i32, i32, vp, vp, C.POINTER(TrainOut)]
        L.or_assign.argtypes = [vp, i32, sz, sz, vp, sz, i32, i32, vp, vp]
        L.or_balance_stats.argtypes = [vp, sz, sz, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def pins_threads() -> bool:
    """The oracle's thread pool pins worker t to the t-th allowed CPU."""
    return bool(lib().or_pool_pins_threads())


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


# ---- RNG (src/lib.rs:227-274) -------------------------------------------
class OracleRng:
    def __init__(self, seed: int):
        self.st = Rng()
        lib().or_rng_init(C.byref(self.st), C.c_uint64(seed & (2**64 - 1)))

    def next_u64(self) -> int:
        return int(lib().or_rng_next(C.byref(self.st)))

    def below(self, bound: int) -> int:
        return int(lib().or_rng_below(C.byref(self.st), bound))

    def open01(self) -> float:
        return float(lib().or_rng_open01(C.byref(self.st)))

    def open01_many(self, count: int) -> np.ndarray:
        out = np.empty(count, np.float64)
        lib().or_rng_fill_open01(C.byref(self.st), count, _p(out))
        return out

    def permutation(self, n: int) -> np.ndarray:
        p = np.empty(n, np.uint32)
        lib().or_rng_permutation(C.byref(self.st), n, _p(p))
        return p


def splitmix64(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.uint64)
    lib().or_splitmix64_stream(seed, count, _p(out))
    return out


def permutation(seed: int, n: int) -> np.ndarray:
    return OracleRng(seed).permutation(n)


# ---- recipe --------------------------------------------------------------
def resolve_recipe(n: int, d: int, k: int, cfg: Config | None = None) -> Recipe:
    r = Recipe()
    rc = lib().or_resolve_recipe(n, d, k, C.byref(cfg or Config()), C.byref(r))
    if rc:
        raise ValueError("recipe out of the oracle's range")
    return r


def scheduled_alpha(alpha0: float, step: int, total: int) -> float:
    return float(lib().or_scheduled_alpha(alpha0, step, total))


# ---- state ---------------------------------------------------------------
def state_apply(centroids, usage, touched, counts, sums, alpha):
    k, d = centroids.shape
    touched = np.ascontiguousarray(touched, np.uint32)
    lib().or_state_apply(k, d, _p(centroids), _p(usage), _p(touched), len(touched),
                         _p(np.ascontiguousarray(counts, np.uint32)),
                         _p(np.ascontiguousarray(sums, np.float32)), C.c_float(alpha))


def state_revive(centroids, usage, tau, rng: OracleRng) -> int:
    k, d = centroids.shape
    return int(lib().or_state_revive(k, d, _p(centroids), _p(usage), C.c_float(tau), C.byref(rng.st)))


# ---- searches ------------------------------------------------------------
def sqnorm(v) -> float:
    v = np.ascontiguousarray(v, np.float32)
    return float(lib().or_sqnorm(_p(v), v.size))


def cpu_norms(c, bias=None):
    c = np.ascontiguousarray(c, np.float32)
    k, d = c.shape
    out = np.empty(k, np.float32)
    b = None if bias is None else np.ascontiguousarray(bias, np.float32)
    lib().or_cpu_norms(_p(c), k, d, _p(b), _p(out))
    return out


def cpu_assign(x, c, bias=None, simd=None, nthreads=None):
    """CpuAssigner set_centroids [+ set_bias] + assign (src/cpu.rs:84-167)."""
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(c, np.float32)
    n, d = x.shape
    k = c.shape[0]
    norms = cpu_norms(c, bias)
    labels = np.empty(n, np.uint32)
    dist = np.empty(n, np.float32)
    simd = lib().or_simd_best() if simd is None else simd
    rc = lib().or_cpu_assign(simd, nthreads or default_threads(), _p(x), ROWS[x.dtype], n, d, _p(c),
                             _p(norms), k, _p(labels), _p(dist))
    if rc:
        raise ValueError("or_cpu_assign failed")
    return labels, dist


def has_vnni() -> bool:
    return bool(lib().or_i8_has_vnni())


def i8_prepare(c, bits=BITS_I8, bias=None):
    c = np.ascontiguousarray(c, np.float32)
    k, d = c.shape
    c8 = np.zeros((k, d), np.int8)
    c16 = np.zeros((k, d), np.int16)
    off = np.empty(k, np.int64)
    mult = C.c_int64()
    shift = C.c_uint32()
    b = None if bias is None else np.ascontiguousarray(bias, np.float32)
    if lib().or_i8_prepare(bits, _p(c), k, d, _p(b), _p(c8), _p(c16), _p(off), C.byref(mult), C.byref(shift)):
        raise ValueError(f"integer assignment supports d <= {MAX_D}")
    return c8, c16, off, int(mult.value), int(shift.value)


def i8_assign(x, c, bits=BITS_I8, bias=None, vnni=None, nthreads=None):
    """CpuI8Assigner set_centroids [+ set_bias] + assign (src/cpu_i8.rs:73-191)."""
    x = np.ascontiguousarray(x, np.int8)
    c = np.ascontiguousarray(c, np.float32)
    n, d = x.shape
    k = c.shape[0]
    labels = np.empty(n, np.uint32)
    dist = np.empty(n, np.float32)
    vnni = has_vnni() if vnni is None else vnni
    b = None if bias is None else np.ascontiguousarray(bias, np.float32)
    rc = lib().or_i8_search(bits, int(vnni), nthreads or default_threads(), _p(x), n, d, _p(c), k, _p(b),
                            _p(labels), _p(dist))
    if rc:
        raise ValueError(f"integer assignment supports d <= {MAX_D}")
    return labels, dist


def brute_f64(x, c, bias=None, nthreads=None):
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(c, np.float32)
    n, d = x.shape
    k = c.shape[0]
    labels = np.empty(n, np.uint32)
    dist = np.empty(n, np.float64)
    b = None if bias is None else np.ascontiguousarray(bias, np.float32)
    lib().or_brute_f64(nthreads or default_threads(), _p(x), ROWS[x.dtype], n, d, _p(c), k, _p(b),
                       _p(labels), _p(dist))
    return labels, dist


def dist_f64(x, c, labels, nthreads=None):
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(c, np.float32)
    labels = np.ascontiguousarray(labels, np.uint32)
    n, d = x.shape
    out = np.empty(n, np.float64)
    lib().or_dist_f64(nthreads or default_threads(), _p(x), ROWS[x.dtype], n, d, _p(c), _p(labels), _p(out))
    return out


def brute_i64(x, c8):
    x = np.ascontiguousarray(x, np.int8)
    c8 = np.ascontiguousarray(c8, np.int8)
    n, d = x.shape
    labels = np.empty(n, np.uint32)
    dist = np.empty(n, np.int64)
    lib().or_brute_i64(_p(x), n, d, _p(c8), c8.shape[0], _p(labels), _p(dist))
    return labels, dist


# ---- report's nearest-centroid step (src/report.rs) -------------------------
def report_sqdist(a, b) -> np.float32:
    """sqdist (src/report.rs:104-114) of two equal-length f32 vectors."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().or_report_sqdist(_p(a), _p(b), a.shape[0]))


def report_nearest(centroids, sizes=None, rows=None, nthreads=None):
    """Nearest live other centroid of each centroid (src/report.rs:169-183):
    (index u32, L2 distance f32); rows=(r0, r1) restricts to those rows."""
    c = np.ascontiguousarray(centroids, np.float32)
    k, d = c.shape
    r0, r1 = rows if rows is not None else (0, k)
    sz = None if sizes is None else np.ascontiguousarray(sizes, np.uint32)
    nearest = np.empty(r1 - r0, np.uint32)
    dist = np.empty(r1 - r0, np.float32)
    lib().or_report_nearest(nthreads or os.cpu_count() or 1, _p(c), k, d, _p(sz), r0, r1, _p(nearest), _p(dist))
    return nearest, dist


def _total_sorted(v):
    """Ascending f32::total_cmp order."""
    v = np.ascontiguousarray(v, np.float32)
    u = v.view(np.uint32)
    key = np.where(u & np.uint32(0x80000000), ~u, u | np.uint32(0x80000000))
    return v[np.argsort(key, kind="stable")]


def _quantile(sorted_v, q):
    """src/report.rs:83-88 (f64::round: half away from zero)."""
    if len(sorted_v) == 0:
        return np.float32(0.0)
    x = (len(sorted_v) - 1) * q
    i = math.floor(x)
    if x - i >= 0.5:
        i += 1
    return np.float32(sorted_v[i])


def _median_mad(sorted_v):
    """src/report.rs:91-96."""
    med = _quantile(sorted_v, 0.5)
    dev = _total_sorted(np.abs(sorted_v - med).astype(np.float32))
    return med, _quantile(dev, 0.5)


def report(labels, sqdist, centroids, family_radius=1.25, min_local_size=30):
    """report::report (src/report.rs:119-243) for small inputs: a dict of the
    per-cluster fields (arrays over k), families (leader, members, size,
    share; largest first), local_z, global_z and the three medians. Clusters
    of equal size are taken in index order (the reference's unstable sorts
    leave that order open: parity unpinned on size ties)."""
    labels = np.ascontiguousarray(labels, np.uint32)
    sq = np.ascontiguousarray(sqdist, np.float32)
    c = np.ascontiguousarray(centroids, np.float32)
    k, d = c.shape
    n = labels.shape[0]
    f32 = np.float32
    dist = np.sqrt(np.fmax(sq, f32(0.0))).astype(np.float32)
    g_med, g_mad = _median_mad(_total_sorted(dist))
    g_scale = max(f32(f32(1.4826) * g_mad), f32(np.finfo(np.float32).tiny))
    cols = ("size", "rms", "median", "mad", "p90", "p99", "max", "looseness", "nearest_centroid", "nearest_dist",
            "isolation", "family")
    cl = {f: np.zeros(k, np.uint32 if f in ("size", "nearest_centroid", "family") else np.float32) for f in cols}
    for j in range(k):
        m = _total_sorted(dist[labels == j])
        if len(m) == 0:
            continue
        cl["size"][j] = len(m)
        cl["median"][j], cl["mad"][j] = _median_mad(m)
        m64 = m.astype(np.float64)
        cl["rms"][j] = f32(math.sqrt(np.cumsum(m64 * m64)[-1] / len(m)))
        cl["p90"][j] = _quantile(m, 0.9)
        cl["p99"][j] = _quantile(m, 0.99)
        cl["max"][j] = m[-1]
    live = cl["size"] > 0
    nearest, nd = report_nearest(c, cl["size"])
    median_rms = _quantile(_total_sorted(cl["rms"][live]), 0.5)
    median_nearest = _quantile(_total_sorted(nd[live]), 0.5)
    tiny = f32(np.finfo(np.float32).tiny)
    cl["nearest_centroid"][:] = nearest
    cl["nearest_dist"][:] = nd
    with np.errstate(invalid="ignore", divide="ignore"):  # inf / inf is NaN here as in the reference
        cl["looseness"][:] = cl["rms"] / max(median_rms, tiny)
        cl["isolation"][:] = nd / max(median_nearest, tiny)
    radius = f32(f32(family_radius) * median_nearest)
    order = sorted(np.flatnonzero(live).tolist(), key=lambda j: -int(cl["size"][j]))
    family_of = [None] * k
    fams = []
    for lead in order:
        if family_of[lead] is not None:
            continue
        joins = [j for j in order if family_of[j] is None and
                 (j == lead or np.sqrt(report_sqdist(c[lead], c[j])) <= radius)]
        for j in joins:
            family_of[j] = len(fams)
        size = sum(int(cl["size"][j]) for j in joins)
        fams.append({"leader": lead, "members": joins, "size": size, "share": size / n})
    fams = sorted(fams, key=lambda f: -f["size"])
    for f, fam in enumerate(fams):
        for j in fam["members"]:
            cl["family"][j] = f
    global_z = ((dist - g_med) / g_scale).astype(np.float32)
    local_z = global_z.copy()
    for r in range(n):
        j = labels[r]
        if not (int(cl["size"][j]) < min_local_size or cl["mad"][j] <= 0.0):
            local_z[r] = (dist[r] - cl["median"][j]) / (f32(1.4826) * cl["mad"][j])
    return {"clusters": cl, "families": fams, "local_z": local_z, "global_z": global_z, "median_dist": g_med,
            "median_rms": median_rms, "median_nearest_dist": median_nearest}


# ---- quantizer (src/quantize.rs) -------------------------------------------
@dataclass
class Quantizer:
    d: int
    shift: np.ndarray
    scale: float

    @staticmethod
    def fit(sample, target_peak=112.0, quantile=1.0, power_of_two=True, nthreads=1) -> "Quantizer":
        sample = np.ascontiguousarray(sample, np.float32)
        n, d = sample.shape
        shift = np.empty(d, np.float32)
        scale = C.c_float()
        lib().or_quant_fit_mt(int(nthreads), _p(sample), n, d, C.c_float(target_peak), C.c_double(quantile),
                              int(power_of_two), _p(shift), C.byref(scale))
        return Quantizer(d, shift, float(scale.value))

    def encode(self, x, nthreads=1):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.d)
        out = np.empty(x.shape, np.int8)
        lib().or_quant_encode_mt(int(nthreads), _p(x), x.shape[0], self.d, _p(self.shift), C.c_float(self.scale),
                                 _p(out))
        return out

    def decode(self, codes):
        codes = np.ascontiguousarray(codes, np.float32).reshape(-1, self.d)
        out = np.empty(codes.shape, np.float32)
        lib().or_quant_decode(_p(codes), codes.shape[0], self.d, _p(self.shift), C.c_float(self.scale), _p(out))
        return out

    def decode_sqdist(self, v):
        return np.float32(np.float32(v) / np.float32(np.float32(self.scale) * np.float32(self.scale)))

    def clipped_fraction(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.d)
        return float(lib().or_quant_clipped_fraction(_p(x), x.shape[0], self.d, _p(self.shift),
                                                     C.c_float(self.scale)))


# ---- training (src/lib.rs:347-520) ----------------------------------------
@dataclass
class TrainResult:
    centroids: np.ndarray
    labels: np.ndarray
    objective: list
    n_splits: int
    recipe: Recipe = field(repr=False)


def train(x, k, cfg: Config | None = None, assigner=ASSIGNER_F32, simd=None, nthreads=None) -> TrainResult:
    x = np.ascontiguousarray(x)
    n, d = x.shape
    cent = np.empty((k, d), np.float32)
    labels = np.empty(n, np.uint32)
    out = TrainOut()
    simd = lib().or_simd_best() if simd is None else simd
    rc = lib().or_train(_p(x), ROWS[x.dtype], n, d, k, C.byref(cfg or Config()), assigner, simd,
                        nthreads or default_threads(), _p(cent), _p(labels), C.byref(out))
    if rc:
        raise ValueError(f"or_train failed ({rc})")
    obj = [out.objective_[i] for i in range(out.n_objective)]
    return TrainResult(cent, labels, obj, int(out.n_splits), out.recipe)


def assign(x, c, simd=None, nthreads=None):
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(c, np.float32)
    n, d = x.shape
    labels = np.empty(n, np.uint32)
    dist = np.empty(n, np.float32)
    simd = lib().or_simd_best() if simd is None else simd
    lib().or_assign(_p(x), ROWS[x.dtype], n, d, _p(c), c.shape[0], simd, nthreads or default_threads(),
                    _p(labels), _p(dist))
    return labels, dist


def balance_stats(labels, k):
    labels = np.ascontiguousarray(labels, np.uint32)
    a, b = C.c_double(), C.c_double()
    e = C.c_uint64()
    lib().or_balance_stats(_p(labels), len(labels), k, C.byref(a), C.byref(b), C.byref(e))
    return a.value, b.value, int(e.value)


def lcg_f32(n: int, seed: int) -> np.ndarray:
    """The reference tests' LCG data (src/cpu.rs:272-280, src/cuda.rs:226-234)."""
    s = np.uint32(seed)
    out = np.empty(n, np.float32)
    # vectorised LCG: s_{i+1} = a s_i + c (mod 2^32) via jump-ahead powers
    a, c = 1664525, 1013904223
    vals = np.empty(n, np.uint64)
    cur = int(s)
    for i in range(n):
        cur = (cur * a + c) & 0xFFFFFFFF
        vals[i] = cur
    out[:] = (vals >> 8).astype(np.float32) / np.float32(1 << 24) - np.float32(0.5)
    return out


def lcg_i8(n: int, seed: int) -> np.ndarray:
    """src/cpu_i8.rs:284-292: the top byte of the LCG state as i8."""
    a, c = 1664525, 1013904223
    cur = seed
    out = np.empty(n, np.int8)
    for i in range(n):
        cur = (cur * a + c) & 0xFFFFFFFF
        out[i] = np.uint8(cur >> 24).view(np.int8)
    return out
