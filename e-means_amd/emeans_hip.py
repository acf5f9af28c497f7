"""Python mirror of the reference's plugin interface over libemeans_hip.so.

``HipAssigner`` has the methods of the reference traits ``Assigner`` and
``AssignRows<T>`` (src/lib.rs:75-91): ``set_centroids``, ``set_bias``,
``name`` and ``assign``; ``train`` / ``assign`` / ``resolve_recipe`` mirror
src/lib.rs:201-225 and 348-520. All compute runs in the HIP library; this
module only marshals arguments. Failures raise ``EmeansError``, and loading
the module without the built library raises too: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemeans_hip.so")

FP32, BF16_RESCORE, FP8_RESCORE, I8, I16 = 0, 1, 2, 3, 4
ROWS_F32, ROWS_I8, ROWS_U8 = 0, 1, 2
ROWS = {np.dtype(np.float32): ROWS_F32, np.dtype(np.int8): ROWS_I8, np.dtype(np.uint8): ROWS_U8}
E_OK, E_ARG, E_STATE, E_MEMORY, E_DEVICE, E_UNSUPPORTED = 0, 1, 2, 3, 4, 5
PRECISIONS = {"fp32": FP32, "bf16": BF16_RESCORE, "fp8": FP8_RESCORE, "i8": I8, "i16": I16}

EXPORTS = [
    "emh_create", "emh_destroy", "emh_set_centroids", "emh_set_bias", "emh_name", "emh_assign_f32",
    "emh_assign_i8", "emh_assign_u8", "emh_last_error", "emh_set_stream", "emh_assign_device",
    "emh_synchronize", "emh_accumulate_device", "emh_config_default", "emh_resolve_recipe", "emh_train",
    "emh_assign_all", "emh_set_timing", "emh_get_timing", "emh_train_batch", "emh_set_split", "emh_train_device",
    "emh_quant_fit", "emh_quant_encode", "emh_quant_decode", "emh_quant_clipped_fraction", "emh_quant_fit_device",
    "emh_quant_encode_device", "emh_report_nearest", "emh_report_nearest_device", "emh_report_config_default",
    "emh_report", "emh_assign_device_noalloc", "emh_set_tile", "emh_last_tile",
]


class EmeansError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


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

    def __init__(self, nepochs=5.0, batch_law=0, batch_size=None, alpha0=None, unused_threshold=None,
                 balance=0.0, seed=1234):
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
        ("seconds", C.c_double),
    ]


_lib = None


def lib():
    """Load the HIP library; raise if it was not built (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python e-means_amd/build.py` (or __graft_entry__.build())")
    # PyTorch-ROCm ships its own HIP runtime under the same soname
    # (libamdhip64.so.7); the first one loaded serves the whole process and
    # torch only initialises its GPU with its own. Load torch's first when it
    # is installed, so that torch tensors and this library share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, sz, i32, u64 = C.c_void_p, C.c_size_t, C.c_int, C.c_uint64
    L.emh_create.argtypes = [i32, i32, i32, C.POINTER(C.c_void_p)]
    L.emh_destroy.argtypes = [vp]
    L.emh_destroy.restype = None
    L.emh_set_centroids.argtypes = [vp, vp, sz, sz]
    L.emh_set_bias.argtypes = [vp, vp, sz]
    L.emh_name.argtypes = [vp, C.c_char_p, sz]
    for f in ("emh_assign_f32", "emh_assign_i8", "emh_assign_u8"):
        getattr(L, f).argtypes = [vp, vp, sz, vp, vp]
    L.emh_last_error.restype = C.c_char_p
    L.emh_set_stream.argtypes = [vp, vp]
    L.emh_assign_device.argtypes = [vp, vp, i32, sz, vp, vp, vp]
    L.emh_assign_device_noalloc.argtypes = [vp, vp, i32, sz, vp, vp, vp]
    L.emh_synchronize.argtypes = [vp]
    L.emh_set_timing.argtypes = [vp, i32]
    L.emh_get_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    L.emh_accumulate_device.argtypes = [vp, vp, i32, sz, vp, vp, vp, vp]
    L.emh_config_default.argtypes = [C.POINTER(Config)]
    L.emh_config_default.restype = None
    L.emh_resolve_recipe.argtypes = [u64, u64, u64, C.POINTER(Config), C.POINTER(Recipe)]
    L.emh_train.argtypes = [vp, vp, i32, sz, sz, sz, C.POINTER(Config), vp, vp, C.POINTER(TrainOut)]
    L.emh_train_device.argtypes = [vp, vp, i32, sz, sz, sz, C.POINTER(Config), vp, vp, C.POINTER(TrainOut)]
    L.emh_assign_all.argtypes = [vp, vp, i32, sz, sz, vp, sz, vp, vp]
    L.emh_train_batch.argtypes = [vp, C.POINTER(Job), sz, C.POINTER(BatchStats)]
    L.emh_set_split.argtypes = [vp, i32]
    L.emh_set_tile.argtypes = [vp, i32]
    L.emh_last_tile.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    f32 = C.c_float
    L.emh_quant_fit.argtypes = [i32, vp, u64, u64, f32, C.c_double, i32, vp, C.POINTER(f32)]
    L.emh_quant_encode.argtypes = [i32, vp, u64, u64, vp, f32, vp]
    L.emh_quant_decode.argtypes = [i32, vp, u64, u64, vp, f32, vp]
    L.emh_quant_clipped_fraction.argtypes = [i32, vp, u64, u64, vp, f32, C.POINTER(C.c_double)]
    L.emh_quant_fit_device.argtypes = [vp, u64, u64, f32, C.c_double, i32, vp, C.POINTER(f32), vp]
    L.emh_quant_encode_device.argtypes = [vp, u64, u64, vp, f32, vp, vp]
    L.emh_report_nearest.argtypes = [C.c_int, vp, u64, u64, vp, vp, vp]
    L.emh_report_nearest_device.argtypes = [vp, u64, u64, vp, vp, vp, vp, vp]
    L.emh_report_config_default.argtypes = [C.POINTER(ReportConfig)]
    L.emh_report.argtypes = [C.c_int, vp, vp, u64, vp, u64, u64, C.POINTER(ReportConfig), vp, vp, vp, vp, vp, vp,
                             vp, vp, C.POINTER(ReportSummary)]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise EmeansError(rc, lib().emh_last_error().decode())


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def require_gpu():
    """Fail loudly unless the HIP library loads and a GPU is visible."""
    L = lib()
    h = C.c_void_p()
    rc = L.emh_create(0, FP32, 0, C.byref(h))
    if rc != 0:
        raise EmeansError(rc, L.emh_last_error().decode())
    L.emh_destroy(h)


class HipAssigner:
    """``impl Assigner + AssignRows<T>`` on one GPU (src/lib.rs:75-91)."""

    def __init__(self, precision=FP32, device: int = 0, rescore_m: int = 0):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        self._h = C.c_void_p()
        _check(lib().emh_create(device, precision, rescore_m, C.byref(self._h)))
        self.precision = precision
        self.k = self.d = 0

    def close(self):
        if self._h:
            lib().emh_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_centroids(self, centroids):
        c = np.ascontiguousarray(centroids, np.float32)
        if c.ndim != 2:
            raise ValueError("centroids must be k x d")
        _check(lib().emh_set_centroids(self._h, _p(c), c.shape[0], c.shape[1]))
        self.k, self.d = c.shape

    def set_bias(self, bias):
        b = np.ascontiguousarray(bias, np.float32)
        _check(lib().emh_set_bias(self._h, _p(b), b.size))

    def name(self) -> str:
        buf = C.create_string_buffer(512)
        _check(lib().emh_name(self._h, buf, 512))
        return buf.value.decode()

    def assign(self, x):
        x = np.ascontiguousarray(x)
        if x.dtype not in ROWS:
            raise TypeError(f"rows must be f32, i8 or u8, got {x.dtype}")
        if x.ndim != 2 or (self.d and x.shape[1] != self.d):
            raise ValueError("x must be m x d with the centroids' d")
        m = x.shape[0]
        labels = np.empty(m, np.uint32)
        dist = np.empty(m, np.float32)
        fn = {ROWS_F32: "emh_assign_f32", ROWS_I8: "emh_assign_i8", ROWS_U8: "emh_assign_u8"}[ROWS[x.dtype]]
        _check(getattr(lib(), fn)(self._h, _p(x), m, _p(labels), _p(dist)))
        return labels, dist

    # ---- device-pointer extensions (torch tensors as HBM plumbing) ----
    def set_stream(self, stream_handle: int):
        _check(lib().emh_set_stream(self._h, C.c_void_p(stream_handle)))

    def assign_device(self, x, labels, dist, ids=None):
        """x, labels, dist, ids: torch CUDA tensors; asynchronous."""
        rows = {"torch.float32": ROWS_F32, "torch.int8": ROWS_I8, "torch.uint8": ROWS_U8}[str(x.dtype)]
        m = labels.shape[0]
        _check(lib().emh_assign_device(self._h, C.c_void_p(x.data_ptr()), rows, m,
                                       None if ids is None else C.c_void_p(ids.data_ptr()),
                                       C.c_void_p(labels.data_ptr()), C.c_void_p(dist.data_ptr())))

    def assign_device_noalloc(self, x, labels, dist, ids=None):
        """assign_device without allocating: the row-side buffers live in the
        library's static per-GPU scratch (EMH_NOALLOC_SCRATCH_BYTES)."""
        rows = {"torch.float32": ROWS_F32, "torch.int8": ROWS_I8, "torch.uint8": ROWS_U8}[str(x.dtype)]
        m = labels.shape[0]
        _check(lib().emh_assign_device_noalloc(self._h, C.c_void_p(x.data_ptr()), rows, m,
                                               None if ids is None else C.c_void_p(ids.data_ptr()),
                                               C.c_void_p(labels.data_ptr()), C.c_void_p(dist.data_ptr())))

    def accumulate_device(self, x, labels, sums, counts, ids=None):
        rows = {"torch.float32": ROWS_F32, "torch.int8": ROWS_I8, "torch.uint8": ROWS_U8}[str(x.dtype)]
        _check(lib().emh_accumulate_device(self._h, C.c_void_p(x.data_ptr()), rows, labels.shape[0],
                                           None if ids is None else C.c_void_p(ids.data_ptr()),
                                           C.c_void_p(labels.data_ptr()), C.c_void_p(sums.data_ptr()),
                                           C.c_void_p(counts.data_ptr())))

    def synchronize(self):
        _check(lib().emh_synchronize(self._h))

    def set_split(self, nslice: int):
        """Centroid-range split of the search: 0 = autotuned, n >= 1 forced."""
        _check(lib().emh_set_split(self._h, int(nslice)))

    def set_tile(self, fragments: int):
        """Rows per workgroup of the search, 128 x fragments: 0 = autotuned, 1 or 2 forced."""
        _check(lib().emh_set_tile(self._h, int(fragments)))

    def last_tile(self):
        """(fragments per wave, centroid slices) of the last search launch."""
        fr, ns = C.c_int(), C.c_int()
        _check(lib().emh_last_tile(self._h, C.byref(fr), C.byref(ns)))
        return fr.value, ns.value

    def set_timing(self, on: bool):
        _check(lib().emh_set_timing(self._h, int(on)))

    def get_timing(self):
        """(row-encode ms, search ms, rescore ms, calls) summed since the last read."""
        ms = (C.c_double * 3)()
        calls = C.c_uint64()
        _check(lib().emh_get_timing(self._h, ms, C.byref(calls)))
        return ms[0], ms[1], ms[2], int(calls.value)


def resolve_recipe(n: int, d: int, k: int, cfg: Config | None = None) -> Recipe:
    r = Recipe()
    _check(lib().emh_resolve_recipe(n, d, k, C.byref(cfg or Config()), C.byref(r)))
    return r


@dataclass
class Output:
    """Mirror of ``Output`` (src/lib.rs:151-163)."""

    centroids: np.ndarray
    labels: np.ndarray
    objective: list
    n_splits: int
    recipe: Recipe = field(repr=False)
    seconds: float = 0.0


def _device_rows(x):
    """(pointer, row type) of a contiguous 2-D torch CUDA tensor, else None."""
    if type(x).__module__.split(".")[0] != "torch" or not getattr(x, "is_cuda", False):
        return None
    import torch

    kinds = {torch.float32: ROWS_F32, torch.int8: ROWS_I8, torch.uint8: ROWS_U8}
    if x.dtype not in kinds or x.dim() != 2 or not x.is_contiguous():
        raise TypeError("device rows must be a contiguous 2-D f32, i8 or u8 CUDA tensor")
    return x.data_ptr(), kinds[x.dtype]


def train(x, k: int, cfg: Config | None = None, assigner: HipAssigner | None = None) -> Output:
    """train (src/lib.rs:348-491) with the HIP assigner (FP32 by default).
    x: a host array, or a torch CUDA tensor whose rows are used in place
    (emh_train_device: no copy)."""
    a = assigner or HipAssigner(FP32)
    dev = _device_rows(x)
    if dev is None:
        x = np.ascontiguousarray(x)
        if x.dtype not in ROWS or x.ndim != 2:
            raise TypeError("x must be a 2-D f32, i8 or u8 array")
    n, d = x.shape
    cent = np.empty((k, d), np.float32)
    labels = np.empty(n, np.uint32)
    out = TrainOut()
    if dev is None:
        _check(lib().emh_train(a.handle, _p(x), ROWS[x.dtype], n, d, k, C.byref(cfg or Config()), _p(cent),
                               _p(labels), C.byref(out)))
    else:
        import torch

        torch.cuda.synchronize()  # the tensor's producers, on torch's streams
        _check(lib().emh_train_device(a.handle, C.c_void_p(dev[0]), dev[1], n, d, k, C.byref(cfg or Config()),
                                      _p(cent), _p(labels), C.byref(out)))
    obj = [out.objective_[i] for i in range(out.n_objective)]
    return Output(cent, labels, obj, int(out.n_splits), out.recipe, float(out.seconds))


class Job(C.Structure):
    """emh_job: one train() call's arguments for emh_train_batch."""

    _fields_ = [("x", C.c_void_p), ("rows", C.c_int), ("n", C.c_uint64), ("d", C.c_uint64), ("k", C.c_uint64),
                ("cfg", Config), ("centroids_out", C.c_void_p), ("labels_out", C.c_void_p),
                ("out", C.POINTER(TrainOut)), ("x_on_device", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("graph_nodes", C.c_uint64), ("seconds", C.c_double),
                ("host_seconds", C.c_double), ("wait_seconds", C.c_double), ("launch_seconds", C.c_double)]


def train_batch(problems, assigner: HipAssigner | None = None):
    """Trains independent problems in lockstep, one captured-graph replay per
    training step (emh_train_batch). problems: (x, k, cfg) tuples with f32
    rows (i8 rows for the I8 and I16 assigners), host arrays or torch CUDA
    tensors used in place. Each result equals train(x, k, cfg, assigner) on
    the same assigner precision. Returns (list of Output, BatchStats)."""
    a = assigner or HipAssigner(FP32)
    want = np.dtype(np.int8) if a.precision in (I8, I16) else np.dtype(np.float32)
    kind = ROWS[want]
    keep, jobs = [], (Job * len(problems))()
    any_dev = False
    for i, (x, k, cfg) in enumerate(problems):
        dev = _device_rows(x)
        if dev is None:
            if want == np.int8 and np.asarray(x).dtype != want:
                raise TypeError("train_batch with the integer assigners takes i8 rows")
            x = np.ascontiguousarray(x, want)
        elif dev[1] != kind:
            raise TypeError(f"train_batch takes {want.name} rows with this assigner")
        any_dev = any_dev or dev is not None
        n, d = x.shape
        cent = np.empty((k, d), np.float32)
        labels = np.empty(n, np.uint32)
        out = TrainOut()
        keep.append((x, cent, labels, out))
        ptr = _p(x) if dev is None else C.c_void_p(dev[0])
        jobs[i] = Job(ptr, kind, n, d, k, cfg or Config(), _p(cent), _p(labels), C.pointer(out),
                      int(dev is not None))
    if any_dev:
        import torch

        torch.cuda.synchronize()
    st = BatchStats()
    _check(lib().emh_train_batch(a.handle, jobs, len(problems), C.byref(st)))
    res = []
    for x, cent, labels, out in keep:
        obj = [out.objective_[i] for i in range(out.n_objective)]
        res.append(Output(cent, labels, obj, int(out.n_splits), out.recipe, float(out.seconds)))
    return res, st


class Quantizer:
    """Int8Quantizer (src/quantize.rs:43-96) computed on the GPU.

    Calibration defaults to faiss's (target peak 112, quantile 1, power-of-two
    scale); ``Quantizer.clipped(q)`` is Calibration::clipped."""

    def __init__(self, shift, scale: float):
        self.shift = np.ascontiguousarray(shift, np.float32)
        self.scale = float(scale)
        self.d = self.shift.shape[0]

    @classmethod
    def fit(cls, sample, target_peak=112.0, quantile=1.0, power_of_two=True, device=0):
        x = np.ascontiguousarray(sample, np.float32)
        n, d = x.shape
        shift = np.empty(d, np.float32)
        scale = C.c_float()
        _check(lib().emh_quant_fit(device, _p(x), n, d, target_peak, quantile, int(power_of_two), _p(shift),
                                   C.byref(scale)))
        return cls(shift, scale.value)

    @classmethod
    def fit_clipped(cls, sample, quantile, device=0):
        return cls.fit(sample, 127.0, quantile, False, device)

    def encode(self, x, device=0):
        x = np.ascontiguousarray(x, np.float32)
        codes = np.empty(x.shape, np.int8)
        _check(lib().emh_quant_encode(device, _p(x), x.shape[0], self.d, _p(self.shift), self.scale, _p(codes)))
        return codes

    def decode(self, codes, device=0):
        c = np.ascontiguousarray(codes, np.float32)
        out = np.empty(c.shape, np.float32)
        _check(lib().emh_quant_decode(device, _p(c), c.shape[0], self.d, _p(self.shift), self.scale, _p(out)))
        return out

    def clipped_fraction(self, x, device=0):
        x = np.ascontiguousarray(x, np.float32)
        out = C.c_double()
        _check(lib().emh_quant_clipped_fraction(device, _p(x), x.shape[0], self.d, _p(self.shift), self.scale,
                                                C.byref(out)))
        return out.value


def quant_fit_device(x, shift, target_peak=112.0, quantile=1.0, power_of_two=True, stream=0):
    """Fit on a torch CUDA tensor (n x d f32); shift: a d-float CUDA tensor."""
    scale = C.c_float()
    _check(lib().emh_quant_fit_device(C.c_void_p(x.data_ptr()), x.shape[0], x.shape[1], target_peak, quantile,
                                      int(power_of_two), C.c_void_p(shift.data_ptr()), C.byref(scale),
                                      C.c_void_p(stream)))
    return scale.value


def quant_encode_device(x, shift, scale, codes, stream=0):
    _check(lib().emh_quant_encode_device(C.c_void_p(x.data_ptr()), x.shape[0], x.shape[1],
                                         C.c_void_p(shift.data_ptr()), scale, C.c_void_p(codes.data_ptr()),
                                         C.c_void_p(stream)))


def report_nearest(centroids, sizes=None, device=0):
    """The report's nearest-centroid step (src/report.rs:169-183): for each
    centroid the nearest live other centroid (sizes[j] > 0; all live when
    sizes is None) and the L2 distance to it; (i, inf) when there is none."""
    c = np.ascontiguousarray(centroids, np.float32)
    if c.ndim != 2:
        raise ValueError("centroids must be k x d")
    k, d = c.shape
    sz = None
    if sizes is not None:
        sz = np.ascontiguousarray(sizes, np.uint32)
        if sz.shape != (k,):
            raise ValueError("sizes must hold one count per centroid")
    nearest = np.empty(k, np.uint32)
    dist = np.empty(k, np.float32)
    _check(lib().emh_report_nearest(device, _p(c), k, d, _p(sz), _p(nearest), _p(dist)))
    return nearest, dist


def report_nearest_device(centroids, sizes, keys, nearest, dist, stream=0):
    """report_nearest on torch CUDA tensors: centroids k x d f32, sizes k u32
    (int32 storage) or None, keys k int64 workspace, nearest k int32, dist k
    f32; stream-ordered, no synchronisation."""
    k, d = centroids.shape
    for t, n in ((keys, k), (nearest, k), (dist, k)):
        if t.numel() != n or not t.is_contiguous():
            raise ValueError("keys, nearest and dist must be contiguous with k elements")
    _check(lib().emh_report_nearest_device(C.c_void_p(centroids.data_ptr()), k, d,
                                           None if sizes is None else C.c_void_p(sizes.data_ptr()),
                                           C.c_void_p(keys.data_ptr()), C.c_void_p(nearest.data_ptr()),
                                           C.c_void_p(dist.data_ptr()), C.c_void_p(stream)))


class ReportConfig(C.Structure):
    """ReportConfig (src/report.rs:20-35)."""
    _fields_ = [("family_radius", C.c_float), ("min_local_size", C.c_uint64)]

    def __init__(self, family_radius=1.25, min_local_size=30):
        super().__init__(family_radius, min_local_size)


class ReportSummary(C.Structure):
    _fields_ = [("median_dist", C.c_float), ("median_rms", C.c_float), ("median_nearest_dist", C.c_float),
                ("n_families", C.c_uint32)]


# ClusterStats (src/report.rs:37-56), field by field
CLUSTER_STATS = np.dtype([("size", "<u4"), ("rms", "<f4"), ("median", "<f4"), ("mad", "<f4"), ("p90", "<f4"),
                          ("p99", "<f4"), ("max", "<f4"), ("looseness", "<f4"), ("nearest_centroid", "<u4"),
                          ("nearest_dist", "<f4"), ("isolation", "<f4"), ("family", "<u4")])


def _total_key(v):
    u = np.ascontiguousarray(v, np.float32).view(np.uint32)
    return np.where(u & np.uint32(0x80000000), ~u, u | np.uint32(0x80000000))


def _rust_fmt(v, prec):
    """Rust's `{:.prec}` for a float (exact decimal rounding, "NaN"/"inf")."""
    v = float(v)
    if v != v:
        return "NaN"
    if v in (float("inf"), float("-inf")):
        return "inf" if v > 0 else "-inf"
    return f"{v:.{prec}f}"


class Family:
    def __init__(self, leader, members, size, share):
        self.leader, self.members, self.size, self.share = leader, members, size, share

    def __repr__(self):
        return f"Family(leader={self.leader}, members={self.members.tolist()}, size={self.size}, share={self.share})"


class Report:
    """report::Report (src/report.rs:58-81) with its methods (src/report.rs:245-291)."""

    def __init__(self, clusters, families, local_z, global_z, summary):
        self.clusters = clusters  # structured array of CLUSTER_STATS, one per cluster
        self.families = families  # list of Family, largest first
        self.local_z = local_z
        self.global_z = global_z
        self.median_dist = np.float32(summary.median_dist)
        self.median_rms = np.float32(summary.median_rms)
        self.median_nearest_dist = np.float32(summary.median_nearest_dist)

    def score(self, row):
        """max(local_z, global_z) with f32::max's NaN rule (src/report.rs:247-249)."""
        return np.fmax(self.local_z[row], self.global_z[row])

    def top_outliers(self, count):
        """(row, score) of the `count` highest scores, best first
        (src/report.rs:252-261); equal scores by row index."""
        scores = np.fmax(self.local_z, self.global_z)
        order = np.argsort(~_total_key(scores), kind="stable")[: min(count, scores.shape[0])]
        return [(int(r), np.float32(scores[r])) for r in order]

    def rare_prototypes(self, max_size_ratio, min_isolation):
        """(cluster, size, isolation) of small, isolated clusters, most
        isolated first (src/report.rs:265-277); equal isolation by index."""
        c = self.clusters
        live = max(int((c["size"] > 0).sum()), 1)
        mean = np.float32(np.float32(self.local_z.shape[0]) / np.float32(live))
        lim = np.float32(np.float32(max_size_ratio) * mean)
        keep = np.flatnonzero((c["size"] > 0) & (c["size"].astype(np.float32) <= lim) &
                              (c["isolation"] >= np.float32(min_isolation)))
        keep = keep[np.argsort(~_total_key(c["isolation"][keep]), kind="stable")]
        return [(int(j), int(c["size"][j]), np.float32(c["isolation"][j])) for j in keep]

    def write_clusters_tsv(self, w):
        """One line per cluster after a header (src/report.rs:280-290)."""
        w.write("cluster\tsize\trms\tmedian\tmad\tp90\tp99\tmax\tlooseness\tnearest_centroid\tnearest_dist"
                "\tisolation\tfamily\tfamily_share\n")
        for j, c in enumerate(self.clusters):
            share = self.families[int(c["family"])].share if c["size"] > 0 else 0.0
            f5 = [_rust_fmt(c[f], 5) for f in ("rms", "median", "mad", "p90", "p99", "max")]
            w.write(f"{j}\t{int(c['size'])}\t" + "\t".join(f5) + f"\t{_rust_fmt(c['looseness'], 3)}"
                    f"\t{int(c['nearest_centroid'])}\t{_rust_fmt(c['nearest_dist'], 5)}"
                    f"\t{_rust_fmt(c['isolation'], 3)}\t{int(c['family'])}\t{_rust_fmt(share, 6)}\n")


def report(labels, sqdist, centroids, cfg: ReportConfig | None = None, device=0) -> Report:
    """report::report (src/report.rs:119-243): the nearest-centroid and family
    radius steps on the GPU, the rest on the host in the reference's
    arithmetic."""
    labels = np.ascontiguousarray(labels, np.uint32)
    sq = np.ascontiguousarray(sqdist, np.float32)
    c = np.ascontiguousarray(centroids, np.float32)
    if c.ndim != 2 or labels.shape != sq.shape or labels.ndim != 1:
        raise ValueError("report: labels and sqdist are n-vectors, centroids k x d")
    n = labels.shape[0]
    k, d = c.shape
    cfg = cfg or ReportConfig()
    clusters = np.zeros(k, CLUSTER_STATS)
    local_z = np.empty(n, np.float32)
    global_z = np.empty(n, np.float32)
    leader = np.empty(k, np.uint32)
    fsize = np.empty(k, np.uint64)
    share = np.empty(k, np.float64)
    start = np.empty(k + 1, np.uint32)
    members = np.empty(k, np.uint32)
    summ = ReportSummary()
    _check(lib().emh_report(device, _p(labels), _p(sq), n, _p(c), k, d, C.byref(cfg), _p(clusters), _p(local_z),
                            _p(global_z), _p(leader), _p(fsize), _p(share), _p(start), _p(members), C.byref(summ)))
    fams = [Family(int(leader[f]), members[start[f]:start[f + 1]].copy(), int(fsize[f]), float(share[f]))
            for f in range(summ.n_families)]
    return Report(clusters, fams, local_z, global_z, summ)


def read_f32(path, d):
    """A raw little-endian f32 file of n x d rows (examples/cluster.rs:11-14),
    memory-mapped: n = file size / (4 d)."""
    a = np.memmap(path, dtype="<f4", mode="r")
    if a.shape[0] % d:
        raise ValueError(f"{path}: {a.shape[0]} floats is not a multiple of d = {d}")
    return a.reshape(-1, d)


def write_f32(path, a):
    """Rows or centroids as raw little-endian f32 (examples/cluster.rs:22-23)."""
    np.ascontiguousarray(a, "<f4").tofile(path)


def read_u32(path):
    """Labels as raw little-endian u32 (examples/outliers.rs:70-71)."""
    return np.fromfile(path, dtype="<u4")


def write_u32(path, labels):
    np.ascontiguousarray(labels, "<u4").tofile(path)


def assign(x, centroids, assigner: HipAssigner | None = None):
    """assign (src/lib.rs:512-520): exact labels and squared distances."""
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(centroids, np.float32)
    a = assigner or HipAssigner(FP32)
    n, d = x.shape
    labels = np.empty(n, np.uint32)
    dist = np.empty(n, np.float32)
    _check(lib().emh_assign_all(a.handle, _p(x), ROWS[x.dtype], n, d, _p(c), c.shape[0], _p(labels), _p(dist)))
    return labels, dist
