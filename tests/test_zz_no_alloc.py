"""emh_assign_device_noalloc against emh_assign_device, through the C ABI.

The zero-allocation entry point runs the same kernels on buffers carved
from the library's static per-GPU scratch array; its labels and distances
must equal the allocating entry point's bit for bit, on two calls in a row
(the second smaller, so the scratch keeps the first call's contents), and
it must leave the GPU's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest

import emeans_hip as E
import oracle as O

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _mixture(n, d, blobs, seed):
    rng = np.random.default_rng(seed)
    centres = rng.normal(scale=4.0, size=(blobs, d)).astype(np.float32)
    return (centres[rng.integers(0, blobs, n)] + rng.normal(size=(n, d))).astype(np.float32)


def _call(fn, a, x, m, ids=None):
    import torch

    labels = torch.full((m,), -1, dtype=torch.int32, device="cuda")
    dist = torch.full((m,), -1.0, dtype=torch.float32, device="cuda")
    rows = {torch.float32: E.ROWS_F32, torch.int8: E.ROWS_I8, torch.uint8: E.ROWS_U8}[x.dtype]
    rc = getattr(E.lib(), fn)(a.handle, C.c_void_p(x.data_ptr()), rows, m,
                              None if ids is None else C.c_void_p(ids.data_ptr()),
                              C.c_void_p(labels.data_ptr()), C.c_void_p(dist.data_ptr()))
    E._check(rc)
    a.synchronize()
    return labels.cpu().numpy().view(np.uint32), dist.cpu().numpy()


@pytest.mark.parametrize("prec,m_top,d,k", [("bf16", 4, 128, 2000), ("fp8", 8, 256, 1500), ("fp32", 1, 96, 700),
                                            ("i8", 1, 512, 300), ("i16", 1, 64, 257), ("bf16", 8, 300, 400)])
def test_noalloc_equals_assign_device(hip, prec, m_top, d, k):
    import torch

    print(f"noalloc {prec} top-{m_top} d={d} k={k}", flush=True)
    n1, n2 = 50_000, 3_001
    if prec in ("i8", "i16"):
        rng = np.random.default_rng(d)
        xh = rng.integers(-128, 128, size=(n1, d)).astype(np.int8)
        ch = rng.normal(scale=40, size=(k, d)).astype(np.float32)
    else:
        xh = _mixture(n1, d, 700, d)
        ch = xh[O.permutation(1234, n1)[:k]].copy()
    x = torch.from_numpy(xh).cuda()
    a = E.HipAssigner(prec, rescore_m=m_top)  # two handles: the reference one allocates
    b = E.HipAssigner(prec, rescore_m=m_top)
    a.set_centroids(ch)
    b.set_centroids(ch)
    ids = torch.from_numpy(O.permutation(7, n1)[:n2].astype(np.int32)).cuda()
    for m, sel in ((n1, None), (n2, ids)):  # the second call is smaller and gathers rows by id
        want = _call("emh_assign_device", a, x, m, sel)
        got = _call("emh_assign_device_noalloc", b, x, m, sel)
        assert np.array_equal(got[0], want[0]), f"{prec}: labels differ (m={m})"
        assert got[1].tobytes() == want[1].tobytes(), f"{prec}: distances differ (m={m})"
        assert got[0].max() < k


def test_noalloc_allocates_nothing(hip):
    """A no-allocation call on 1M rows (16 MB of candidate lists at BF16
    top-4) leaves the device's free memory as it was. A small call first
    loads the kernels, so one-time runtime set-up is not counted."""
    import torch

    n, d, k = 1_000_000, 64, 512
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(n, d, generator=g, device="cuda")
    labels = torch.empty(n, dtype=torch.int32, device="cuda")
    dist = torch.empty(n, dtype=torch.float32, device="cuda")
    a = E.HipAssigner("bf16", rescore_m=4)
    a.set_centroids(x[:k].cpu().numpy())
    a.assign_device_noalloc(x[:1000], labels[:1000], dist[:1000])
    a.synchronize()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    a.assign_device_noalloc(x, labels, dist)
    a.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 == free0, f"free memory moved by {free0 - free1} bytes"
    assert int(labels.max()) < k


def test_noalloc_rejects_rows_past_the_scratch(hip):
    """FP8 top-8 at d = 600 (rows over 512 bytes, encoded into the scratch)
    needs 32 + 640 + 8 bytes per row: 200,000 rows exceed the 64 MiB
    scratch and are refused before any launch."""
    import torch

    n, d = 200_000, 600
    x = torch.zeros(n, d, device="cuda")
    labels = torch.empty(n, dtype=torch.int32, device="cuda")
    dist = torch.empty(n, dtype=torch.float32, device="cuda")
    a = E.HipAssigner("fp8", rescore_m=8)
    a.set_centroids(np.ones((64, d), np.float32))
    with pytest.raises(E.EmeansError) as ei:
        a.assign_device_noalloc(x, labels, dist)
    assert ei.value.code == E.E_ARG and "scratch" in str(ei.value)
    a.assign_device_noalloc(x[:1000], labels[:1000], dist[:1000])  # a fitting call still works
    a.synchronize()
    assert int(labels[:1000].max()) < 64
