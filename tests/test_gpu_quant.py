"""Int8Quantizer on the GPU (src/quantize.rs:43-96): the reference's known-
answer test, and bit-exact agreement with the CPU oracle's restatement for
both calibrations on larger, heavy-tailed inputs with NaN and Inf values."""
import numpy as np
import pytest

import emeans_hip as E
import oracle as O

pytestmark = pytest.mark.gpu


def test_faiss_calibration_known_answer(hip):
    # src/quantize.rs:102-116
    sample = np.array([[15.0, -3.0], [5.0, -3.0], [10.0, -2.0], [10.0, -4.0]], np.float32)
    q = E.Quantizer.fit(sample)
    assert q.shift.tolist() == [10.0, -3.0]
    assert q.scale == 16.0
    codes = q.encode(sample)
    assert codes.reshape(-1).tolist() == [80, 0, -80, 0, 0, 16, 0, -16]
    back = q.decode(codes.astype(np.float32))
    assert np.array_equal(back, sample)
    assert q.encode(np.array([[1000.0, -3.0]], np.float32)).reshape(-1).tolist() == [127, 0]
    assert q.clipped_fraction(np.array([[1000.0, -3.0]], np.float32)) == 0.5


def _heavy(n, d, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_t(3, size=(n, d)) * rng.uniform(0.5, 20, size=d) + rng.normal(scale=50, size=d)).astype(np.float32)
    return x


# d = 37: the scalar kernels; d = 100 (25 quads, 10 rows per workgroup pass)
# and d = 1028 (257 quads, one row per pass): the 4-wide kernels and the
# LDS-staged mean with a partial last dimension strip and row tile
@pytest.mark.parametrize("d", [37, 100, 1028])
@pytest.mark.parametrize("cal", ["faiss", "clipped99", "clipped50", "clipped0"])
def test_quantizer_matches_cpu(hip, cal, d):
    x = _heavy(20000 if d < 1000 else 5003, d, 3)
    if cal == "faiss":
        args = dict(target_peak=112.0, quantile=1.0, power_of_two=True)
    else:
        args = dict(target_peak=127.0, quantile={"clipped99": 0.99, "clipped50": 0.5, "clipped0": 0.0}[cal],
                    power_of_two=False)
    print(f"quantizer {cal} d={d}", flush=True)
    g = E.Quantizer.fit(x, **args)
    c = O.Quantizer.fit(x, **args)
    assert np.array_equal(g.shift, c.shift) and g.scale == c.scale
    y = _heavy(3000, d, 4)
    y[0, 0], y[1, 1], y[2, 2] = np.nan, np.inf, -np.inf
    assert np.array_equal(g.encode(y), c.encode(y))
    codes = g.encode(y).astype(np.float32)
    assert np.array_equal(g.decode(codes), c.decode(codes))
    assert g.clipped_fraction(y) == c.clipped_fraction(y)


def test_quantizer_device_entry_points(hip):
    import torch

    x = _heavy(5000, 64, 5)
    xd = torch.from_numpy(x).cuda()
    shift = torch.empty(64, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    scale = E.quant_fit_device(xd, shift)
    codes = torch.empty(x.shape, dtype=torch.int8, device="cuda")
    E.quant_encode_device(xd, shift, scale, codes)
    torch.cuda.synchronize()
    c = O.Quantizer.fit(x)
    assert np.array_equal(shift.cpu().numpy(), c.shift) and scale == c.scale
    assert np.array_equal(codes.cpu().numpy(), c.encode(x))
