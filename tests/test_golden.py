"""The committed golden fixtures (tests/golden/, written by make_golden.py
from the oracle) against the oracle, against the library's own host recipe,
and against the GPU searches.

The fixtures are the oracle's outputs: what pins them to the reference is
checked here too, against values fixed outside this repository — faiss
1.16.0's printed recipes (src/lib.rs:526-540), the published splitmix64
seed-0 outputs, and the exact integer brute force the reference's own I8
test uses (src/cpu_i8.rs:294-324)."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F32_SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 16), (37, 70, 33), (130, 257, 100), (64, 1030, 384), (9, 6, 1027)]
I8_SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 64), (37, 70, 33), (130, 257, 100), (20, 1030, 512), (9, 6, 1027)]


def _json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _npz(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_rng_fixture_reproduced():
    g = _json("rng.json")
    # splitmix64 from seed 0: Vigna's published first output
    assert int(g["splitmix64"]["0"][0]) == 0xE220A8397B1DCDAF
    for seed in (0, 1, 1234):
        s = str(seed)
        assert [str(int(v)) for v in O.splitmix64(seed, 4)] == g["splitmix64"][s]
        r = O.OracleRng(seed)
        assert [str(r.next_u64()) for _ in range(8)] == g["xoshiro256pp"][s]
        r = O.OracleRng(seed)
        assert [float(r.open01()).hex() for _ in range(4)] == g["open01_hex"][s]
        r = O.OracleRng(seed)
        for bound, want in g["below"][s]:
            got = r.below(bound)
            assert got == want and 0 <= got < bound
        p = [int(v) for v in O.permutation(seed, 10)]
        assert p == g["permutation"][s] and sorted(p) == list(range(10))
    p = O.permutation(7, 100_000).astype(np.uint64)
    assert str(int((p * np.arange(1, 100_001, dtype=np.uint64)).sum())) == g["permutation_7_100000_weighted_sum"]


def _recipe_fields(r):
    return {"batch_size": int(r.batch_size), "alpha0_hex": float(r.alpha0).hex(),
            "epoch_rows": [int(v) for v in r.epoch_rows], "total_steps": int(r.total_steps),
            "unused_threshold_hex": float(r.unused_threshold).hex()}


def test_recipe_fixture_reproduced_and_faiss_values():
    g = _json("recipes.json")
    for name, v in g.items():
        r = O.resolve_recipe(v["n"], v["d"], v["k"], O.Config(nepochs=v["nepochs"]))
        assert _recipe_fields(r) == {k: v[k] for k in _recipe_fields(r)}, name
    # faiss 1.16.0's printed values (src/lib.rs:526-540)
    assert g["lib_test_a"]["batch_size"] == 16384 and abs(float.fromhex(g["lib_test_a"]["alpha0_hex"]) - 0.728183) < 1e-5
    assert g["lib_test_b"]["epoch_rows"] == [1_000_000, 500_000, 1_000_000] and g["lib_test_b"]["batch_size"] == 4096
    assert abs(float.fromhex(g["lib_test_b"]["alpha0_hex"]) - 0.319215) < 1e-5
    assert g["lib_test_c"]["batch_size"] == 32 and abs(float.fromhex(g["lib_test_c"]["alpha0_hex"]) - 0.144604) < 1e-5
    assert float.fromhex(g["lib_test_d"]["unused_threshold_hex"]) == 0.0


def test_library_recipe_equals_fixture():
    """The library's host recipe (the C++ restatement the training driver
    runs) resolves every fixture shape to the same bits."""
    import emeans_hip as E

    g = _json("recipes.json")
    for name, v in g.items():
        r = E.resolve_recipe(v["n"], v["d"], v["k"], E.Config(nepochs=v["nepochs"]))
        got = {"batch_size": int(r.batch_size), "alpha0_hex": float(r.alpha0).hex(),
               "epoch_rows": [int(r.epoch_rows_[i]) for i in range(r.n_epochs)], "total_steps": int(r.total_steps),
               "unused_threshold_hex": float(r.unused_threshold).hex()}
        assert got == {k: v[k] for k in got}, name


def test_search_fixtures_reproduced():
    f32, i8 = _npz("search_f32.npz"), _npz("search_i8.npz")
    for n, k, d in F32_SHAPES:
        x = O.lcg_f32(n * d, 1).reshape(n, d)
        c = O.lcg_f32(k * d, 2).reshape(k, d)
        lab, best = O.brute_f64(x, c)
        assert np.array_equal(lab, f32[f"labels_{n}_{k}_{d}"]) and np.array_equal(best, f32[f"best_{n}_{k}_{d}"])
        # the CPU search of src/cpu.rs within its own bound of the fixture
        cl, cd = O.cpu_assign(x, c)
        got = O.dist_f64(x, c, cl)
        tol = 1e-4 * (1 + best)
        assert np.all(got <= best + tol) and np.all(np.abs(cd - best) <= tol)
    for n, k, d in I8_SHAPES:
        x = O.lcg_i8(n * d, 1).reshape(n, d)
        ci = O.lcg_i8(k * d, 2).reshape(k, d)
        want_l, want_d = i8[f"labels_{n}_{k}_{d}"], i8[f"dist_{n}_{k}_{d}"]
        # the exact integer brute force of src/cpu_i8.rs:300-306, in numpy
        diff = x.astype(np.int64)[:, None, :] - ci.astype(np.int64)[None, :, :]
        s = (diff * diff).sum(-1)
        assert np.array_equal(s.argmin(1), want_l) and np.array_equal(s.min(1), want_d)
        for bits in (O.BITS_I8, O.BITS_I16):
            gl, gd = O.i8_assign(x, ci.astype(np.float32), bits=bits)
            assert np.array_equal(gl, want_l) and np.array_equal(gd, want_d.astype(np.float32))


@pytest.mark.gpu
def test_gpu_searches_match_fixtures(hip):
    """The HIP searches on the fixture inputs: FP32 within src/cpu.rs's bound
    of the fixture's FP64 best; I8 and I16 labels and distances equal to the
    fixture's exact ones."""
    E = hip
    f32, i8 = _npz("search_f32.npz"), _npz("search_i8.npz")
    for n, k, d in F32_SHAPES:
        x = O.lcg_f32(n * d, 1).reshape(n, d)
        c = O.lcg_f32(k * d, 2).reshape(k, d)
        best = f32[f"best_{n}_{k}_{d}"]
        a = E.HipAssigner("fp32")
        a.set_centroids(c)
        lab, dist = a.assign(x)
        got = O.dist_f64(x, c, lab)
        tol = 1e-4 * (1 + best)
        assert np.all(got <= best + tol) and np.all(np.abs(dist - best) <= tol), (n, k, d)
    for n, k, d in I8_SHAPES:
        x = O.lcg_i8(n * d, 1).reshape(n, d)
        c = O.lcg_i8(k * d, 2).reshape(k, d).astype(np.float32)
        want_l, want_d = i8[f"labels_{n}_{k}_{d}"], i8[f"dist_{n}_{k}_{d}"].astype(np.float32)
        for prec in ("i8", "i16"):
            a = E.HipAssigner(prec)
            a.set_centroids(c)
            lab, dist = a.assign(x)
            assert np.array_equal(lab, want_l) and np.array_equal(dist, want_d), (prec, n, k, d)
