"""Writes the golden fixtures of this directory from the oracle (the C
restatement of the reference crate in oracle/). They pin the restatement's
outputs so that a later change to it, or to the library that must agree with
it, shows up as a diff:

- rng.json: splitmix64 and xoshiro256++ outputs, open01 and Lemire draws and
  Fisher-Yates permutations (src/lib.rs:227-274) for seeds 0, 1 and 1234;
- recipes.json: the resolved recipe (src/lib.rs:165-225) of each benchmark
  configuration and of the reference's own recipe test shapes
  (src/lib.rs:526-540);
- search_f32.npz: FP64 brute-force labels and best distances of the seven
  shapes of src/cpu.rs:284-309 on the reference tests' LCG data;
- search_i8.npz: exact integer labels and distances of the seven shapes of
  src/cpu_i8.rs:294-324 (integer-valued centroids), which CentroidBits::I8
  and I16 must reproduce.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

F32_SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 16), (37, 70, 33), (130, 257, 100), (64, 1030, 384), (9, 6, 1027)]
I8_SHAPES = [(1, 1, 1), (5, 3, 7), (13, 9, 64), (37, 70, 33), (130, 257, 100), (20, 1030, 512), (9, 6, 1027)]
# (name, n, d, k, nepochs): the benchmark configurations' training shapes and
# src/lib.rs:526-540's recipe test shapes
RECIPES = [("config1", 200_000, 96, 1024, 5.0), ("config3", 1_000_000, 512, 4096, 5.0),
           ("config4", 10_000_000, 256, 65536, 2.0), ("config5_small", 5_000, 64, 64, 3.0),
           ("config5_large", 500_000, 256, 4096, 3.0), ("lib_test_a", 200_000, 96, 1024, 5.0),
           ("lib_test_b", 1_000_000, 8, 100, 2.5), ("lib_test_c", 4000, 4, 8, 5.0), ("lib_test_d", 1000, 2, 10, 5.0),
           ("fractional_epochs", 12_345, 20, 100, 2.5)]


def rng_fixture():
    out = {"splitmix64": {}, "xoshiro256pp": {}, "open01_hex": {}, "below": {}, "permutation": {}}
    bounds = [1, 2, 3, 10, 1000, 2**33 + 1]
    for seed in (0, 1, 1234):
        out["splitmix64"][str(seed)] = [str(int(v)) for v in O.splitmix64(seed, 4)]
        r = O.OracleRng(seed)
        out["xoshiro256pp"][str(seed)] = [str(r.next_u64()) for _ in range(8)]
        r = O.OracleRng(seed)
        out["open01_hex"][str(seed)] = [float(r.open01()).hex() for _ in range(4)]
        r = O.OracleRng(seed)
        out["below"][str(seed)] = [[b, r.below(b)] for b in bounds for _ in range(2)]
        out["permutation"][str(seed)] = [int(v) for v in O.permutation(seed, 10)]
    p = O.permutation(7, 100_000).astype(np.uint64)
    out["permutation_7_100000_weighted_sum"] = str(int((p * np.arange(1, 100_001, dtype=np.uint64)).sum()))
    return out


def recipe_fixture():
    out = {}
    for name, n, d, k, ep in RECIPES:
        r = O.resolve_recipe(n, d, k, O.Config(nepochs=ep))
        out[name] = {"n": n, "d": d, "k": k, "nepochs": ep, "batch_size": int(r.batch_size),
                     "alpha0_hex": float(r.alpha0).hex(), "epoch_rows": [int(v) for v in r.epoch_rows],
                     "total_steps": int(r.total_steps), "unused_threshold_hex": float(r.unused_threshold).hex()}
    return out


def search_f32_fixture():
    arrs = {}
    for n, k, d in F32_SHAPES:
        x = O.lcg_f32(n * d, 1).reshape(n, d)
        c = O.lcg_f32(k * d, 2).reshape(k, d)
        lab, best = O.brute_f64(x, c)
        arrs[f"labels_{n}_{k}_{d}"] = lab.astype(np.uint32)
        arrs[f"best_{n}_{k}_{d}"] = best.astype(np.float64)
    return arrs


def search_i8_fixture():
    arrs = {}
    for n, k, d in I8_SHAPES:
        x = O.lcg_i8(n * d, 1).reshape(n, d)
        ci = O.lcg_i8(k * d, 2).reshape(k, d)
        lab, dist = O.brute_i64(x, ci)
        arrs[f"labels_{n}_{k}_{d}"] = np.asarray(lab, np.uint32)
        arrs[f"dist_{n}_{k}_{d}"] = np.asarray(dist, np.int64)
    return arrs


def main():
    with open(os.path.join(HERE, "rng.json"), "w") as f:
        json.dump(rng_fixture(), f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "recipes.json"), "w") as f:
        json.dump(recipe_fixture(), f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "search_f32.npz"), **search_f32_fixture())
    np.savez_compressed(os.path.join(HERE, "search_i8.npz"), **search_i8_fixture())


if __name__ == "__main__":
    main()
