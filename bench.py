"""Benchmark of the nearest-centroid search (assign-only search at scale).

Workload: 2,000,000 x 128 f32 rows from a Gaussian mixture (16384 blob
centres ~ N(0, 16 I), each row = a uniformly chosen centre + N(0, I), seed 0),
k = 16384 centroids drawn from the rows; every row resident in HBM before the
timed region. One step = one pass of the search over all n rows: for the
BF16/FP8 searches that is the MFMA search kernel plus the FP32 rescoring of
the top-m candidates (plus the per-row FP8 encoding), for FP32 the fused
argmin kernel. Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision bf16|fp8|fp32|i8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "e-means_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

# MI355X dense peaks (AMD published): TFLOP/s (TOP/s for int8)
PEAK = {"fp32": 157.3, "bf16": 2516.6, "fp8": 5033.2, "i8": 5033.2}
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp8", "fp32", "i8"])
    ap.add_argument("--rescore-m", type=int, default=4)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--d", type=int, default=128)
    ap.add_argument("--k", type=int, default=16384)
    ap.add_argument("--blobs", type=int, default=16384)
    ap.add_argument("--judge-rows", type=int, default=20000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP32 / FP8 comparison passes")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (HBM counter calibration)")
    return ap.parse_args()


def copy_probe(nbytes: int):
    """A plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for it
    calibrates the counters of the search kernels."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}), flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_data(n, d, k, blobs, device):
    import torch

    g = torch.Generator(device=device).manual_seed(0)
    centres = torch.randn(blobs, d, generator=g, device=device) * 4.0
    which = torch.randint(0, blobs, (n,), generator=g, device=device)
    x = centres[which] + torch.randn(n, d, generator=g, device=device)
    pick = torch.randperm(n, generator=torch.Generator().manual_seed(1234))[:k]
    c = x[pick.to(device)].contiguous()
    return x.contiguous(), c


def run_search(prec, m, x, c_host, steps, warmup, rows_i8=False):
    """Time `steps` searches of all rows; per-kernel times from HIP events."""
    import torch

    import emeans_hip as E

    a = E.HipAssigner(prec, rescore_m=m)
    a.set_stream(torch.cuda.current_stream().cuda_stream)
    a.set_centroids(c_host)
    n = x.shape[0]
    labels = torch.empty(n, dtype=torch.int32, device=x.device)
    dist = torch.empty(n, dtype=torch.float32, device=x.device)
    for _ in range(warmup):
        a.assign_device(x, labels, dist)
    torch.cuda.synchronize()
    a.set_timing(True)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(steps):
        a.assign_device(x, labels, dist)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    enc, search, rescore, calls = a.get_timing()
    a.set_timing(False)
    name = a.name()
    return {
        "ms_per_step": ms,
        "rows_per_s": n / (ms * 1e-3),
        "search_ms": search / max(calls, 1),
        "rescore_ms": rescore / max(calls, 1),
        "encode_ms": enc / max(calls, 1),
        "name": name,
        "labels": labels,
        "dist": dist,
    }


def judge(x, c, res, rows):
    """FP64 brute force (torch, on the GPU) on a fixed subsample of rows."""
    import torch

    g = torch.Generator().manual_seed(7)
    ids = torch.randperm(x.shape[0], generator=g)[:rows].to(x.device)
    xs = x[ids].double()
    cd = c.double()
    d2 = (xs * xs).sum(1, keepdim=True) - 2.0 * (xs @ cd.T) + (cd * cd).sum(1)[None]
    best = d2.min(1)
    lab = res["labels"][ids].long()
    d64 = ((xs - cd[lab]) ** 2).sum(1)
    best_exact = ((xs - cd[best.indices]) ** 2).sum(1)
    wrong = d64 > best_exact + 1e-6 * (1 + best_exact)
    # error relative to 1 + d, the form of the reference's tolerance (src/cpu.rs:303)
    rel = (res["dist"][ids].double() - d64).abs() / (1.0 + d64)
    return {
        "label_disagreement": float((lab != best.indices).double().mean()),
        "true_label_errors": float(wrong.double().mean()),
        "dist_err_max_rel_1pd": float(rel.max()),
        "judged_rows": rows,
    }


def cpu_baseline(x, c, seconds):
    """The oracle's restatement of the reference CPU search (src/cpu.rs,
    AVX-512 4x4 micro-kernel, 256 KiB blocks) on the host cores, on a bounded
    sample of the same rows."""
    import numpy as np

    import oracle as O

    cores = O.default_threads()
    simd = O.lib().or_simd_best()
    xh = x.cpu().numpy()
    ch = c.cpu().numpy()
    sample = 512
    t = time.perf_counter()
    O.cpu_assign(xh[:sample], ch, nthreads=cores)
    dt = time.perf_counter() - t
    sample = int(min(xh.shape[0], max(sample, sample * seconds / max(dt, 1e-3))))
    t = time.perf_counter()
    O.cpu_assign(xh[:sample], ch, nthreads=cores)
    dt = time.perf_counter() - t
    return {
        "value": sample / dt,
        "unit": "rows/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {sample} rows of the same workload against all {ch.shape[0]} centroids "
                  f"({'AVX-512' if simd == O.SIMD_AVX512 else 'portable'} kernel, {cores} threads, {dt:.1f} s)",
        "gflops": 2.0 * sample * ch.shape[0] * ch.shape[1] / dt / 1e9,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (rows would shard by "
              "range with no exchange; see the sharding note in DESIGN.md). Run with --gpus 1.")
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU")
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    dev = torch.device("cuda:0")
    log(f"data: n={args.n} d={args.d} k={args.k}")
    x, c = make_data(args.n, args.d, args.k, args.blobs, dev)
    c_host = c.cpu().numpy()
    torch.cuda.synchronize()

    variants = {}
    prec = args.precision
    m = args.rescore_m if prec in ("bf16", "fp8") else 1
    xq = x
    if prec == "i8":
        xq = x.clamp(-128, 127).round().to(torch.int8)
    log(f"timing {prec} (top-{m})")
    head = run_search(prec, m, xq, c_host, args.steps, args.warmup)
    variants[prec] = head
    if not args.no_compare and prec != "fp32":
        log("timing fp32")
        variants["fp32"] = run_search("fp32", 1, x, c_host, max(2, args.steps // 4), 1)
    if not args.no_compare and prec == "bf16":
        log("timing fp8 (top-8)")
        variants["fp8"] = run_search("fp8", 8, x, c_host, args.steps, args.warmup)

    quality = {}
    if prec != "i8":
        for name, v in variants.items():
            q = judge(x, c, v, min(args.judge_rows, args.n))
            q["mse"] = float(v["dist"].double().mean())
            quality[name] = q
        if "fp32" in quality:
            for name, q in quality.items():
                q["mse_rel_vs_fp32"] = q["mse"] / quality["fp32"]["mse"] - 1.0

    flops = 2.0 * args.n * args.k * args.d
    achieved = flops / (head["search_ms"] * 1e-3) / 1e12
    peak = PEAK[prec]
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved, 2),
        "peak": peak,
        "unit": "TFLOP/s" if prec != "i8" else "TOP/s",
        "frac": round(achieved / peak, 4),
        "traffic": None,
        "kernel": f"search_kernel ({prec})",
        "kernel_ms": round(head["search_ms"], 4),
    }
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(x, c, args.cpu_seconds)

    out = {
        "metric": "assigned rows/s & % roofline, BF16/FP8+rescore vs FP32, at matched MSE",
        "value": round(head["rows_per_s"], 1),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16": "bf16", "fp8": "fp8", "fp32": "f32", "i8": "int8"}[prec],
        "data": "synthetic (Gaussian mixture, generated on the GPU with torch, seed 0)",
        "config": {
            "workload": f"assign-only search at scale: {args.n}x{args.d} f32 rows, mixture of {args.blobs} blobs, "
                        f"k={args.k} centroids drawn from the rows",
            "n": args.n, "d": args.d, "k": args.k,
            "precision": prec + ("+fp32 rescore" if prec in ("bf16", "fp8") else ""),
            "rescore_m": m,
            "rows_resident_in_hbm": True,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "variants": {
            k_: {
                "rows_per_s": round(v["rows_per_s"], 1),
                "ms_per_step": round(v["ms_per_step"], 4),
                "search_ms": round(v["search_ms"], 4),
                "rescore_ms": round(v["rescore_ms"], 4),
                "encode_ms": round(v["encode_ms"], 4),
                "search_tflops": round(flops / (v["search_ms"] * 1e-3) / 1e12, 2),
                "name": v["name"],
                **quality.get(k_, {}),
            }
            for k_, v in variants.items()
        },
    }
    if "fp32" in variants and prec != "fp32":
        out["speedup_vs_fp32"] = round(head["rows_per_s"] / variants["fp32"]["rows_per_s"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
