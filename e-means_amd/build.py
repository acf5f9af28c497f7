"""Builds libemeans_hip.so (HIP for gfx950) in place, next to this file.

    python e-means_amd/build.py [--jobs N]

Each .hip source compiles to an object under e-means_amd/_build/, then one
shared library is linked. No CUDA, no hipify: the sources are HIP for CDNA4.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libemeans_hip.so")
# the heaviest translation units (most template instantiations) first
SOURCES = ["emh_search_reg_bf16.hip", "emh_search_reg_fp8.hip", "emh_search_reg_int.hip", "emh_search_reg_f32.hip",
           "emh_search_lds.hip", "emh_search_jobs.hip", "emh_search.hip", "emh_accum.hip", "emh_api.hip",
           "emh_train.hip", "emh_batch.hip", "emh_quant.hip", "emh_report.hip"]
ARCH = os.environ.get("EMH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",          # the host recipe is f32 without contraction, as Rust
    "-fhip-fp32-correctly-rounded-divide-sqrt",  # f32 / and sqrtf round to nearest, as IEEE
    "-fvisibility=hidden",
    "-Wno-unused-result",
    "-Wno-unused-value",
    # packed FP32 VALU ops stay off (hipcc's host pass ignores the feature)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in arch VGPRs: the search epilogues read every score
    # with VALU ops, which cannot take AGPR operands
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(jobs: int = min(8, os.cpu_count() or 4)) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 4))
    print(build(ap.parse_args().jobs))
