"""Compile the gfx950 HIP library ``mlxlearn_amd/_lib/libmlxlearn_amd.so``.

Every translation unit is compiled with hipcc for gfx950 only, in parallel,
then linked into one shared library exporting the C ABI of
``include/mlxlearn_amd.h``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "mlxlearn_amd", "_lib")
OUT = os.path.join(OUT_DIR, "libmlxlearn_amd.so")
OBJ_DIR = os.path.join(HERE, "_build")
SOURCES = ["mlk_capi.hip", "mlk_index.hip", "mlk_screen.hip", "mlk_knn.hip", "mlk_radius.hip", "mlk_kmeans.hip",
           "mlk_tune.hip", "mlk_diag.hip", "mlk_noalloc.hip", "mlk_order.hip"]
# per-source extra flags: the screen's values are finite by construction, so
# its max-tree needs no NaN canonicalisation
EXTRA = {"mlk_screen.hip": ["-fno-honor-nans"]}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed float32 VALU forms are disabled for every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    # MFMA accumulators in VGPRs: the epilogue reads them every tile
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(line for line in proc.stderr.splitlines() if noise not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    if err.strip() and verbose:
        print(err, file=sys.stderr)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
