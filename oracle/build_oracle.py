"""Build the oracle (test infrastructure) and, when the reference tree is
present, the reference's own native CPU engines into ``oracle/_ref``.

The oracle is the checker for tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; the product library never links it.

Reference engines compiled from their own sources where they lie (never
copied): ``src/mlxlearn/_kernels/_neighbors_cpu_native.cpp`` (exact k-NN for
d <= 32) and ``_radius_cpu_native.cpp`` + ``_radius_tree.cpp`` (exact radius
engine), with the flags of the reference's ``setup.py:50-73``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = "/root/reference/src/mlxlearn/_kernels"
ORACLE_SO = os.path.join(HERE, "_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def build_oracle(verbose: bool = False) -> str:
    cmd = ["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
           "-pthread", os.path.join(HERE, "exact_oracle.c"), "-o", ORACLE_SO, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return ORACLE_SO


def build_reference(verbose: bool = False) -> list[str]:
    """Compile the reference's native CPU engines; skipped when the tree is absent."""
    if not os.path.isdir(REF_ROOT):
        return []
    os.makedirs(REF_DIR, exist_ok=True)
    include = sysconfig.get_paths()["include"]
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    flags = ["-O3", "-std=c++17", "-fno-fast-math", "-ffp-contract=off", "-shared", "-fPIC",
             "-pthread", f"-I{include}"]
    jobs = [
        ("_neighbors_cpu_native", ["_neighbors_cpu_native.cpp"]),
        ("_radius_cpu_native", ["_radius_cpu_native.cpp", "_radius_tree.cpp"]),
    ]
    built = []
    for name, sources in jobs:
        out = os.path.join(REF_DIR, name + suffix)
        srcs = [os.path.join(REF_ROOT, s) for s in sources]
        if os.path.exists(out) and all(os.path.getmtime(out) >= os.path.getmtime(s) for s in srcs):
            built.append(out)
            continue
        cmd = ["g++", *flags, *srcs, "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        built.append(out)
    return built


if __name__ == "__main__":
    build_oracle(verbose=True)
    print(build_reference(verbose=True))
    sys.exit(0)
