#!/usr/bin/env python3
"""Benchmark: exact k-NN queries/s on 1,000,000 x 128 (mixture of 256 Gaussians,
1% exact duplicate rows), 10,000 queries, k = 10, on one MI355X.

One step = one full exact query of all 10,000 rows through the C ABI with the
bfloat16 MFMA screen + float64 exact re-rank (the measured `value`). The same
step with the plain float32 MFMA screen is timed beside it, the two index
matrices are compared (must agree 100%), and a query sample is checked against
the exact oracle. Inputs are resident in HBM before timing starts.

Also reported:
  roofline      screen kernel: algorithmic flops 2*Q*N*D per query call over the
                summed device time of the screen launches (HIP events on the
                library's stream), against the dense bf16 MFMA peak; `traffic`
                from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/.
  cpu_baseline  numpy float64 restatement of the reference CPU path
                (CPUIndex._gram) on a bounded query sample, host cores.
  autotune      before timing, each precision's plan is autotuned on the
                queries (mlk_autotune; --no-autotune keeps the heuristic plan);
                the heuristic plan's time is reported beside the tuned one.
  step_ms       median and interquartile range of the timed steps, and the
                first (cold) call, which builds the screen mirrors.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "mlx-learn_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "exact kNN queries/s 1M×128 k=10: BF16+FP64-rerank vs FP32, 100% index match"
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per screen launch from profiles/<name> (written from rocprofv3 --pmc runs)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(X, Q, k, sample, budget_s=20.0):
    """Time the float64 restatement of the reference CPU path on a bounded query sample."""
    import numpy as np

    import oracle
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    port = oracle.GramPort(X)
    prep = time.perf_counter() - t0
    done, elapsed, block = 0, 0.0, 64
    while done < sample and elapsed < budget_s:
        q = Q[done:done + block]
        t0 = time.perf_counter()
        port.query(q, k, block=block)
        elapsed += time.perf_counter() - t0
        done += len(q)
    del port
    return {"value": done / elapsed, "unit": "queries/s", "cores": cores, "kind": "port",
            "sample": f"{done} queries x {X.shape[0]} rows x {X.shape[1]} (numpy float64 BLAS CPUIndex._gram "
                      f"restatement, prep {prep:.1f}s excluded)"}


def copy_calibration(nbytes):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings scale the screen's."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel() * 4, "launches": 3}), flush=True)


def mfma_calibration(blocks=1024, iterations=20000):
    """The library's register-only bf16 MFMA loop (4 waves per SIMD): the
    attainable MFMA rate at the current clock, and the reference point of the
    MFMA busy-cycle counters of a rocprofv3 --pmc run around it."""
    import ctypes
    import torch
    from mlxlearn_amd import _lib
    lib = _lib.load()
    out = ctypes.c_double()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.mlk_mfma_probe(blocks, iterations, ctypes.byref(out), stream), "mlk_mfma_probe")
    print(json.dumps({"mfma_probe_tflops": round(out.value, 1), "blocks": blocks, "iterations": iterations,
                      "frac_of_2500": round(out.value / PEAK_BF16_TFLOPS, 4)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=10_000)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--oracle-sample", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "deep"],
                    help="c2 (default) is the headline; c1, c3, c4, c5 are the other configurations; "
                         "deep is 1,024-feature rows (the deep screen)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="time the heuristic plans instead of autotuned ones (plans imported from "
                         "MLXLEARN_PLAN_FILE are still used)")
    ap.add_argument("--export-plans", default="", metavar="PATH",
                    help="write the process's tuned plans to PATH at the end (for MLXLEARN_PLAN_FILE, so that "
                         "profiled runs repeat them without autotuning under the profiler)")
    ap.add_argument("--mfma-calibration", action="store_true",
                    help="only run the library's register-only MFMA probe (MFMA counter calibration) and exit")
    ap.add_argument("--no-iter2", dest="iter2", action="store_false",
                    help="c5: skip the second Lloyd iteration measured after the timed steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.mfma_calibration:
        mfma_calibration()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} (WORLD_SIZE={world}) is not supported. "
              "The database would shard by contiguous row ranges with an exact (distance, index) merge.",
              file=sys.stderr)
        sys.exit(2)
    import bench_workloads
    if args.workload != "c2":
        line = bench_workloads.run(args)
        print(json.dumps(line), flush=True)
        export_plans(args)
        return

    import numpy as np
    import torch

    import mlxlearn_amd
    import oracle
    from mlxlearn_amd import _lib

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    t0 = time.perf_counter()
    X, Q = oracle.gen_c2(seed=args.seed, n=args.n, m=args.queries)
    log(f"generated {X.shape} + {Q.shape} in {time.perf_counter() - t0:.1f}s")
    Xd = torch.from_numpy(X).cuda()
    Qd = torch.from_numpy(Q).cuda()
    index = mlxlearn_amd.DeviceNeighborIndex(Xd)
    n, d = X.shape
    m, k = Q.shape[0], args.k

    def run(precision):
        dist, idx = index.query(Qd, k, precision=precision)
        return dist, idx, _lib.stats()

    results, tuned = {}, {}
    for precision in ("bfloat16", "float32"):
        first = bench_workloads.cold(lambda: run(precision))
        if args.autotune:
            # plan search on these queries: untimed preparation, like fit
            tuned[precision] = bench_workloads.tune(index, Qd, k, precision=precision, max_trials=24)
        screen_ms = []

        def step():
            out = run(precision)
            screen_ms.append(out[2]["screen_ms"])
            return out

        # the query returns after its own read-back: no device-wide wait per step
        wall, (dist, idx, stats), spread = bench_workloads.timed(step, args.steps, args.warmup, sync_each=False)
        screen_ms = screen_ms[-args.steps:]
        spread["first_call_ms"] = first
        results[precision] = {"ms_per_step": wall * 1e3, "qps": m / wall, "idx": idx.cpu().numpy(),
                              "dist": dist.cpu().numpy(), "screen_ms": float(np.mean(screen_ms)), "stats": stats,
                              "spread": spread}
        log(f"{precision}: {wall * 1e3:.2f} ms/step, {m / wall:,.0f} q/s, screen {np.mean(screen_ms):.2f} ms, "
            f"stats {stats}")

    bf, fp = results["bfloat16"], results["float32"]
    agree = float((bf["idx"] == fp["idx"]).all(axis=1).mean())
    sample = min(args.oracle_sample, m)
    want_d, want_i = oracle.knn(X, Q[:sample], k)
    oracle_match = float((bf["idx"][:sample] == want_i).all(axis=1).mean())
    oracle_match_fp32 = float((fp["idx"][:sample] == want_i).all(axis=1).mean())
    rel = np.abs(bf["dist"][:sample] - want_d) / np.maximum(want_d, 1e-300)
    log(f"index agreement bf16 vs fp32 {agree:.6f}; vs oracle ({sample} queries) bf16 {oracle_match} fp32 "
        f"{oracle_match_fp32}; max rel dist err {rel.max():.3g}")

    flops = 2.0 * m * n * d
    achieved = flops / (bf["screen_ms"] * 1e-3) / 1e12
    traffic = read_traffic("screen_bf16_c2.txt")
    line = {
        "metric": METRIC,
        "value": round(bf["qps"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bf["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded mixture of 256 Gaussians, 1% exact duplicate rows)",
        "config": {"workload": "c2: exact kNN, database 1,000,000x128 float32, 10,000 queries, k=10, "
                               "bf16 MFMA screen + fp64 exact rerank",
                   "n": n, "d": d, "queries": m, "k": k, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4), "traffic": traffic,
                     "kernel": "k_screen<bf16,128>", "launches_per_step": bf["stats"]["rounds"],
                     "screen_ms_per_step": round(bf["screen_ms"], 3)},
        "fp32_path": {"value": round(fp["qps"], 1), "ms_per_step": round(fp["ms_per_step"], 3),
                      "screen_ms_per_step": round(fp["screen_ms"], 3),
                      "screen_tflops": round(flops / (fp["screen_ms"] * 1e-3) / 1e12, 2),
                      "frac_fp32_peak": round(flops / (fp["screen_ms"] * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4)},
        "index_match": {"bf16_vs_fp32": agree, "bf16_vs_oracle": oracle_match, "fp32_vs_oracle": oracle_match_fp32,
                        "oracle_sample": sample, "max_rel_dist_err": float(rel.max())},
        "screen_stats": {key: bf["stats"][key] for key in ("candidates", "reranked", "fallback_queries", "rounds",
                                                           "splits", "capacity")},
        "step_ms": {"bf16": bf["spread"], "fp32": fp["spread"]},
        "autotune": {"bf16": tuned.get("bfloat16"), "fp32": tuned.get("float32")} if tuned else None,
    }
    index.close()
    del Xd, Qd
    torch.cuda.empty_cache()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(X, Q, k, args.cpu_sample)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    export_plans(args)


def export_plans(args):
    if args.export_plans:
        from mlxlearn_amd import _lib
        with open(args.export_plans, "w") as fh:
            fh.write(_lib.export_plans())


if __name__ == "__main__":
    main()
