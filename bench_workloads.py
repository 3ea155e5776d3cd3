"""The other benchmark configurations of bench.py (``--workload c1|c3|c4|c5``).

Each prints one JSON line like the headline (c2): whole-job throughput with
inputs resident in HBM, the screen kernel's roofline fraction (algorithmic
flops over the summed device time of its launches, from HIP events on the
library's stream), an oracle check on a sample, and a CPU baseline timed on a
bounded sample on the host cores.

* c1  NearestNeighbors(k=5) on 10,000 x 64 N(0,1), 1,000 queries, plain fp32 MFMA screen.
* c3  KNeighborsClassifier(k=3, uniform) on 60,000 x 784 integer pixels, 10,000
      queries: fp16 screen (integer data are exact in fp16) + exact rank + device votes.
* c4  radius_neighbors on 2,000,000 x 32 (uniform cube + 20% clumps), 256
      batches of 1..4096 queries in one call (CSR split per batch), float32 request
      (screened in bf16 by the library; results do not depend on it),
      rows sorted by (distance, index).
* c5  KMeans: one exact assignment of 10,000,000 x 96 rows to 65,536 centroids
      (bf16 screen, k = 1) + one float64 centroid update.
* deep  NearestNeighbors(k=10) on 250,000 x 1,024 rows (the c2 mixture generator
      at 1,024 features), 10,000 queries: rows too wide for register-held queries,
      screened by the deep kernel in 256-feature chunks (bf16; fp16 reported beside).
"""
from __future__ import annotations

import ctypes
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "mlx-learn_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK = {"float32": 157.3, "bfloat16": 2500.0, "float16": 2500.0}  # TFLOP/s, MI355X dense


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cold(fn):
    """Milliseconds of a first call (it builds the screen mirrors and workspace)."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return round((time.perf_counter() - t0) * 1e3, 3)


def timed(fn, steps, warmup, sync_each=True):
    """Mean seconds per step over `steps` timed steps (bracketed by device
    synchronisation), the last output, and the spread of the per-step times
    (median and interquartile range).

    sync_each=False: steps whose call returns only once its results are
    complete (a k-NN query returns after reading its status back) are not
    followed by a device-wide synchronisation each; it would only add a host
    round trip per step. The timed region is synchronised on both sides either
    way."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    per = []
    out = None
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        out = fn()
        if sync_each:
            torch.cuda.synchronize()
        per.append(time.perf_counter() - ts)
    total = time.perf_counter() - t0
    q1, med, q3 = np.percentile(np.array(per) * 1e3, [25, 50, 75])
    spread = {"median_ms": round(float(med), 3), "iqr_ms": round(float(q3 - q1), 3), "max_ms": round(max(per) * 1e3, 3),
              "runs": steps}
    return total / steps, out, spread


def tune(index, Q, k=None, **kw):
    """Autotune the plan for these queries (untimed preparation); a summary for the JSON line."""
    plan = index.autotune(Q, k, **kw)
    log(f"autotune {kw}: {plan}")
    return {"tuned_ms": round(plan["tuned_ms"], 3), "default_ms": round(plan["default_ms"], 3),
            "trials": plan["trials"], "plan": {key: plan[key] for key in ("splits", "capacity", "query_tiles",
                                                                          "stages", "rounds")}}


def base_line(args, metric, value, unit, ms, dtype, data, config):
    return {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": data, "config": config}


def roofline(flops, screen_ms, precision, kernel, launches):
    ach = flops / (screen_ms * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": round(ach, 2), "peak": PEAK[precision], "unit": "TFLOP/s",
            "frac": round(ach / PEAK[precision], 4), "traffic": None, "kernel": kernel,
            "launches_per_step": launches, "screen_ms_per_step": round(screen_ms, 3)}


def knn_step(index, Qd, k, precision):
    from mlxlearn_amd import _lib
    def step():
        d, i = index.query(Qd, k, precision=precision)
        return d, i, _lib.stats()
    return step


def knn_graph_step(index, Qd, k, precision):
    """The same call replayed from a captured HIP graph (fixed query and output buffers)."""
    import torch
    from mlxlearn_amd import _lib
    out = (torch.empty((Qd.shape[0], k), dtype=torch.float64, device="cuda"),
           torch.empty((Qd.shape[0], k), dtype=torch.int64, device="cuda"))

    def step():
        d, i = index.query(Qd, k, precision=precision, graph=True, out=out)
        return d, i, _lib.stats()
    return step


def run_c1(args):
    import torch
    import mlxlearn_amd
    import oracle
    X, Q = oracle.gen_c1(seed=args.seed)
    index = mlxlearn_amd.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    Qd = torch.from_numpy(Q).cuda()
    res, spread, tuned = {}, {}, {}
    for prec in ("float32", "bfloat16"):
        first = cold(knn_step(index, Qd, 5, prec))
        if args.autotune:
            tuned[prec] = tune(index, Qd, 5, precision=prec, max_trials=16)
        t, (d, i, st), spread[prec] = timed(knn_step(index, Qd, 5, prec), args.steps, args.warmup, sync_each=False)
        spread[prec]["first_call_ms"] = first
        res[prec] = (t, d.cpu().numpy(), i.cpu().numpy(), st)
        log(f"c1 {prec}: {t * 1e3:.3f} ms/step, stats {st}")
    # the eager call's screen time (device events) for the roofline
    st_eager = res["float32"][3]
    # serving-loop form: the fp32 call captured once in a HIP graph, then replayed
    gstep = knn_graph_step(index, Qd, 5, "float32")
    tg, (dg, ig, _), spread["float32_graph"] = timed(gstep, max(args.steps, 20), max(args.warmup, 2),
                                                     sync_each=False)
    log(f"c1 float32 graph replay: {tg * 1e3:.3f} ms/step")
    wd, wi = oracle.knn(X, Q, 5)
    t, d, i, st = res["float32"]
    match = float((i == wi).all(axis=1).mean())
    match_graph = float((ig.cpu().numpy() == wi).all(axis=1).mean())
    match_bf = float((res["bfloat16"][2] == wi).all(axis=1).mean())
    flops = 2.0 * Q.shape[0] * X.shape[0] * X.shape[1]
    line = base_line(args, "exact kNN queries/s 10k x 64 k=5 (plain fp32 path)", Q.shape[0] / tg, "queries/s",
                     tg * 1e3, "fp32", "synthetic (seeded N(0,1))",
                     {"workload": "c1: NearestNeighbors(k=5), database 10,000x64 float32, 1,000 queries, fp32 MFMA "
                                  "screen + fp64 exact rerank, call replayed from a captured HIP graph",
                      "n": 10000, "d": 64, "queries": 1000, "k": 5})
    line["roofline"] = roofline(flops, st_eager["screen_ms"], "float32", "k_screen<fp32,64>", st_eager["rounds"])
    line["eager_path"] = {"value": round(Q.shape[0] / t, 1), "ms_per_step": round(t * 1e3, 3),
                          "index_match_graph_vs_oracle": match_graph}
    line["step_ms"] = spread
    line["autotune"] = tuned or None
    line["bf16_path"] = {"value": round(Q.shape[0] / res["bfloat16"][0], 1),
                         "ms_per_step": round(res["bfloat16"][0] * 1e3, 3)}
    line["index_match"] = {"fp32_vs_oracle": match, "bf16_vs_oracle": match_bf, "oracle_sample": Q.shape[0],
                           "max_rel_dist_err": float(np.max(np.abs(d - wd) / np.maximum(wd, 1e-300)))}
    if not args.no_cpu:
        port = oracle.GramPort(X)
        t0 = time.perf_counter()
        port.query(Q, 5, block=256)
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(Q.shape[0] / tc, 1), "unit": "queries/s",
                                "cores": len(os.sched_getaffinity(0)), "kind": "port",
                                "sample": "all 1,000 queries (numpy float64 BLAS CPUIndex._gram restatement)"}
    return line


def numpy_votes(idx, codes, n_classes):
    votes = np.zeros((idx.shape[0], n_classes))
    for j in range(idx.shape[1]):
        np.add.at(votes, (np.arange(idx.shape[0]), codes[idx[:, j]]), 1.0)
    return votes.argmax(1)


def run_c3(args):
    import torch
    import mlxlearn_amd
    import oracle
    from mlxlearn_amd import _lib
    X, y, Q = oracle.gen_c3(seed=args.seed)
    k, n_classes = 3, 10
    index = mlxlearn_amd.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    Qd = torch.from_numpy(Q).cuda()
    codes = torch.as_tensor(y.astype(np.int32)).cuda()
    lib = _lib.load()
    m = Q.shape[0]
    pred = torch.empty(m, dtype=torch.int64, device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def step(prec):
        def f():
            _, i = index.query(Qd, k, return_distance=False, precision=prec)
            st = _lib.stats()
            _lib.check(lib.mlk_knn_votes(ptr(i), ctypes.c_void_p(0), m, k, ptr(codes), n_classes,
                                         _lib.WEIGHTS_UNIFORM, ptr(pred), ctypes.c_void_p(0),
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
            return i, st
        return f

    res, spread, tuned = {}, {}, {}
    for prec in ("float16", "float32"):
        first = cold(step(prec))
        if args.autotune:
            tuned[prec] = tune(index, Qd, k, precision=prec, max_trials=16)
        t, (i, st), spread[prec] = timed(step(prec), args.steps, args.warmup)
        spread[prec]["first_call_ms"] = first
        res[prec] = (t, i.cpu().numpy(), pred.cpu().numpy(), st)
        log(f"c3 {prec}: {t * 1e3:.3f} ms/step, stats {st}")
    sample = min(args.oracle_sample, m)
    _, wi = oracle.knn(X, Q[:sample], k)
    want_pred = numpy_votes(wi, y, n_classes)
    t, i, p, st = res["float16"]
    flops = 2.0 * m * X.shape[0] * X.shape[1]
    line = base_line(args, "KNeighborsClassifier.predict queries/s 60k x 784 k=3", m / t, "queries/s", t * 1e3,
                     "fp16", "synthetic (seeded 28x28 integer blob images, 10 labels)",
                     {"workload": "c3: KNeighborsClassifier(k=3, uniform), database 60,000x784 integer pixels, "
                                  "10,000 queries, fp16 MFMA screen + exact rank + device votes",
                      "n": X.shape[0], "d": X.shape[1], "queries": m, "k": k})
    line["roofline"] = roofline(flops, max(st["screen_ms"], 1e-9), "float16", "k_screen<fp16,784>",
                                st["rounds"]) if st["rounds"] else None
    t32, i32, p32, st32 = res["float32"]
    line["fp32_path"] = {"value": round(m / t32, 1), "ms_per_step": round(t32 * 1e3, 3),
                         "screen": {0x00: "fp32", 0x10: "bf16", 0x20: "fp16"}[int(st32["screen_precision"])],
                         "roofline": roofline(flops, max(st32["screen_ms"], 1e-9), "float32", "k_screen<fp32,896>",
                                              st32["rounds"]) if st32["rounds"] else None,
                         "index_match_vs_oracle": float((i32[:sample] == wi).all(axis=1).mean()),
                         "pred_vs_oracle": float((p32[:sample] == want_pred).mean()),
                         "pred_vs_fp16": float((p32 == p).mean())}
    line["step_ms"] = spread
    line["autotune"] = tuned or None
    line["index_match"] = {"fp16_vs_oracle": float((i[:sample] == wi).all(axis=1).mean()),
                           "fp16_vs_fp32": float((i == res["float32"][1]).all(axis=1).mean()),
                           "pred_vs_oracle": float((p[:sample] == want_pred).mean()), "oracle_sample": sample}
    line["screen_stats"] = {key: st[key] for key in ("candidates", "reranked", "fallback_queries", "rounds")}
    if not args.no_cpu:
        port = oracle.GramPort(X)
        nq = min(args.cpu_sample, m)
        t0 = time.perf_counter()
        port.query(Q[:nq], k, block=128)
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(nq / tc, 1), "unit": "queries/s", "cores": len(os.sched_getaffinity(0)),
                                "kind": "port", "sample": f"{nq} queries x 60,000 rows x 784 (numpy float64 BLAS "
                                                         "CPUIndex._gram restatement; votes excluded)"}
    return line


def run_c4(args):
    import torch
    import mlxlearn_amd
    import oracle
    from mlxlearn_amd import _lib
    X, Q, sizes, radius = oracle.gen_c4(seed=args.seed)
    R2 = radius * radius
    index = mlxlearn_amd.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    Qd = torch.from_numpy(Q).cuda()
    offsets = np.concatenate([[0], np.cumsum(sizes)])

    def step_for(prec):
        def step():
            # every batch in one call: the CSR row pointers split it per batch
            indptr, indices, dist = index.radius(Qd, R2, sort_results=True, precision=prec)
            return indptr, indices, dist, _lib.stats()
        return step

    # the float32 screen of the float32 request, timed beside the bf16 screen
    # (identical CSR: the screen precision changes only the candidate pairs)
    first32 = cold(step_for("float32"))
    t32, (ip32, ind32, dist32, st32), spread32 = timed(step_for("float32"), max(2, args.steps // 2), 1)
    spread32["first_call_ms"] = first32
    ip32, ind32 = ip32.cpu().numpy(), ind32.cpu().numpy()
    del dist32
    step = step_for("bfloat16")
    first = cold(step)
    tuned = tune(index, Qd, radius_sq=R2, sort_results=True, precision="bfloat16", max_trials=10) if args.autotune else None
    t, (indptr, indices, dist, st), spread = timed(step, args.steps, args.warmup)
    spread["first_call_ms"] = first
    indptr = indptr.cpu().numpy()
    log(f"c4: {t * 1e3:.3f} ms/step, nnz {indptr[-1]}, stats {st}")
    # oracle check: the first query of a spread of batches
    picks = [int(offsets[b]) for b in range(0, len(sizes), max(1, len(sizes) // 16))]
    ind_h, dist_h = indices.cpu().numpy(), dist.cpu().numpy()
    ok = 0
    for q in picks:
        wp, wi, wd = oracle.radius(X, Q[q:q + 1], R2, sort_results=True)
        got_i = ind_h[indptr[q]:indptr[q + 1]]
        got_d = dist_h[indptr[q]:indptr[q + 1]]
        ok += int(np.array_equal(got_i, wi) and np.allclose(got_d, wd, rtol=1e-12, atol=0))
    m = Q.shape[0]
    flops = 2.0 * m * X.shape[0] * X.shape[1]
    sp = {0x00: ("float32", "fp32"), 0x10: ("bfloat16", "bf16"), 0x20: ("float16", "fp16")}[int(st["screen_precision"])]
    line = base_line(args, "radius_neighbors queries/s 2M x 32, 256 batches, sorted CSR", m / t, "queries/s", t * 1e3,
                     sp[1], "synthetic (seeded uniform cube + 20% Gaussian clumps, calibrated radius)",
                     {"workload": f"c4: radius_neighbors(r={radius}, sort_results=True), database 2,000,000x32 float32, "
                                  f"{len(sizes)} batches ({m} queries) in one call, {sp[1]} MFMA screen of the float32 "
                                  "queries + exact float64 band",
                      "n": X.shape[0], "d": X.shape[1], "queries": m, "batches": len(sizes)})
    line["roofline"] = roofline(flops, st["screen_ms"], sp[0], f"k_screen<{sp[1]},32,pairs>", 1)
    line["fp32_path"] = {"value": round(m / t32, 1), "ms_per_step": round(t32 * 1e3, 3),
                         "screen": {0x00: "fp32", 0x10: "bf16", 0x20: "fp16"}[int(st32["screen_precision"])],
                         "roofline": roofline(flops, st32["screen_ms"], "float32", "k_screen<fp32,32,pairs>", 1),
                         "candidates": st32["candidates"], "step_ms": spread32,
                         "csr_equal_to_bf16": bool(np.array_equal(ip32, indptr) and np.array_equal(ind32, ind_h))}
    line["step_ms"] = spread
    line["autotune"] = tuned
    line["result"] = {"nnz": int(indptr[-1]), "median_per_query": float(np.median(np.diff(indptr))),
                      "oracle_rows_checked": len(picks), "oracle_rows_equal": ok, "candidates": st["candidates"]}
    if not args.no_cpu:
        # the reference's own radius engine (brute_query, compiled from its sources
        # into oracle/_ref) on a bounded sample of the queries, every host core
        nq = min(256, m)
        cores = oracle.default_threads()
        if oracle.reference_radius_engine() is not None:
            t0 = time.perf_counter()
            oracle.reference_radius(X, Q[:nq], R2, sort_results=True, threads=cores)
            tc = time.perf_counter() - t0
            kind, what = "reference", "the reference's brute_query (_radius_cpu_native.cpp:219-310)"
        else:
            t0 = time.perf_counter()
            oracle.radius(X, Q[:nq], R2, sort_results=True)
            tc = time.perf_counter() - t0
            kind, what = "port", "C restatement of the canonical float64 brute force with the exact band"
        line["cpu_baseline"] = {"value": round(nq / tc, 1), "unit": "queries/s", "cores": cores, "kind": kind,
                                "sample": f"{nq} queries x 2,000,000 rows, sorted CSR ({what}, {cores} threads)"}
    return line


def run_c5(args):
    import torch
    import oracle
    from mlxlearn_amd.cluster._k_means import assign_device, update_device
    n = args.n if args.n != 1_000_000 else 10_000_000
    X, C = oracle.gen_c5(seed=args.seed, n=n)
    kc = C.shape[0]
    Xd = torch.from_numpy(X).cuda()
    Cd = torch.from_numpy(C).cuda()
    from mlxlearn_amd import _lib

    def step():
        labels, resid = assign_device(Xd, Cd, precision="bfloat16")
        st = _lib.stats()
        sums, counts, inertia = update_device(Xd, labels, kc, resid)
        return labels, resid, counts, inertia, st

    first = cold(step)
    tuned = None
    if args.autotune:
        # a plan tuned on a centres index serves every assignment of this shape
        import mlxlearn_amd
        centres = mlxlearn_amd.DeviceNeighborIndex(Cd)
        tuned = tune(centres, Xd, 1, precision="bfloat16", max_trials=12)
        centres.close()
    t, (labels, resid, counts, inertia, st), spread = timed(step, args.steps, args.warmup)
    spread["first_call_ms"] = first
    log(f"c5: {t * 1e3:.3f} ms/step, stats {st}")
    sample = min(4096, n)
    rows = np.random.default_rng(1).choice(n, sample, replace=False)
    rows.sort()
    wl, wr = oracle.kmeans_assign(X[rows], C)
    lab = labels.cpu().numpy()
    res = resid.cpu().numpy()
    flops = 2.0 * n * kc * X.shape[1]
    line = base_line(args, "KMeans assignment + update rows/s 10M x 96, 65,536 centroids", n / t, "rows/s", t * 1e3,
                     "bf16", "synthetic (seeded mixture of 4,096 Gaussians, centroids sampled from rows)",
                     {"workload": f"c5: KMeans one exact assignment (bf16 MFMA screen, k=1) + one float64 update, "
                                  f"{n}x96 float32 rows, {kc} centroids", "n": n, "d": X.shape[1], "clusters": kc})
    line["roofline"] = roofline(flops, st["screen_ms"], "bfloat16", "k_screen<bf16,96>", st["rounds"])
    line["step_ms"] = spread
    line["autotune"] = tuned
    line["check"] = {"labels_vs_oracle": float((lab[rows] == wl).mean()), "oracle_sample_rows": sample,
                     "max_rel_residual_err": float(np.max(np.abs(res[rows] - wr) / np.maximum(wr, 1e-300))),
                     "inertia": float(inertia.item()), "inertia_fsum_rel_err":
                         abs(float(inertia.item()) - math.fsum(res.tolist())) / max(math.fsum(res.tolist()), 1e-300)}
    # a Lloyd iteration after the first: the updated centres, assigned with
    # certified reuse of the labels above (lloyd_bounds.reusable_labels) and
    # without it; both give the same labels
    if not args.iter2:
        return finish_c5(args, line, X, C, kc)
    sums, counts, _ = update_device(Xd, labels, kc, resid)
    C1 = torch.where((counts > 0)[:, None], sums / counts.clamp(min=1).to(torch.float64)[:, None],
                     Cd.to(torch.float64)).to(torch.float32).contiguous()
    reused = []
    _, (l_reuse, _), s_reuse = timed(lambda: assign_device(Xd, C1, precision="bfloat16", previous=labels,
                                                           out_reused=reused), 5, 1)
    _, (l_full, _), s_full = timed(lambda: assign_device(Xd, C1, precision="bfloat16"), 5, 1)
    # medians of five calls, with the slowest call beside them
    line["lloyd_iteration_2"] = {"assign_reuse_ms": s_reuse["median_ms"], "assign_reuse_max_ms": s_reuse["max_ms"],
                                 "assign_full_ms": s_full["median_ms"], "assign_full_max_ms": s_full["max_ms"],
                                 "reused_rows": reused[-1], "reused_fraction": round(reused[-1] / n, 4),
                                 "labels_equal": bool(torch.equal(l_reuse, l_full))}
    log(f"c5 iteration 2: {line['lloyd_iteration_2']}")
    return finish_c5(args, line, X, C, kc)


def finish_c5(args, line, X, C, kc):
    import oracle
    if not args.no_cpu:
        # the reference's assignment on the CPU: float32 Gram argmin (BLAS sgemm on
        # every host core) + the float64 check, on a bounded sample of rows
        port = oracle.KMeansPort(C)
        nr = 8192
        t0 = time.perf_counter()
        port.assign(X[:nr])
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(nr / tc, 1), "unit": "rows/s", "cores": len(os.sched_getaffinity(0)),
                                "kind": "port", "sample": f"{nr} rows x {kc} centroids (numpy float32 Gram argmin "
                                                          "+ float64 check, e_means.py:50-69 restated; update "
                                                          "excluded)"}
    return line


def run_deep(args):
    import torch
    import mlxlearn_amd
    import oracle
    n, d, m, k = 250_000, 1024, 10_000, 10
    X, Q = oracle.gen_c2(seed=args.seed, n=n, m=m, d=d)
    index = mlxlearn_amd.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    Qd = torch.from_numpy(Q).cuda()
    res, spread, tuned = {}, {}, {}
    for prec in ("bfloat16", "float16"):
        first = cold(knn_step(index, Qd, k, prec))
        if args.autotune:
            tuned[prec] = tune(index, Qd, k, precision=prec, max_trials=16)
        t, (dd, ii, st), spread[prec] = timed(knn_step(index, Qd, k, prec), args.steps, args.warmup,
                                              sync_each=False)
        spread[prec]["first_call_ms"] = first
        res[prec] = (t, dd.cpu().numpy(), ii.cpu().numpy(), st)
        log(f"deep {prec}: {t * 1e3:.3f} ms/step, stats {st}")
    sample = np.random.default_rng(args.seed).choice(m, args.oracle_sample, replace=False)
    wd, wi = oracle.knn(X, Q[sample], k)
    t, dd, ii, st = res["bfloat16"]
    dp = (d + 255) // 256 * 256  # deep kernels pad rows to 256-feature chunks
    flops = 2.0 * m * n * dp
    line = base_line(args, "exact kNN queries/s 250k x 1024 k=10 (deep screen, bf16 + fp64 rerank)", m / t,
                     "queries/s", t * 1e3, "bf16", "synthetic (seeded Gaussian mixture, 256 centres, 1% duplicates)",
                     {"workload": "deep: NearestNeighbors(k=10), database 250,000x1024 float32, 10,000 queries, "
                                  "bf16 deep MFMA screen (256-feature chunks) + fp64 exact rerank",
                      "n": n, "d": d, "queries": m, "k": k})
    line["roofline"] = roofline(flops, st["screen_ms"], "bfloat16", "k_screen<bf16,256,1,2,2,deep>", st["rounds"])
    line["step_ms"] = spread
    line["autotune"] = tuned or None
    line["fp16_path"] = {"value": round(m / res["float16"][0], 1), "ms_per_step": round(res["float16"][0] * 1e3, 3),
                         "screen_frac": round(flops / (res["float16"][3]["screen_ms"] * 1e-3) / 1e12 / PEAK["float16"], 4)}
    line["screen_stats"] = {"candidates": st["candidates"], "reranked": st["reranked"],
                            "fallback_queries": st["fallback_queries"], "rounds": st["rounds"], "splits": st["splits"]}
    line["index_match"] = {"bf16_vs_oracle": float((ii[sample] == wi).all(axis=1).mean()),
                           "fp16_vs_oracle": float((res["float16"][2][sample] == wi).all(axis=1).mean()),
                           "oracle_sample": int(args.oracle_sample),
                           "max_rel_dist_err": float(np.max(np.abs(dd[sample] - wd) / np.maximum(wd, 1e-300)))}
    if not args.no_cpu:
        port = oracle.GramPort(X)
        qs = Q[:64]
        t0 = time.perf_counter()
        port.query(qs, k, block=64)
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(len(qs) / tc, 1), "unit": "queries/s",
                                "cores": len(os.sched_getaffinity(0)), "kind": "port",
                                "sample": "64 of the 10,000 queries (numpy float64 BLAS CPUIndex._gram restatement)"}
    return line


def run(args):
    return {"c1": run_c1, "c3": run_c3, "c4": run_c4, "c5": run_c5, "deep": run_deep}[args.workload](args)
