#!/usr/bin/env python3
"""Benchmark of the matching hot path on one MI355X.

Default workload (config 4, the batched multi-query configuration): seeded R-MAT graph, scale 20,
edge factor 10, (a, b, c) = (.57, .19, .19), deduplicated, 16 uniform labels, resident in HBM; the
admitted random-walk queries of the config-4 query stream (asdmatch_amd/workload_manifest.json;
admission rule in asdmatch_amd/workloads.py).  One step = every admitted query filtered, planned
and counted exactly.  Timed side by side:
  - plain GPU drop-in: one filter + plan + count call sequence per query (asdm_filter_cfl,
    asdm_plan_create, asdm_count), as the reference's main.cpp runs a query;
  - batched + learned plans (the value): one asdm_batch_count_learned call per step, plans
    ordered by the trained estimator (asdmatch_amd/data/mpnn_estimator.bin).
Every count is checked against the manifest's exact count.

Prints one JSON line: queries/s (value), embeddings/s, the roofline of the dominant kernel (its
algorithmic bytes over its HIP-event-timed launches in the timed region) and the reference CPU
path timed on this host on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--workload c4|c2]

--workload c2 runs the config-2 queries (Chung-Lu 1M / 15M) one by one through the plain path.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ESTIMATOR = os.path.join(ROOT, "asdmatch_amd", "data", "mpnn_estimator.bin")
REF_LOCAL = os.path.join(ROOT, "oracle", "_ref", "ref_driver_local")
TRAFFIC_FILES = {"c2": os.path.join(ROOT, "profiles", "c2_traffic.txt"),
                 "c4": os.path.join(ROOT, "profiles", "c4_traffic.txt")}


def start_cpu_baseline(g, q, tmpdir):
    """Reference ALL_LEVEL build (its exact path), one thread, on one admitted config-2 query."""
    if not os.path.exists(REF_LOCAL):
        return None
    dpath = os.path.join(tmpdir, "c2.graph")
    qpath = os.path.join(tmpdir, "c2_query.graph")
    g.write(dpath)
    q.write(qpath)
    return subprocess.Popen([REF_LOCAL, "-q", qpath, "-d", dpath], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                            text=True)


def finish_cpu_baseline(proc, expect, timeout):
    if proc is None:
        return None
    try:
        out, _ = proc.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        proc.kill()
        proc.communicate()
        return {"error": f"reference did not finish within {timeout} s"}
    vals = {}
    for key, pat in (("filter", r"Filtering time \(s\): ([0-9.eE+-]+)"), ("plan", r"Planning Time: ([0-9.eE+-]+)"),
                     ("exec", r"Execution Time: ([0-9.eE+-]+)"), ("count", r"Number of matches: (\d+)")):
        m = re.search(pat, out)
        if m:
            vals[key] = float(m.group(1)) if key != "count" else int(m.group(1))
    if "exec" not in vals:
        return {"error": "reference output not parsed", "tail": out[-300:]}
    secs = vals.get("filter", 0) + vals.get("plan", 0) + vals["exec"]
    return {"value": 1.0 / secs, "unit": "queries/s", "cores": 1, "kind": "reference",
            "seconds": secs, "count": vals.get("count"), "count_matches_gpu": vals.get("count") == expect,
            "embeddings_per_s": (vals.get("count") or 0) / secs}


def read_traffic(kernel, workload="c2"):
    """Calibrated HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary."""
    path = TRAFFIC_FILES[workload]
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line:
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    if vals.get("kernel") != kernel or "traffic_bytes_per_launch" not in vals:
        return None
    return float(vals["traffic_bytes_per_launch"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("c2", "c4"), default="c4",
                    help="c4: config-4 batch, batched + learned plans vs plain (default); c2: config-2 queries "
                         "one by one through the plain path")
    ap.add_argument("--batch-shape", default=None,
                    help="LEAF_LANES,CLASS_SCALE,TILE_ITERS: use this batch launch shape instead of autotuning")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many MiB (PMC calibration) and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the matching path runs on a single MI355X here "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch

        n = args.copy_probe << 20
        x = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "launches": 5}))
        return

    import asdmatch_amd as A
    from asdmatch_amd import workloads as W

    if args.workload == "c4":
        return bench_batch(args)

    t0 = time.time()
    g = W.graph("c2")
    adm = W.admitted("c2")
    queries = [(W.query(g, s, seed), seed, cnt) for s, seed, cnt in adm]
    gen_s = time.time() - t0

    tmpdir = tempfile.mkdtemp(prefix="asdm_bench_")
    proc = None
    if not args.no_cpu_baseline and queries:
        sample = min(queries, key=lambda t: t[2])  # the smallest admitted query
        proc = start_cpu_baseline(g, sample[0], tmpdir)

    ctx = A.Context(0)
    ctx.upload(g)

    def step():
        total = 0
        for q, seed, expect in queries:
            cs = ctx.filter(q, g)
            plan = cs.plan(q)
            c = ctx.count(plan, cs)
            if c != expect:
                raise SystemExit(f"count mismatch for query seed {seed}: {c} != {expect}")
            total += c
        return total

    for _ in range(args.warmup):
        step()
    ctx.set_profiling(True)  # HIP events around every launch, resolved after the timed region
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    t1 = time.perf_counter()
    emb = 0
    for _ in range(args.steps):
        emb += step()
    hip.hipDeviceSynchronize()
    t2 = time.perf_counter()
    kstats = ctx.kernel_stats()
    ctx.set_profiling(False)

    elapsed = t2 - t1
    nq = len(queries) * args.steps
    roofline = roofline_of(kstats, "c2")

    cpu = finish_cpu_baseline(proc, min(queries, key=lambda t: t[2])[2] if queries else None, timeout=600)
    if cpu and "value" in cpu:
        cpu["sample"] = ("reference ALL_LEVEL build (exact enumeration path), single thread, on the smallest "
                         "admitted config-2 query; time = its Filtering + Planning + Execution lines "
                         "(graph load excluded, as the GPU graph is resident)")

    out = {
        "metric": "embeddings/s + queries/s, batched+learned-plan vs plain GPU, exact uint64 counts",
        "value": round(nq / elapsed, 4),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 ids / u64 counts",
        "data": "synthetic",
        "config": {"workload": "config 2: Chung-Lu 1M V / 15M E, beta 2.5, 32 labels Zipf(1.2); admitted "
                               "random-walk queries, plain GPU engine (filter + plan + exact count)",
                   "queries_per_step": len(queries), "query_sizes": [q.n for q, _, _ in queries],
                   "admission_row_limit": W.load_manifest()["c2"]["row_limit"]},
        "embeddings_per_s": round(emb / elapsed, 1),
        "embeddings_per_step": emb // max(1, args.steps),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in ks.items()} for ks in kstats],
        "setup_s": round(gen_s, 2),
    }
    print(json.dumps(out), flush=True)


def roofline_of(kstats, workload):
    """Roofline object of the dominant kernel (most device time) of a kernel table."""
    ks = [k for k in kstats if k["launches"] and k["ms"] > 0]
    if not ks:
        return None
    dom = max(ks, key=lambda k: k["ms"])
    bytes_per_launch = dom["bytes"] / dom["launches"]
    avg_s = dom["ms"] / 1e3 / dom["launches"]
    achieved = bytes_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": read_traffic(dom["name"], workload),
            "kernel": dom["name"], "alg_bytes_per_launch": round(bytes_per_launch),
            "avg_launch_ms": round(avg_s * 1e3, 4), "launches": dom["launches"]}


def start_cpu_batch_baseline(g, sample, tmpdir):
    """Reference ALL_LEVEL build (its exact path), one single-threaded process per sampled query,
    the processes run side by side on the host cores."""
    if not os.path.exists(REF_LOCAL) or not sample:
        return None
    dpath = os.path.join(tmpdir, "c4.graph")
    g.write(dpath)
    procs = []
    for k, (q, _) in enumerate(sample):
        qpath = os.path.join(tmpdir, f"c4_query_{k}.graph")
        q.write(qpath)
        procs.append(subprocess.Popen([REF_LOCAL, "-q", qpath, "-d", dpath], stdout=subprocess.PIPE,
                                      stderr=subprocess.STDOUT, text=True))
    return procs


def finish_cpu_batch_baseline(procs, sample, total_embeddings, nq, timeout):
    if not procs:
        return None
    secs, ok, emb = 0.0, True, 0
    for p, (_, expect) in zip(procs, sample):
        r = finish_cpu_baseline(p, expect, timeout)
        if not r or "value" not in r:
            return r
        secs += r["seconds"]
        ok &= bool(r["count_matches_gpu"])
        emb += expect
    rate = emb / secs  # embeddings per second of the reference's exact path
    return {"value": round(len(sample) / secs, 6), "unit": "queries/s", "cores": 1, "kind": "reference",
            "seconds": round(secs, 3), "queries": len(sample), "counts_match_gpu": ok,
            "embeddings_per_s": round(rate, 1),
            "full_batch_queries_per_s_extrapolated": round(nq / (total_embeddings / rate), 9)}


def bench_batch(args):
    """Config-4 batch: the admitted queries counted (a) one by one through the plain path and
    (b) in one asdm_batch_count_learned call per step with learned plans (the value)."""
    import asdmatch_amd as A
    from asdmatch_amd import workloads as W

    t0 = time.time()
    g = W.graph("c4")
    adm = W.admitted("c4")
    qs = [W.query(g, s, seed) for s, seed, _ in adm]
    want = [c for _, _, c in adm]
    gen_s = time.time() - t0

    # CPU baseline sample: the four admitted queries with the fewest embeddings (the reference
    # enumerates every embedding; the whole batch would take it days), run while the GPU works
    tmpdir = tempfile.mkdtemp(prefix="asdm_bench_")
    order = sorted(range(len(qs)), key=lambda i: want[i])
    sample = [(qs[i], want[i]) for i in order[:4]]
    procs = None if args.no_cpu_baseline else start_cpu_batch_baseline(g, sample, tmpdir)

    ctx = A.Context(0)
    ctx.upload(g)
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    m = None
    if os.path.exists(ESTIMATOR):
        m = A.Mpnn(ctx)
        m.load(ESTIMATOR)

    def plain():
        out = []
        for q in qs:
            cs = ctx.filter(q, g)
            out.append(0 if cs is None else ctx.count(cs.plan(q), cs))
        return out

    def batched(mm):
        c, st, lat = ctx.batch_count_learned(g, qs, mm)
        if [int(x) for x in c] != want or any(st):
            raise SystemExit("batched count mismatch")
        return lat

    for _ in range(args.warmup):
        if plain() != want:
            raise SystemExit("plain count mismatch")
        batched(None)
        batched(m)
    # launch shape of the batched engine, picked by the library's autotuner (untimed) on every fourth
    # query of the batch, which keeps the tuner's 36 shapes to seconds
    if args.batch_shape:
        lanes, scale, iters = args.batch_shape.split(",")
        ctx.set_batch_tuning(int(lanes), float(scale), int(iters))
        tune = ctx.batch_tuning() + (None,)
    elif args.warmup:
        tune = ctx.batch_autotune(g, qs[::4], m, reps=1)
    else:
        tune = ctx.batch_tuning() + (None,)
    hip.hipDeviceSynchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        got = plain()
    hip.hipDeviceSynchronize()
    t2 = time.perf_counter()
    if got != want:
        raise SystemExit("plain count mismatch")
    for _ in range(args.steps):
        batched(None)
    hip.hipDeviceSynchronize()
    t3 = time.perf_counter()
    # the value: batched + learned plans, HIP events around every launch (resolved afterwards)
    ctx.set_profiling(True)
    lats = []
    hip.hipDeviceSynchronize()
    t4 = time.perf_counter()
    for _ in range(args.steps):
        lats.extend(batched(m).tolist())
    hip.hipDeviceSynchronize()
    t5 = time.perf_counter()
    kstats = ctx.kernel_stats()
    ctx.set_profiling(False)

    nq = len(qs) * args.steps
    plain_qps, batch_qps, learned_qps = nq / (t2 - t1), nq / (t3 - t2), nq / (t5 - t4)
    lats.sort()
    cpu = finish_cpu_batch_baseline(procs, sample, sum(want), len(qs), timeout=600)
    if cpu and "value" in cpu:
        cpu["sample"] = ("reference ALL_LEVEL build (exact enumeration path), one thread per query, on the four "
                         "admitted config-4 queries with the fewest embeddings; time = the sum of their "
                         "Filtering + Planning + Execution lines (graph load excluded, as the GPU graph is "
                         "resident); the extrapolation divides the batch's embeddings by the sample's rate")
    out = {
        "metric": "embeddings/s + queries/s, batched+learned-plan vs plain GPU, exact uint64 counts",
        "value": round(learned_qps, 3), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round((t5 - t4) / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32 ids / u64 counts", "data": "synthetic",
        "config": {"workload": "config 4: R-MAT scale 20, edge factor 10, (a,b,c)=(.57,.19,.19), 16 uniform "
                               "labels; admitted random-walk queries, one asdm_batch_count_learned call per step",
                   "queries_per_step": len(qs), "query_sizes": sorted({q.n for q in qs}),
                   "admission_row_limit": W.load_manifest()["c4"]["row_limit"],
                   "learned_plans": m is not None,
                   "batch_launch_shape": {"leaf_lanes": tune[0], "class_scale": tune[1], "tile_iters": tune[2],
                                          "autotuned_on": "every fourth query" if tune[3] is not None else None}},
        "plain_gpu_queries_per_s": round(plain_qps, 3),
        "batched_plain_plan_queries_per_s": round(batch_qps, 3),
        "batched_learned_plan_queries_per_s": round(learned_qps, 3),
        "batched_learned_vs_plain": round(learned_qps / plain_qps, 3),
        "embeddings_per_s": round(sum(want) * args.steps / (t5 - t4), 1),
        "latency_ms_p50": round(lats[len(lats) // 2], 3), "latency_ms_p99": round(lats[int(len(lats) * 0.99)], 3),
        "roofline": roofline_of(kstats, "c4"),
        "cpu_baseline": cpu,
        "kernels": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in ks.items()} for ks in kstats],
        "setup_s": round(gen_s, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
