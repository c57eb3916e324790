#!/usr/bin/env python3
"""Benchmark of the matching hot path on one MI355X.

Default workload (config 4, the batched multi-query configuration): seeded R-MAT graph, scale 20,
edge factor 10, (a, b, c) = (.57, .19, .19), deduplicated, 16 uniform labels, resident in HBM; every
admitted random-walk query of the 1,024-slot config-4 query stream (asdmatch_amd/workload_manifest.json,
admission rule in asdmatch_amd/workloads.py; the JSON line states the admission per size class).
One step = every admitted query filtered, planned and counted exactly.  Timed side by side:
  - plain GPU drop-in: one filter + plan + count call sequence per query (asdm_filter_cfl,
    asdm_plan_create, asdm_count), as the reference's main.cpp runs a query;
  - batched + learned plans (the value): one asdm_batch_count_learned call per step, plans
    ordered by the trained estimator (asdmatch_amd/data/mpnn_estimator.bin); queries the planner
    sends to the bag-table engine are counted by it inside the call.
Every count is checked against the manifest's exact count.

--workload c3: config 3's admitted 12-vertex query on the Chung-Lu 200k / 4M graph, full local-count
matrix (LOCAL_COUNT) per step.  --workload c2: the admitted config-2 queries one by one.

Prints one JSON line: queries/s (value), the roofline of the dominant kernel (its algorithmic bytes
over its HIP-event-timed launches in the timed region) and the reference CPU path timed on this
host's cores on a bounded sample of the same workload.  Progress goes to stderr.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--workload c4|c3|c2]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ESTIMATOR = os.path.join(ROOT, "asdmatch_amd", "data", "mpnn_estimator.bin")
REF_BIN = os.path.join(ROOT, "oracle", "_ref")
TRAFFIC_FILES = {w: os.path.join(ROOT, "profiles", f"{w}_traffic.txt") for w in ("c2", "c3", "c4", "c4_isolated")}
METRIC = "embeddings/s + queries/s, batched+learned-plan vs plain GPU, exact uint64 counts"


def say(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ reference CPU path
def parse_ref(out):
    vals = {}
    for key, pat in (("filter", r"Filtering time \(s\): ([0-9.eE+-]+)"), ("plan", r"Planning Time: ([0-9.eE+-]+)"),
                     ("exec", r"Execution Time: ([0-9.eE+-]+)"), ("count", r"Number of matches: (\d+)")):
        m = re.search(pat, out)
        if m:
            vals[key] = float(m.group(1)) if key != "count" else int(m.group(1))
    return vals


def run_reference_pool(binary, jobs, dpath, workers, timeout):
    """Runs the reference once per (query path, expected count) job, `workers` processes at a time,
    as `xargs -P workers` would.  Per job: the reference's own Filtering + Planning + Execution
    seconds (graph load excluded) and its count.  Returns (results, wall-clock makespan)."""
    exe = os.path.join(REF_BIN, binary)

    def one(job):
        qpath, expect = job
        t0 = time.perf_counter()
        try:
            r = subprocess.run([exe, "-q", qpath, "-d", dpath], capture_output=True, text=True, timeout=timeout)
            v = parse_ref(r.stdout)
        except subprocess.TimeoutExpired:
            v = {}
        wall = time.perf_counter() - t0
        secs = v.get("filter", 0) + v.get("plan", 0) + v["exec"] if "exec" in v else None
        return {"seconds": secs, "wall": wall, "count": v.get("count"), "expect": expect}

    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        res = list(ex.map(one, jobs))
    return res, time.perf_counter() - t0


def compute_makespan(res, workers):
    """Makespan of the jobs' compute seconds (graph load excluded) on `workers` cores, scheduled in
    the pool's order (each job to the first free worker)."""
    free = [0.0] * workers
    for r in res:
        k = min(range(workers), key=lambda i: free[i])
        free[k] += r["seconds"] or 0.0
    return max(free)


def cpu_baseline_sample(g, sample, tmpdir, workers, timeout, label):
    """Both reference builds on the same sample: the ALL_LEVEL + LOCAL_COUNT build (exact
    enumeration; oracle/_ref/ref_driver_local) and the default count build (factorized counting;
    oracle/_ref/ref_driver, whose multi-bag counts are known to differ).  Throughput = sample size
    over the compute makespan on `workers` cores."""
    if not os.path.exists(os.path.join(REF_BIN, "ref_driver_local")) or not sample:
        return None
    dpath = os.path.join(tmpdir, "data.graph")
    g.write(dpath)
    jobs = []
    for k, (q, expect) in enumerate(sample):
        qpath = os.path.join(tmpdir, f"query_{k}.graph")
        q.write(qpath)
        jobs.append((qpath, expect))
    say(f"cpu baseline: {len(jobs)} queries, {workers} processes, ALL_LEVEL build")
    exact, wall_e = run_reference_pool("ref_driver_local", jobs, dpath, workers, timeout)
    say(f"cpu baseline: count build")
    cbuild, wall_c = run_reference_pool("ref_driver", jobs, dpath, workers, timeout)
    done = [r for r in exact if r["seconds"] is not None]
    if len(done) < len(exact):
        return {"error": f"{len(exact) - len(done)} of {len(exact)} reference runs did not finish in {timeout} s"}
    ms = compute_makespan(exact, workers)
    msc = compute_makespan(cbuild, workers) if all(r["seconds"] is not None for r in cbuild) else None
    return {
        "value": round(len(exact) / ms, 4), "unit": "queries/s", "cores": workers, "kind": "reference",
        "sample": label + f"; reference ALL_LEVEL build (exact enumeration), one single-threaded process per query, "
                          f"{workers} at a time; throughput = queries / makespan of their Filtering + Planning + "
                          "Execution seconds (graph load excluded, as the GPU graph is resident)",
        "queries": len(exact), "makespan_s": round(ms, 3), "wall_makespan_s_incl_graph_load": round(wall_e, 2),
        "summed_query_s": round(sum(r["seconds"] for r in exact), 3),
        "counts_match_gpu": all(r["count"] == r["expect"] for r in exact),
        "count_build": {"queries_per_s": round(len(cbuild) / msc, 4) if msc else None,
                        "makespan_s": round(msc, 3) if msc else None,
                        "counts_equal_exact": sum(1 for r in cbuild if r["count"] == r["expect"]),
                        "counts_differ": sum(1 for r in cbuild if r["count"] is not None and r["count"] != r["expect"]),
                        "unfinished": sum(1 for r in cbuild if r["count"] is None)},
    }


# ------------------------------------------------------------------ roofline
def read_traffic(kernel, workload, alg_bytes_per_launch):
    """Calibrated HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, when
    that summary was measured on this workload, for this kernel, on launches of the same average
    algorithmic size (within 5 %).  Otherwise (None, the reason): a figure measured on other
    launches says nothing about these."""
    path = TRAFFIC_FILES.get(workload)
    if not path or not os.path.exists(path):
        return None, "no PMC summary for this workload"
    blocks = []  # one block of "key: value" lines per kernel, each opened by its "kernel:" line
    with open(path) as f:
        for line in f:
            if ":" in line:
                k, v = line.split(":", 1)
                if k.strip() == "kernel":
                    blocks.append({})
                if blocks:
                    blocks[-1][k.strip()] = v.strip()
    vals = next((b for b in blocks if b.get("kernel") == kernel), None)
    if vals is None or "traffic_bytes_per_launch" not in vals:
        return None, f"the PMC summary covers {', '.join(b.get('kernel', '?') for b in blocks)}, not {kernel}"
    ref = float(vals.get("alg_bytes_per_launch", 0))
    if not ref or abs(alg_bytes_per_launch - ref) > 0.05 * ref:
        return None, (f"the PMC summary was measured on launches of {ref:.4g} algorithmic bytes, these average "
                      f"{alg_bytes_per_launch:.4g}")
    return float(vals["traffic_bytes_per_launch"]), vals.get("source")


def roofline_of(kstats, workload):
    """Roofline object of the dominant kernel (most device time) of a kernel table."""
    ks = [k for k in kstats if k["launches"] and k["ms"] > 0 and k["bytes"] > 0]
    if not ks:
        return None
    dom = max(ks, key=lambda k: k["ms"])
    bytes_per_launch = dom["bytes"] / dom["launches"]
    avg_s = dom["ms"] / 1e3 / dom["launches"]
    achieved = bytes_per_launch / avg_s / 1e9
    traffic, tnote = read_traffic(dom["name"], workload, bytes_per_launch)
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": tnote,
            "kernel": dom["name"], "alg_bytes_per_launch": round(bytes_per_launch),
            "avg_launch_ms": round(avg_s * 1e3, 4), "launches": dom["launches"]}


def device_aggregate(kstats, steps, wall_s):
    """Algorithmic bytes of every timed kernel per step over the step's wall time: the whole device's
    rate while several streams share it (the table's host entries and derived counters excluded)."""
    b = sum(k["bytes"] for k in kstats
            if k["ms"] > 0 and not k["name"].startswith("host") and not k["name"].endswith("_touched"))
    gbs = b / steps / wall_s / 1e9
    return {"alg_bytes_per_step": round(b / steps), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 5)}


def kernel_table(kstats):
    return [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in ks.items()} for ks in kstats]


def hip_sync():
    import ctypes

    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def base_line(args, value, ms_per_step, config):
    return {"metric": METRIC, "value": value, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32 ids / u64 counts", "data": "synthetic", "config": config}


# ------------------------------------------------------------------ config 4
def bench_batch(args):
    import asdmatch_amd as A
    from asdmatch_amd import workloads as W

    t0 = time.time()
    g = W.graph("c4")
    man = W.load_manifest()["c4"]
    sl = W.slots("c4")
    qs = [W.query(g, sl[i], seed) for i, seed, _ in man["queries"]]
    want = [c for _, _, c in man["queries"]]
    engines = man.get("engines", ["frontier"] * len(qs))
    if args.max_queries:
        qs, want, engines = qs[: args.max_queries], want[: args.max_queries], engines[: args.max_queries]
    fidx = [i for i, e in enumerate(engines) if e == "frontier"]
    sizes = {}
    for s in sl:
        sizes.setdefault(s.k, [0, 0, 0])[0] += 1
    for (i, _, _), e in zip(man["queries"], man.get("engines", [])):
        sizes[sl[i].k][1] += 1
        sizes[sl[i].k][2] += e == "bags"
    gen_s = time.time() - t0
    say(f"c4: graph and {len(qs)} admitted queries ({len(fidx)} frontier, {len(qs) - len(fidx)} bag engine) "
        f"in {gen_s:.1f} s")

    ctx = A.Context(0)
    ctx.upload(g)
    if args.bag_workers:
        ctx.set_bag_workers(args.bag_workers)
    if args.table_form:
        ctx.set_table_form(args.table_form)
    m = None
    if os.path.exists(ESTIMATOR):
        m = A.Mpnn(ctx)
        m.load(ESTIMATOR)

    def plain(idx):
        out = []
        for i in idx:
            cs = ctx.filter(qs[i], g)
            out.append(0 if cs is None else ctx.count(cs.plan(qs[i]), cs))
        return out

    def batched(idx, mm):
        c, st, lat = ctx.batch_count_learned(g, [qs[i] for i in idx], mm)
        if [int(x) for x in c] != [want[i] for i in idx] or any(st):
            raise SystemExit("batched count mismatch")
        return lat

    every = list(range(len(qs)))
    for w in range(args.warmup):
        say(f"warmup step {w + 1}/{args.warmup}: batched + learned plans")
        batched(every, m)
    # launch shape of the batched engine, picked by the library's autotuner (untimed) on every
    # fourth frontier query, which keeps the tuner's 36 shapes to seconds
    fq = [qs[i] for i in fidx]
    if args.only_timed:
        tune = ctx.batch_tuning() + (None,)
    elif args.batch_shape:
        lanes, scale, iters = args.batch_shape.split(",")
        ctx.set_batch_tuning(int(lanes), float(scale), int(iters))
        tune = ctx.batch_tuning() + (None,)
    elif args.warmup and fq:
        say("autotuning the batch launch shape")
        tune = ctx.batch_autotune(g, fq[::4], m, reps=1)
    else:
        tune = ctx.batch_tuning() + (None,)

    # frontier-admitted subset: plain vs batched (plain plans) vs batched (learned plans)
    sub = {}
    if fidx and not args.only_timed:
        say("frontier subset: plain, batched, batched + learned")
        reps = max(3, args.steps)
        hip_sync()
        a = time.perf_counter()
        for _ in range(reps):
            if plain(fidx) != [want[i] for i in fidx]:
                raise SystemExit("plain count mismatch")
        hip_sync()
        b = time.perf_counter()
        for _ in range(reps):
            batched(fidx, None)
        hip_sync()
        c = time.perf_counter()
        for _ in range(reps):
            batched(fidx, m)
        hip_sync()
        d = time.perf_counter()
        n = len(fidx) * reps
        sub = {"queries": len(fidx), "steps": reps, "plain_gpu_queries_per_s": round(n / (b - a), 3),
               "batched_plain_plan_queries_per_s": round(n / (c - b), 3),
               "batched_learned_plan_queries_per_s": round(n / (d - c), 3),
               "batched_learned_vs_plain": round((b - a) / (d - c), 3),
               "embeddings_per_s": round(sum(want[i] for i in fidx) * reps / (d - c), 1)}

    # the whole admitted batch: plain, then batched + learned plans (the value, profiled).  The
    # plain path is a side measurement: at most two passes, so a long run (--steps 20) spends its
    # time on the value's K steps
    plain_steps = 0 if args.only_timed else max(1, min(args.steps, 2))
    say(f"timed: plain path over {len(qs)} queries x {plain_steps}")
    hip_sync()
    t1 = time.perf_counter()
    got = want
    for _ in range(plain_steps):
        got = plain(every)
    hip_sync()
    t2 = time.perf_counter()
    if got != want:
        raise SystemExit("plain count mismatch")
    say(f"timed: batched + learned plans over {len(qs)} queries x {args.steps}")
    ctx.set_profiling(True)
    lats = []
    hip_sync()
    t4 = time.perf_counter()
    for _ in range(args.steps):
        lats.extend(batched(every, m).tolist())
    hip_sync()
    t5 = time.perf_counter()
    kstats = ctx.kernel_stats()
    ctx.set_profiling(False)
    nq = len(qs) * args.steps
    plain_qps = len(qs) * plain_steps / (t2 - t1) if plain_steps else None
    learned_qps = nq / (t5 - t4)
    lats.sort()
    # The timed step counts the bag-engine queries on several streams at once, so its kernels
    # share the device and their launch durations include that sharing.  One more step, untimed,
    # with a single worker gives the same kernels' durations when each has the device alone.
    iso = None
    plain_plan_qps = None
    if not args.only_timed:
        # the same batched call with the plain planner's plans: what the learned plans themselves
        # change, apart from batching
        say("untimed: one batched step with plain plans")
        hip_sync()
        t8 = time.perf_counter()
        batched(every, None)
        hip_sync()
        plain_plan_qps = len(qs) / (time.perf_counter() - t8)
        workers = ctx.bag_workers()
        say("untimed: one batched step with one bag worker (isolated kernel durations)")
        ctx.set_bag_workers(1)
        ctx.set_profiling(True)
        t6 = time.perf_counter()
        batched(every, m)
        hip_sync()
        t7 = time.perf_counter()
        iso = (ctx.kernel_stats(), t7 - t6)
        ctx.set_profiling(False)
        ctx.set_bag_workers(workers)

    cpu = None
    if not args.no_cpu_baseline and not args.only_timed:
        # the 2P admitted queries with the fewest embeddings (the reference enumerates every one)
        workers = min(16, os.cpu_count() or 1)
        order = sorted(fidx, key=lambda i: want[i])[: 2 * workers]
        cpu = cpu_baseline_sample(g, [(qs[i], want[i]) for i in order], tempfile.mkdtemp(prefix="asdm_bench_"),
                                  workers, 600, f"the {len(order)} admitted config-4 queries with the fewest embeddings")
    admitted = {str(k): {"slots": v[0], "admitted": v[1], "bag_engine": v[2]} for k, v in sorted(sizes.items())}
    out = base_line(args, round(learned_qps, 4), round((t5 - t4) / args.steps * 1e3, 3), {
        "workload": "config 4: R-MAT scale 20, edge factor 10, (a,b,c)=(.57,.19,.19), 16 uniform labels; every admitted "
                    "random-walk query of the 1,024 slots, one asdm_batch_count_learned call per step",
        "slots_total": len(sl), "queries_per_step": len(qs),
        "admitted_fraction": round(len(man["queries"]) / len(sl), 4),
        "admitted_by_size": admitted,
        "not_admitted_note": "slots whose draws are neither within the frontier row limit nor countable by the "
                             "bag-table engine (a quotient of treewidth > 2, more than 2^14 partition terms, or over "
                             "the 2 s guard); sizes 21-32 have none admitted",
        "admission_row_limit": man["row_limit"], "learned_plans": m is not None, "bag_workers": ctx.bag_workers(), "table_form": args.table_form,
        "batch_launch_shape": {"leaf_lanes": tune[0], "class_scale": tune[1], "tile_iters": tune[2],
                               "autotuned_on": "every fourth frontier query" if tune[3] is not None else None}})
    out.update({
        "plain_gpu_queries_per_s": round(plain_qps, 4) if plain_qps else None, "plain_gpu_steps": plain_steps,
        "batched_learned_plan_queries_per_s": round(learned_qps, 4),
        "batched_learned_vs_plain": round(learned_qps / plain_qps, 3) if plain_qps else None,
        "batched_plain_plan_queries_per_s": round(plain_plan_qps, 4) if plain_plan_qps else None,
        "learned_vs_plain_plans": round(learned_qps / plain_plan_qps, 3) if plain_plan_qps else None,
        "embeddings_per_s": None,
        "embeddings_note": "counts are exact mod 2^64 like the reference's size_t; bag-engine queries have far more "
                           "than 2^64 embeddings, so their counts are residues and no embedding rate is summed over "
                           "them (see frontier_subset.embeddings_per_s)",
        "frontier_subset": sub,
        "latency_ms_p50": round(lats[len(lats) // 2], 3), "latency_ms_p99": round(lats[int(len(lats) * 0.99)], 3),
        "latency_note": "per query, from the call start to the host synchronization after its last level",
        "roofline": roofline_of(kstats, "c4"),
        "roofline_note": "timed step: the bag-engine queries run on several streams at once, so launch durations "
                         "include the sharing of the device; roofline_isolated is the same from an untimed step with "
                         "one bag worker",
        "roofline_isolated": ({**roofline_of(iso[0], "c4_isolated"), "step_s": round(iso[1], 3)} if iso else None),
        "device_aggregate": device_aggregate(kstats, args.steps, (t5 - t4) / args.steps),
        "cpu_baseline": cpu, "kernels": kernel_table(kstats),
        "kernels_isolated": kernel_table(iso[0]) if iso else None, "setup_s": round(gen_s, 2)})
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ config 3
def bench_local(args):
    import numpy as np

    import asdmatch_amd as A
    from asdmatch_amd import workloads as W

    t0 = time.time()
    g = W.graph("c3")
    (slot, seed, count), = [tuple(x) for x in W.load_manifest()["c3"]["queries"]]
    q = W.query(g, W.slots("c3")[slot], seed)
    gen_s = time.time() - t0
    ctx = A.Context(0)
    if args.table_form:
        ctx.set_table_form(args.table_form)
    ctx.upload(g)

    def step():
        cs = ctx.filter(q, g)
        plan = cs.plan(q, local=True)
        mat, c = ctx.local_count(plan, cs)
        if c != count:
            raise SystemExit(f"count mismatch: {c} != {count}")
        return mat, plan.uses_bags

    for w in range(args.warmup):
        say(f"warmup step {w + 1}/{args.warmup}")
        mat, bags = step()
        if not all(int(s) == count for s in mat.sum(axis=0, dtype=np.uint64)):
            raise SystemExit("column sums differ from the count")
    say(f"timed: {args.steps} local counts")
    ctx.set_profiling(True)
    hip_sync()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        mat, bags = step()
    hip_sync()
    t2 = time.perf_counter()
    kstats = ctx.kernel_stats()
    ctx.set_profiling(False)
    secs = (t2 - t1) / args.steps
    cpu = None
    if not args.no_cpu_baseline and os.path.exists(os.path.join(REF_BIN, "ref_driver_local")):
        tmpdir = tempfile.mkdtemp(prefix="asdm_bench_")
        dpath, qpath, csv = (os.path.join(tmpdir, n) for n in ("c3.graph", "c3_query.graph", "c3.csv"))
        g.write(dpath)
        q.write(qpath)
        limit = 90
        say(f"cpu baseline: reference LOCAL_COUNT build on the same query, {limit} s wall-clock limit")
        t3 = time.perf_counter()
        try:
            r = subprocess.run([os.path.join(REF_BIN, "ref_driver_local"), "-q", qpath, "-d", dpath, "-c", csv],
                               capture_output=True, text=True, timeout=limit)
            v = parse_ref(r.stdout)
            finished = "exec" in v
        except subprocess.TimeoutExpired:
            finished, v = False, {}
        wall = time.perf_counter() - t3
        cpu = {"value": round(1.0 / limit, 4) if not finished else round(1.0 / (v["filter"] + v["plan"] + v["exec"]), 4),
               "unit": "queries/s", "cores": 1, "kind": "reference", "bound": "upper" if not finished else "measured",
               "sample": f"the same 12-vertex query and graph, reference ALL_LEVEL + LOCAL_COUNT build, one thread, "
                         + (f"stopped after {limit} s of wall time, graph load included (the value is the upper "
                            f"bound 1/{limit}: the build enumerates the {count} embeddings one by one)"
                            if not finished else "finished"),
               "wall_s": round(wall, 1)}
    out = base_line(args, round(1.0 / secs, 4), round(secs * 1e3, 3), {
        "workload": "config 3: Chung-Lu 200,000 V / 4,000,000 E, beta 2.5, 16 uniform labels; the admitted 12-vertex "
                    "random-walk query (first draw of the seed sequence), full |Vg| x |Vq| LOCAL_COUNT matrix per step",
        "query_vertices": q.n, "query_edges": int(len(q.edges())), "engine": "bags" if bags else "frontier",
        "slots_total": 1, "admitted": 1})
    out.update({"count": count, "embeddings_per_s": round(count / secs, 1),
                "matrix": {"rows": g.n, "cols": q.n, "column_sums_equal_count": True},
                "roofline": roofline_of(kstats, "c3"), "cpu_baseline": cpu, "kernels": kernel_table(kstats),
                "setup_s": round(gen_s, 2)})
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ config 2
def bench_c2(args):
    import asdmatch_amd as A
    from asdmatch_amd import workloads as W

    t0 = time.time()
    g = W.graph("c2")
    adm = W.admitted("c2")
    queries = [(W.query(g, s, seed), seed, cnt) for s, seed, cnt in adm]
    gen_s = time.time() - t0
    ctx = A.Context(0)
    if args.table_form:
        ctx.set_table_form(args.table_form)
    ctx.upload(g)

    def step():
        total = 0
        for q, seed, expect in queries:
            cs = ctx.filter(q, g)
            c = ctx.count(cs.plan(q), cs)
            if c != expect:
                raise SystemExit(f"count mismatch for query seed {seed}: {c} != {expect}")
            total += c
        return total

    for w in range(args.warmup):
        say(f"warmup step {w + 1}/{args.warmup}")
        step()
    say(f"timed: {args.steps} steps of {len(queries)} queries")
    ctx.set_profiling(True)
    hip_sync()
    t1 = time.perf_counter()
    emb = 0
    for _ in range(args.steps):
        emb += step()
    hip_sync()
    t2 = time.perf_counter()
    kstats = ctx.kernel_stats()
    ctx.set_profiling(False)
    elapsed = t2 - t1
    cpu = None
    if not args.no_cpu_baseline and queries:
        smallest = min(queries, key=lambda t: t[2])
        cpu = cpu_baseline_sample(g, [(smallest[0], smallest[2])], tempfile.mkdtemp(prefix="asdm_bench_"), 1, 900,
                                  "the admitted config-2 query with the fewest embeddings")
    man = W.load_manifest()["c2"]
    out = base_line(args, round(len(queries) * args.steps / elapsed, 4), round(elapsed / args.steps * 1e3, 3), {
        "workload": "config 2: Chung-Lu 1M V / 15M E, beta 2.5, 32 labels Zipf(1.2); admitted random-walk queries, "
                    "plain GPU engine (filter + plan + exact count)",
        "slots_total": 16, "queries_per_step": len(queries), "query_sizes": [q.n for q, _, _ in queries],
        "admission_row_limit": man["row_limit"]})
    out.update({"embeddings_per_s": round(emb / elapsed, 1), "embeddings_per_step": emb // max(1, args.steps),
                "roofline": roofline_of(kstats, "c2"), "cpu_baseline": cpu, "kernels": kernel_table(kstats),
                "setup_s": round(gen_s, 2)})
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed steps (c4: one batched call over the whole batch each, ~3 s; its rate varies by about 5 %% "
                         "between single steps)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("c2", "c3", "c4"), default="c4",
                    help="c4: config-4 batch, batched + learned plans vs plain (default); c3: config-3 local counts; "
                         "c2: config-2 queries one by one through the plain path")
    ap.add_argument("--batch-shape", default=None,
                    help="LEAF_LANES,CLASS_SCALE,TILE_ITERS: use this batch launch shape instead of autotuning")
    ap.add_argument("--max-queries", type=int, default=0, help="c4 only: the first N admitted queries (profiling)")
    ap.add_argument("--only-timed", action="store_true",
                    help="c4 only: warmup and the timed batched steps alone (no plain path, frontier subset, "
                         "autotuning, isolated step or CPU baseline), for profilers")
    ap.add_argument("--bag-workers", type=int, default=0, help="c4 only: bag-engine worker streams (default: library's)")
    ap.add_argument("--table-form", type=int, default=0, choices=(0, 1, 2),
                    help="relation-product form of the bag engine (0 per product, 1 LDS rows, 2 zero-fill + scatter)")
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
    {"c4": bench_batch, "c3": bench_local, "c2": bench_c2}[args.workload](args)


if __name__ == "__main__":
    main()
