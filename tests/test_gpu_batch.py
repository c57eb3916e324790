"""Batched multi-query engine (asdm_batch_count): the same exact counts as the reference's golden
outputs and as the single-query engine."""
import os

import numpy as np
import pytest

import asdmatch_amd as A
from tests import common as T

pytestmark = pytest.mark.gpu


def test_batch_matches_reference_on_yeast(ctx):
    gold = T.yeast_golden()["queries"]
    g = A.Graph.load(T.yeast_data_path())
    names = list(gold)
    qs = [A.Graph.load(T.yeast_query_path(n)) for n in names]
    counts, lat = ctx.batch_count(g, qs)
    assert [int(c) for c in counts] == [gold[n]["count"] for n in names]
    assert np.all(lat > 0)


def test_batch_matches_reference_on_synthetic(ctx):
    for case in T.synth_golden():
        g = T.make_graph(case)
        recs = [r for r in case["queries"] if r["count"] is not None]
        qs = [g.random_walk_query(r["k"], r["density"], r["max_extra"], r["seed"]) for r in recs]
        counts, _ = ctx.batch_count(g, qs)
        assert [int(c) for c in counts] == [r["count"] for r in recs], case["name"]


def test_config4_counts_match_reference(ctx):
    """Admitted config-4 queries pinned to the reference itself: the counts that the reference's
    exact (ALL_LEVEL) build printed for the admitted queries it finished within its time limit
    (tests/golden/c4_reference.json), through the batched engine with learned plans and through
    the plain path."""
    import json

    from asdmatch_amd import workloads as W

    with open(f"{T.GOLDEN}/c4_reference.json") as f:
        ref = [r for r in json.load(f)["queries"] if r["status"] == "ok"]
    assert len(ref) >= 30
    g = W.graph("c4")
    sl = W.slots("c4")
    qs = [W.query(g, sl[r["slot"]], r["seed"]) for r in ref]
    want = [r["count"] for r in ref]
    m = A.Mpnn(ctx)
    m.load(os.path.join(os.path.dirname(A.__file__), "data", "mpnn_estimator.bin"))
    counts, st, lat = ctx.batch_count_learned(g, qs, m)
    assert not any(st)
    assert [int(c) for c in counts] == want
    assert [A.count_embeddings(ctx, q, g) for q in qs[::4]] == want[::4]


def test_batch_latency_is_per_query(ctx):
    """lat_ms holds each query's own completion time: positive, within the call's duration, and not
    one value for the whole batch when its queries finish at different depths."""
    import time

    case = [c for c in T.synth_golden() if c["name"] == "rmat_small"][0]
    g = T.make_graph(case)
    qs = [g.random_walk_query(r["k"], r["density"], r["max_extra"], r["seed"]) for r in case["queries"]]
    t0 = time.perf_counter()
    counts, lat = ctx.batch_count(g, qs)
    call_ms = (time.perf_counter() - t0) * 1e3
    assert [int(c) for c in counts] == [r["count"] for r in case["queries"]]
    assert np.all(lat > 0) and np.all(lat <= call_ms + 1)
    assert len(set(np.round(lat, 4).tolist())) > 1


def test_batch_config4_admitted_queries(ctx):
    """Every admitted config-4 query, counted by the batched engine (frontier and bag-table queries
    together), against the manifest's counts.  The manifest was written by this library's admission
    run, so this is agreement between engines and runs; the reference-pinned subset is
    test_config4_counts_match_reference."""
    from asdmatch_amd import workloads as W

    g = W.graph("c4")
    adm = W.admitted("c4")
    qs = [W.query(g, s, seed) for s, seed, _ in adm]
    counts, _ = ctx.batch_count(g, qs)
    assert [int(c) for c in counts] == [c for _, _, c in adm]


def test_batch_budget_fallback_gives_same_counts(ctx):
    case = [c for c in T.synth_golden() if c["name"] == "pl_small_l4"][0]
    g = T.make_graph(case)
    qs = [g.random_walk_query(r["k"], r["density"], r["max_extra"], r["seed"]) for r in case["queries"]]
    counts, _ = ctx.batch_count(g, qs, budget_bytes=1 << 12)  # forces the per-query fallback
    assert [int(c) for c in counts] == [r["count"] for r in case["queries"]]


def test_config4_heaviest_queries_single_and_batched(ctx):
    """Admitted config-4 queries with the largest counts: the single-query engine, the batched engine
    and the batched engine with an untrained estimator's arbitrary plans must agree with the
    manifest's counts (agreement between engines; the manifest is this library's own output)."""
    from asdmatch_amd import workloads as W

    g = W.graph("c4")
    adm = W.admitted("c4")
    heavy = sorted(adm, key=lambda t: -t[2])[:8]
    qs = [W.query(g, s, seed) for s, seed, _ in heavy]
    want = [c for _, _, c in heavy]
    assert [A.count_embeddings(ctx, q, g) for q in qs] == want
    counts, _ = ctx.batch_count(g, qs)
    assert [int(c) for c in counts] == want
    m = A.Mpnn(ctx, seed=5)
    counts, status, _ = ctx.batch_count_learned(g, qs, m)
    assert [int(c) for c in counts] == want and not any(status)


def test_autotune_and_launch_shapes_keep_counts(ctx):
    """Every launch shape gives the same counts (the autotuner checks this across its whole grid)
    and the tuner keeps a shape from that grid on the context."""
    case = [c for c in T.synth_golden() if c["name"] == "pl_small_l4"][0]
    g = T.make_graph(case)
    recs = [r for r in case["queries"] if r["count"] is not None]
    qs = [g.random_walk_query(r["k"], r["density"], r["max_extra"], r["seed"]) for r in recs]
    want = [r["count"] for r in recs]
    default = ctx.batch_tuning()
    try:
        for shape in ((1, 0.5, 1), (64, 2.0, 8), (4, 1.0, 64)):
            ctx.set_batch_tuning(*shape)
            assert ctx.batch_tuning() == shape
            counts, _ = ctx.batch_count(g, qs)
            assert [int(c) for c in counts] == want, shape
        for bad in ((3, 1.0, 4), (16, 0.0, 4), (16, 1.0, 0), (16, 1.0, 65)):
            with pytest.raises(A.AsdmError):
                ctx.set_batch_tuning(*bad)
        lanes, scale, iters, ms = ctx.batch_autotune(g, qs, reps=1)
        assert lanes in (4, 16, 64) and scale in (0.25, 0.5, 1.0, 2.0) and iters in (2, 4, 8) and ms > 0
        assert ctx.batch_tuning() == (lanes, scale, iters)
        counts, _ = ctx.batch_count(g, qs)
        assert [int(c) for c in counts] == want
    finally:
        ctx.set_batch_tuning(*default)


def test_bag_workers_keep_counts(ctx):
    """The bag-engine queries of a batch counted by 1, 2 and 4 worker contexts (one stream each):
    the same exact counts, every query stamped with its own completion time, and the setting
    bounded to 1..16."""
    from asdmatch_amd import workloads as W

    man = W.load_manifest()["c4"]
    g = W.graph("c4")
    sl = W.slots("c4")
    bag = [(i, seed, c) for (i, seed, c), e in zip(man["queries"], man["engines"]) if e == "bags"][::6]
    assert len(bag) >= 40
    qs = [W.query(g, sl[i], seed) for i, seed, _ in bag]
    want = [c for _, _, c in bag]
    default = ctx.bag_workers()
    try:
        for n in (1, 2, 4):
            ctx.set_bag_workers(n)
            assert ctx.bag_workers() == n
            counts, st, lat = ctx.batch_count_learned(g, qs)
            assert [int(c) for c in counts] == want and not any(st), n
            assert np.all(lat > 0) and len(set(np.round(lat, 3).tolist())) > len(qs) // 2, n
        for bad in (0, 17):
            with pytest.raises(A.AsdmError):
                ctx.set_bag_workers(bad)
    finally:
        ctx.set_bag_workers(default)
