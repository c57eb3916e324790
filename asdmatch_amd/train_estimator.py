"""Trains the learned cardinality estimator on the GPU and writes its weights and a report.

    python -m asdmatch_amd.train_estimator [--pairs 200000] [--epochs 24]
        [--out asdmatch_amd/data/mpnn_estimator.bin] [--report estimator_report.json]

Training pairs (sub-query, exact count) come from this library's GPU engine: connected induced
sub-queries of (a) the config-2 query slots' random-walk draws on the config-2 graph (Chung-Lu,
1M vertices, 15M edges, 32 Zipf labels) and (b) random-walk queries of 3-12 vertices on the seeded
training graphs of asdmatch_amd.learned.TRAIN_GRAPHS.  Queries are split 90/10 into train and
held-out sets before their sub-queries are drawn, so no held-out sub-query shares a parent query
with a training pair.

Evaluation on config 4 (R-MAT scale 20, never trained on): the admitted queries' exact counts and
connected sub-queries of them; median and p95 q-error of the network and of the plain
independence estimator.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import Context, Mpnn, mpnn_features, plain_estimate
from . import learned as LR
from . import workloads as W

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "mpnn_estimator.bin")


def log(*a):
    print(*a, flush=True)


def c2_queries(g, per_slot: int):
    out = []
    for i, s in enumerate(W.slots("c2")):
        for j in range(per_slot):
            try:
                out.append(W.query(g, s, W.draw_seed("c2", i, 100 + j)))
            except Exception:  # a walk that cannot be drawn is skipped
                pass
    return out


def collect(ctx, target_pairs: int, seed: int, seconds: float):
    rng = np.random.default_rng(seed)
    # (name, graph factory, query factory, share of the pairs, row limit, chunk); the config-2
    # graph's sub-queries are the costliest to count exactly, so they get a lower row limit
    sources = [("c2", lambda: W.graph("c2"), lambda g: c2_queries(g, 12), 0.1, 1 << 22, 256)]
    for k, (name, mk) in enumerate(LR.TRAIN_GRAPHS.items()):
        # 12,000 queries per graph: sub-queries over the row limit are dropped, so 8,000 left the
        # total near 150k pairs
        sources.append((name, mk, lambda g, k=k: LR.random_queries(g, 12000, 3, 12, 9100 + k), 0.3, 1 << 24, 1024))
    train, held = [], []
    for name, mk, mkq, share, limit, chunk in sources:
        t0 = time.perf_counter()
        g = mk()
        qs = mkq(g)
        cut = int(len(qs) * 0.9)
        want = int(target_pairs * share)
        per_q = max(2, int(np.ceil(want / max(1, len(qs)))))
        budget = seconds * share
        log(f"{name}: {len(qs)} queries, {per_q} sub-queries each (graph {time.perf_counter() - t0:.1f} s)")
        train.append(LR.make_pairs(ctx, g, qs[:cut], per_q, rng, row_limit=limit, chunk=chunk, log=log,
                                   seconds=0.9 * budget))
        held.append(LR.make_pairs(ctx, g, qs[cut:], per_q, rng, row_limit=limit, chunk=chunk, log=log,
                                  seconds=0.1 * budget))
        log(f"{name}: {len(train[-1])} train / {len(held[-1])} held-out pairs in {time.perf_counter() - t0:.1f} s")
        del g
    return LR.PairSet.concat(train), LR.PairSet.concat(held)


def c4_eval(ctx, m, rng):
    """q-errors on config 4: the admitted queries the frontier engine counts (exact counts from the
    manifest, all below 2^64; the bag-engine queries' counts are residues mod 2^64 of far larger
    numbers, so no q-error can be taken against them) and connected sub-queries of them (counted
    here)."""
    g = W.graph("c4")
    ctx.upload(g)
    man = W.load_manifest()["c4"]
    sl = W.slots("c4")
    adm = [(sl[i], seed, c) for (i, seed, c), e in zip(man["queries"], man["engines"]) if e == "frontier"]
    qs = [W.query(g, s, seed) for s, seed, _ in adm]
    css = ctx.filter_many(qs, g)
    feats, plain, counts = [], [], []
    for q, cs, (_, _, c) in zip(qs, css, adm):
        if cs is None:
            continue
        full = np.array([(1 << q.n) - 1], np.uint64)
        feats.append(mpnn_features(g, cs, full))
        plain.append(plain_estimate(cs, full)[0])
        counts.append(c)
    full = LR.PairSet(LR.GraphBatch.concat(feats), np.array(counts, np.uint64), np.array(plain),
                      np.array([q.n for q in qs]))
    subs = LR.make_pairs(ctx, g, qs, 8, rng, row_limit=1 << 24, chunk=1024)
    return LR.evaluate(m, full), LR.evaluate(m, subs)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=200_000)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--report", default=None)
    ap.add_argument("--data-seconds", type=float, default=600.0,
                    help="time budget of pair generation, split across the sources by their share")
    ap.add_argument("--eval-only", default=None, metavar="WEIGHTS",
                    help="only evaluate these weights on config 4 (admitted queries and their sub-queries)")
    a = ap.parse_args(argv)
    ctx = Context(0)
    if a.eval_only:
        m = Mpnn(ctx)
        m.load(a.eval_only)
        c4_full, c4_sub = c4_eval(ctx, m, np.random.default_rng(a.seed + 1))
        rep = {"weights": os.path.basename(a.eval_only), "c4_admitted_full": c4_full, "c4_subqueries": c4_sub}
        log(json.dumps(rep, indent=1))
        if a.report:
            with open(a.report, "w") as f:
                json.dump(rep, f, indent=1)
        del m
        ctx.close()
        return
    t0 = time.perf_counter()
    tr, te = collect(ctx, a.pairs, a.seed, a.data_seconds)
    t_data = time.perf_counter() - t0
    log(f"pairs: {len(tr)} train, {len(te)} held-out ({t_data:.1f} s)")
    m = Mpnn(ctx, seed=a.seed)
    t1 = time.perf_counter()
    losses = LR.train(m, tr, a.epochs, lr=a.lr, batch=a.batch, seed=a.seed, log=log)
    t_train = time.perf_counter() - t1
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    m.save(a.out)
    rep = {"pairs_train": len(tr), "pairs_heldout": len(te), "data_s": t_data, "train_s": t_train,
           "epochs": a.epochs, "loss_first": losses[0], "loss_last": losses[-1],
           "train": LR.evaluate(m, tr), "heldout": LR.evaluate(m, te)}
    c4_full, c4_sub = c4_eval(ctx, m, np.random.default_rng(a.seed + 1))
    rep["c4_admitted_full"] = c4_full
    rep["c4_subqueries"] = c4_sub
    log(json.dumps(rep, indent=1))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    del m
    ctx.close()


if __name__ == "__main__":
    main()
