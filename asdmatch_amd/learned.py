"""Training data, training and evaluation of the learned cardinality estimator.

The estimator (asdm_mpnn_*) predicts log2(1 + #embeddings) of a (sub-)query from node features
computed on its candidate space.  It takes the role of the reference's sampling estimator behind
subsetToCard (plan/estimator.cpp:7, :173-291): the planner asks it for the cardinality of every
connected vertex subset it considers and orders the enumeration greedily by those values.

Training pairs are (sub-query, exact count) produced by this library's own GPU engine:
connected induced sub-queries of random-walk queries on seeded graphs, filtered in one batched
pass (asdm_filter_many), featurized on their own candidate spaces and counted exactly by the
batched engine under a per-query row limit (queries over it are dropped, not guessed).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

import numpy as np

from . import (ASDM_E_EMPTY_CANDIDATES, ASDM_OK, AsdmError, Context, Graph, GraphBatch, Mpnn, mpnn_features,
               plain_estimate)

# Seeded training graphs: two Chung-Lu power-law graphs with the c2 label law (32 labels,
# Zipf 1.2) at smaller scale, one with uniform labels, and an R-MAT graph; none of them is a
# benchmark graph.
TRAIN_GRAPHS: Dict[str, Callable[[], Graph]] = {
    "cl32z": lambda: Graph.chung_lu(200_000, 3_000_000, 2.5, 32, 1.2, 0x7EA10001),
    "cl16u": lambda: Graph.chung_lu(100_000, 1_500_000, 2.5, 16, 0.0, 0x7EA10002),
    "rmat16u": lambda: Graph.rmat(17, 10, 0.57, 0.19, 0.19, 16, 0.0, 0x7EA10003),
}


def connected_masks(off: np.ndarray, nbr: np.ndarray, rng: np.random.Generator, count: int,
                    min_size: int = 2) -> List[int]:
    """Random connected vertex subsets (bitmasks) of a query graph, grown one random frontier
    vertex at a time from a random start to a random target size; the full vertex set included."""
    n = len(off) - 1
    full = (1 << n) - 1
    out = {full}
    for _ in range(count * 3):
        if len(out) > count or n <= min_size:
            break
        k = int(rng.integers(min_size, n + 1))
        start = int(rng.integers(0, n))
        mask, members = 1 << start, [start]
        while len(members) < k:
            front = sorted({int(w) for u in members for w in nbr[off[u]:off[u + 1]] if not (mask >> int(w)) & 1})
            if not front:
                break
            w = front[int(rng.integers(0, len(front)))]
            mask |= 1 << w
            members.append(w)
        out.add(mask)
    return sorted(out)


def induced(q: Graph, mask: int) -> Graph:
    """The sub-query induced by mask, vertices renumbered in ascending original id."""
    off, nbr, lab = q.csr()
    ids = [u for u in range(len(lab)) if (mask >> u) & 1]
    pos = {u: i for i, u in enumerate(ids)}
    edges = [(pos[u], pos[int(w)]) for u in ids for w in nbr[off[u]:off[u + 1]] if u < int(w) and int(w) in pos]
    return Graph.from_edges([int(lab[u]) for u in ids], np.array(edges, np.uint32).reshape(-1, 2))


@dataclass
class PairSet:
    """Estimator inputs with exact counts (and the plain estimator's value for comparison)."""
    batch: GraphBatch
    counts: np.ndarray  # uint64 exact counts
    plain: np.ndarray   # independence-model estimates on the same candidate spaces
    sizes: np.ndarray   # vertices per graph

    @property
    def targets(self) -> np.ndarray:
        return np.log2(1.0 + self.counts.astype(np.float64)).astype(np.float32)

    def __len__(self):
        return len(self.counts)

    def take(self, idx) -> "PairSet":
        idx = np.asarray(idx)
        return PairSet(self.batch.subset(idx), self.counts[idx], self.plain[idx], self.sizes[idx])

    @staticmethod
    def concat(sets) -> "PairSet":
        sets = [s for s in sets if len(s)]
        return PairSet(GraphBatch.concat([s.batch for s in sets]), np.concatenate([s.counts for s in sets]),
                       np.concatenate([s.plain for s in sets]), np.concatenate([s.sizes for s in sets]))


def make_pairs(ctx: Context, g: Graph, queries: List[Graph], per_query: int, rng: np.random.Generator,
               row_limit: int = 1 << 24, chunk: int = 512, log: Optional[Callable] = None,
               seconds: Optional[float] = None) -> PairSet:
    """(sub-query, exact count) pairs: per_query connected induced sub-queries of each query (and
    the query itself), counted by the batched engine; sub-queries over row_limit are dropped, as
    are those with an empty candidate set (the planner never sees them).  With `seconds`, no new
    chunk starts after that much time."""
    t_start = time.perf_counter()
    subs = []
    for q in queries:
        off, nbr, _ = q.csr()
        for mask in connected_masks(off, nbr, rng, per_query):
            subs.append(induced(q, mask))
    parts = []
    ctx.upload(g)
    ctx.set_row_limit(row_limit)
    try:
        for c0 in range(0, len(subs), chunk):
            if seconds is not None and time.perf_counter() - t_start > seconds:
                if log:
                    log(f"  time budget reached after {c0} of {len(subs)} sub-queries")
                break
            sq = subs[c0:c0 + chunk]
            t0 = time.perf_counter()
            css = ctx.filter_many(sq, g)
            counts, st, _ = ctx.batch_count_learned(g, sq)
            keep = [i for i in range(len(sq)) if st[i] == ASDM_OK and css[i] is not None]
            feats, plain = [], []
            for i in keep:
                full = np.array([(1 << sq[i].n) - 1], np.uint64)
                feats.append(mpnn_features(g, css[i], full))
                plain.append(plain_estimate(css[i], full)[0])
            if keep:
                parts.append(PairSet(GraphBatch.concat(feats), counts[keep].astype(np.uint64), np.array(plain),
                                     np.array([sq[i].n for i in keep])))
            if log:
                dropped = int(np.sum(st == -8))
                empty = int(np.sum(st == ASDM_E_EMPTY_CANDIDATES))
                log(f"  pairs {c0 + len(sq)}/{len(subs)}: kept {len(keep)}, over limit {dropped}, empty {empty}, "
                    f"{(time.perf_counter() - t0) * 1e3:.0f} ms")
            del css
    finally:
        ctx.set_row_limit(0)
    return PairSet.concat(parts) if parts else PairSet(GraphBatch([0], np.zeros((0, 8)), [0], []),
                                                      np.zeros(0, np.uint64), np.zeros(0), np.zeros(0, int))


def random_queries(g: Graph, n: int, kmin: int, kmax: int, seed: int) -> List[Graph]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = int(rng.integers(kmin, kmax + 1))
        dens = float(rng.random())
        try:
            out.append(g.random_walk_query(k, dens, 0xFFFFFFFF, seed * 1000 + i))
        except AsdmError:
            pass
    return out


def train(m: Mpnn, data: PairSet, epochs: int, lr: float = 3e-3, batch: int = 512, seed: int = 0,
          log: Optional[Callable] = None) -> List[float]:
    """Adam on 0.5 * mean (y - log2(1 + count))^2 over shuffled mini-batches; the learning rate
    follows a cosine from lr down to lr / 10.  Returns the mean loss of each epoch."""
    rng = np.random.default_rng(seed)
    t = data.targets
    n = len(data)
    losses = []
    steps = epochs * ((n + batch - 1) // batch)
    step = 0
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot = 0.0
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            cur = lr * (0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * step / max(1, steps))))
            tot += m.train_step(data.batch.subset(idx), t[idx], cur) * len(idx)
            step += 1
        losses.append(tot / n)
        if log:
            log(f"  epoch {ep}: loss {losses[-1]:.4f}")
    return losses


def q_errors(est, true) -> np.ndarray:
    e = np.maximum(np.asarray(est, np.float64), 1.0)
    t = np.maximum(np.asarray(true, np.float64), 1.0)
    return np.maximum(e / t, t / e)


def evaluate(m: Mpnn, data: PairSet) -> dict:
    """Median / p95 q-error of the network and of the plain independence estimator."""
    y = m.forward(data.batch).astype(np.float64)
    est = np.exp2(np.clip(y, 0, 64)) - 1.0
    ql = q_errors(est, data.counts)
    qp = q_errors(data.plain, data.counts)
    return {"n": len(data), "learned_median": float(np.median(ql)), "learned_p95": float(np.percentile(ql, 95)),
            "plain_median": float(np.median(qp)), "plain_p95": float(np.percentile(qp, 95))}
