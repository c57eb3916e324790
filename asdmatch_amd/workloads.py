"""Seeded synthetic workloads of the benchmark configurations (graphs and random-walk queries).

Every graph and query is a pure function of its seed (libasdm's portable xoshiro256** streams).
Configurations (sizes and label laws as specified for the benchmark):
  c1  Chung-Lu 3,000 V / 12,500 E, beta 2.5, 70 labels Zipf(1.0); 8-vertex walk query, induced edges
  c2  Chung-Lu 1,000,000 V / 15,000,000 E, beta 2.5, 32 labels Zipf(1.2); 16 walk queries of 8, 12
      and 16 vertices (6/5/5), alternately sparse (walk tree + at most 2 induced edges) and dense
      (all induced edges)
  c3  Chung-Lu 200,000 V / 4,000,000 E, beta 2.5, 16 uniform labels; one 12-vertex walk query
  c4  R-MAT scale 20, edge factor 10, (a,b,c) = (.57,.19,.19), deduplicated, 16 uniform labels;
      1,024 walk queries with sizes uniform in 4..32 and edge density uniform in [tree, induced]

Query admission.  Random-walk queries on these power-law graphs mostly have astronomically many
embeddings (10^13 and far beyond for 8-vertex trees on c2), which no exact method enumerates.  A
query slot therefore draws walks from a fixed seed sequence and admits the first that one of the two
exact engines counts within its limit: the frontier engine within `row_limit` materialized partial
embeddings (c4: counted by the batched engine, all slots of a draw round in one batch), the
bag-table engine (queries the planner sends to it) within a time guard.  The admitted seeds, counts
and engines are recorded in workload_manifest.json so every run measures the same queries.  The row
limit depends only on the data, the query and the deterministic planner; the time guard of the
bag-table engine (2 s on c4) also depends on device speed.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import AsdmError, Context, Graph

MANIFEST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workload_manifest.json")

SEEDS = {"c1": 0xA5D00001, "c2": 0xA5D00002, "c3": 0xA5D00003, "c4": 0xA5D00004}
UNCAPPED = 0xFFFFFFFF


def graph(cfg: str) -> Graph:
    s = SEEDS[cfg]
    if cfg == "c1":
        return Graph.chung_lu(3000, 12500, 2.5, 70, 1.0, s)
    if cfg == "c2":
        return Graph.chung_lu(1_000_000, 15_000_000, 2.5, 32, 1.2, s)
    if cfg == "c3":
        return Graph.chung_lu(200_000, 4_000_000, 2.5, 16, 0.0, s)
    if cfg == "c4":
        return Graph.rmat(20, 10, 0.57, 0.19, 0.19, 16, 0.0, s)
    raise ValueError(cfg)


@dataclass
class Slot:
    k: int
    density: float
    max_extra: int


def slots(cfg: str) -> List[Slot]:
    if cfg == "c1":
        return [Slot(8, 1.0, UNCAPPED)]
    if cfg == "c2":
        sizes = [8] * 6 + [12] * 5 + [16] * 5
        return [Slot(k, 1.0, 2 if i % 2 == 0 else UNCAPPED) for i, k in enumerate(sizes)]
    if cfg == "c3":
        return [Slot(12, 1.0, UNCAPPED)]
    if cfg == "c4":
        rng = np.random.Generator(np.random.PCG64(SEEDS["c4"]))
        ks = rng.integers(4, 33, size=1024)
        ds = rng.random(1024)
        return [Slot(int(k), float(d), UNCAPPED) for k, d in zip(ks, ds)]
    raise ValueError(cfg)


def draw_seed(cfg: str, slot: int, j: int) -> int:
    return SEEDS[cfg] + 1_000_000 * (slot + 1) + j


def query(g: Graph, s: Slot, seed: int) -> Graph:
    return g.random_walk_query(s.k, s.density, s.max_extra, seed)


def admit(ctx: Context, g: Graph, cfg: str, row_limit: int, max_draws: int = 40, local: bool = False,
          guard_ms: float = 10000, log=print) -> list:
    """Walks each slot's seed sequence and returns [(slot index, seed, count, engine)] of admitted
    queries, engine being the one the planner chose ("frontier" or "bags").

    A frontier draw is rejected when it exceeds row_limit; guard_ms only protects the run against a
    draw whose work is extreme (such draws are logged as 'guard')."""
    out = []
    ctx.set_row_limit(row_limit)
    ctx.set_time_limit(guard_ms)
    try:
        for i, s in enumerate(slots(cfg)):
            got = None
            for j in range(max_draws):
                seed = draw_seed(cfg, i, j)
                try:
                    q = query(g, s, seed)
                except AsdmError:
                    log(f"  {cfg} slot {i} draw {j}: no walk")
                    continue
                cs = ctx.filter(q, g)
                if cs is None:
                    log(f"  {cfg} slot {i} draw {j}: empty candidate set")
                    continue
                plan = cs.plan(q, local=local)
                engine = "bags" if plan.uses_bags else "frontier"
                try:
                    t0 = time.perf_counter()
                    c = (ctx.local_count(plan, cs)[1] if local else ctx.count(plan, cs))
                    dt = time.perf_counter() - t0
                except AsdmError as e:
                    if e.code != -8:
                        raise
                    log(f"  {cfg} slot {i} draw {j}: rejected by the {engine} engine "
                        f"({'guard' if 'time' in str(e) else 'rows'})")
                    continue
                got = (i, seed, c, engine)
                log(f"  {cfg} slot {i} draw {j}: admitted ({engine} engine), count {c}, {dt * 1e3:.2f} ms")
                break
            log(f"{cfg} slot {i} k={s.k} -> {got}")
            if got:
                out.append(got)
    finally:
        ctx.set_row_limit(0)
        ctx.set_time_limit(0)
    return out


def admit_batched(ctx: Context, g: Graph, cfg: str, row_limit: int, max_draws: int, log=print,
                  guard_ms: float = 2000, state: Optional[dict] = None, stop_after_s: float = 1e9) -> list:
    """The admission rule for a many-query configuration, evaluated with the batched engine: draw j
    of every slot not yet admitted is counted in one asdm_batch_count_learned call (plain plans);
    a slot admits its first draw that the call counts, i.e. one that stays within the per-query
    row limit on materialized partial embeddings (frontier engine) or, for a query the planner
    sends to the bag-table engine, within guard_ms.  Returns [(slot index, seed, count, engine)].

    state (optional) carries the progress between calls: {"next_draw": j, "got": [...]}; the call
    returns after the draw round that passes stop_after_s, with state updated."""
    sl = slots(cfg)
    state = state if state is not None else {}
    got = {int(x[0]): tuple(x) for x in state.get("got", [])}
    t_start = time.perf_counter()
    ctx.set_row_limit(row_limit)
    ctx.set_time_limit(guard_ms)
    try:
        for j in range(state.get("next_draw", 0), max_draws):
            todo, qs = [], []
            for i, s in enumerate(sl):
                if i in got:
                    continue
                try:
                    qs.append(query(g, s, draw_seed(cfg, i, j)))
                    todo.append(i)
                except AsdmError:
                    pass
            if not qs:
                break
            t0 = time.perf_counter()
            counts, st, _ = ctx.batch_count_learned(g, qs)
            nbag = 0
            for i, q, c, code in zip(todo, qs, counts, st):
                if code == 0:
                    cs = ctx.filter(q, g)
                    engine = "bags" if cs.plan(q).uses_bags else "frontier"
                    nbag += engine == "bags"
                    got[i] = (i, draw_seed(cfg, i, j), int(c), engine)
            log(f"  {cfg} draw {j}: {len(qs)} slots tried, {int((st == 0).sum())} admitted ({nbag} by the bag "
                f"engine), {len(got)} in total, {time.perf_counter() - t0:.1f} s")
            state["next_draw"] = j + 1
            state["got"] = [list(got[i]) for i in sorted(got)]
            if time.perf_counter() - t_start > stop_after_s:
                break
    finally:
        ctx.set_row_limit(0)
        ctx.set_time_limit(0)
    return [got[i] for i in sorted(got)]


def load_manifest() -> dict:
    with open(MANIFEST) as f:
        return json.load(f)


def admitted(cfg: str) -> list:
    """[(Slot, seed, expected count)] of the recorded admitted queries of cfg."""
    m = load_manifest()[cfg]
    sl = slots(cfg)
    return [(sl[i], seed, count) for i, seed, count in m["queries"]]


ROW_LIMITS = {"c1": 1 << 40, "c2": 1 << 30, "c3": 1 << 27, "c4": 1 << 24}
MAX_DRAWS = {"c1": 4, "c2": 30, "c3": 40, "c4": 8}
BATCHED_RULE = {"c4"}  # configurations admitted with the batched engine (many query slots)


def build_manifest(cfgs=("c1", "c2", "c3", "c4"), log=print, resume: bool = False, stop_after_s: float = 1e9) -> dict:
    """Runs the admission rule on the GPU and writes workload_manifest.json.  With resume, a
    batched configuration continues from the draw round its manifest entry's "progress" records
    (the admission of a large batch may span several runs, each stopping after stop_after_s)."""
    m = load_manifest() if os.path.exists(MANIFEST) else {}
    ctx = Context(0)
    for cfg in cfgs:
        g = graph(cfg)
        ctx.upload(g)
        local = cfg == "c3"
        if cfg in BATCHED_RULE:
            state = dict(m.get(cfg, {}).get("progress", {})) if resume else {}
            got = admit_batched(ctx, g, cfg, ROW_LIMITS[cfg], MAX_DRAWS[cfg], log=log, state=state,
                                stop_after_s=stop_after_s)
            q = [x[:3] for x in got]
            engines = [x[3] for x in got]
            m[cfg] = {"row_limit": ROW_LIMITS[cfg], "local_plan": local, "queries": q, "engines": engines,
                      "slots": len(slots(cfg)), "rule": "batched", "progress": state,
                      "complete": state.get("next_draw", 0) >= MAX_DRAWS[cfg]}
            with open(MANIFEST, "w") as f:
                json.dump(m, f, indent=1)
            del g
            continue
        else:
            got = admit(ctx, g, cfg, ROW_LIMITS[cfg], max_draws=MAX_DRAWS[cfg], local=local, log=log)
            q = [x[:3] for x in got]
            engines = [x[3] for x in got]
        m[cfg] = {"row_limit": ROW_LIMITS[cfg], "local_plan": local, "queries": q, "engines": engines,
                  "slots": len(slots(cfg)), "rule": "batched" if cfg in BATCHED_RULE else "single"}
        with open(MANIFEST, "w") as f:
            json.dump(m, f, indent=1)
        del g
    ctx.close()
    return m


if __name__ == "__main__":
    import sys

    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    resume = "--resume" in sys.argv
    stop = float(next((a.split("=")[1] for a in sys.argv if a.startswith("--stop-after=")), 1e9))
    build_manifest(tuple(args) or ("c1", "c2", "c3", "c4"), log=lambda *a: print(*a, flush=True), resume=resume,
                   stop_after_s=stop)
