"""Query-graph paths (setQueryGraphWeights + allSourcesShortestPaths + reconstructPath) and
cartesian-step candidates (optimizedCartesianProduct) through the C ABI, against the reference's
own outputs (tests/golden/paths.json, from oracle/_ref/ref_cs_dump --paths) and the oracle."""
import hashlib
import json
import os

import numpy as np
import pytest

import asdmatch_amd as A
from oracle import oracle as O
from tests import common as T

pytestmark = pytest.mark.gpu


def _cases():
    with open(os.path.join(T.GOLDEN, "paths.json")) as f:
        return json.load(f)


def _data_graph(source, cache={}):
    if source not in cache:
        if source == "yeast":
            cache[source] = A.Graph.load(T.yeast_data_path())
        elif source == "c2":
            from asdmatch_amd import workloads as W

            cache[source] = W.graph("c2")
        else:
            cache[source] = T.make_graph({c["name"]: c for c in T.synth_golden()}[source])
    return cache[source]


def test_query_paths_match_reference(ctx):
    """Distances, successor matrices and reconstructed paths equal the reference's on every fixture
    case, config-2 queries with wrapped products included; a pair whose reference reconstruction
    would not end returns ASDM_E_UNSUPPORTED."""
    loops = 0
    for c in _cases():
        q = A.Graph.from_edges(c["labels"], c["edges"])
        cs = ctx.filter(q, _data_graph(c["source"]))
        assert cs is not None
        dist, nxt = cs.query_paths()
        assert dist.tolist() == c["dist"], (c["source"], c["query"])
        assert nxt.tolist() == c["next"], (c["source"], c["query"])
        for i in range(c["nq"]):
            for j in range(c["nq"]):
                want = c["paths"][i][j]
                if want is None:
                    with pytest.raises(A.AsdmError) as e:
                        cs.query_path(i, j)
                    assert e.value.code == -7
                    loops += 1
                else:
                    assert cs.query_path(i, j) == want, (c["source"], c["query"], i, j)
    assert loops > 0


def test_cartesian_candidates_match_reference(ctx):
    """The device's layer-by-layer walk equals the reference's cartesian candidate sets on the
    sampled (i, j, v) of the yeast and synthetic queries, and the oracle on more samples."""
    rng = np.random.default_rng(11)
    n = extra = 0
    for c in _cases():
        if c["source"] == "c2":
            continue
        g = _data_graph(c["source"])
        q = A.Graph.from_edges(c["labels"], c["edges"])
        cs = ctx.filter(q, g)
        for r in c["cart"]:
            got = cs.cartesian_candidates(r["v"], c["paths"][r["i"]][r["j"]]).tolist()
            assert len(got) == r["n"], (c["source"], c["query"], r["i"], r["j"], r["v"])
            assert hashlib.sha256(" ".join(map(str, got)).encode()).hexdigest() == r["sha256"]
            n += 1
        # more (i, j, v) against the oracle, including vertices that are not candidates
        _, _, glab = g.csr()
        ocs = O.OCandSpace(O.OGraph.from_edges(c["labels"], c["edges"]), O.OGraph.from_edges(glab, g.edges()))
        for _ in range(12):
            i, j = (int(x) for x in rng.integers(0, c["nq"], 2))
            path = c["paths"][i][j]
            if i == j or not path or len(path) < 2:
                continue
            cand = ocs.candidates(i)
            v = int(cand[rng.integers(0, len(cand))]) if rng.random() < 0.8 else int(rng.integers(0, g.n))
            assert cs.cartesian_candidates(v, path).tolist() == ocs.cartesian(v, path).tolist()
            extra += 1
    assert n >= 500 and extra >= 50


def test_cartesian_candidates_edge_cases(ctx):
    """A path graph with a closed answer, a vertex outside C(path[0]), a non-arc step, and the
    output capacity."""
    g = A.Graph.from_edges([0, 1, 0, 1], [[0, 1], [1, 2], [2, 3]])
    q = A.Graph.from_edges([0, 1, 0], [[0, 1], [1, 2]])
    cs = ctx.filter(q, g)
    assert cs.cartesian_candidates(0, [0, 1, 2]).tolist() == [2]
    assert cs.cartesian_candidates(2, [0, 1, 2]).tolist() == [0]
    assert cs.cartesian_candidates(1, [0, 1, 2]).tolist() == []
    assert cs.cartesian_candidates(0, [0, 2]).tolist() == []  # 0 and 2 are not adjacent in the query
    assert cs.query_path(0, 2) == [0, 1, 2]


def _cartesian_order(q, rng):
    """An order of q's vertices whose second vertex is not adjacent to the first (a cartesian
    step); later vertices follow adjacency where they can, so there may be more such steps."""
    off, nbr, _ = q.csr()
    n = q.n
    adj = [set(nbr[off[u]:off[u + 1]].tolist()) for u in range(n)]
    for _ in range(50):
        a = int(rng.integers(n))
        far = [w for w in range(n) if w != a and w not in adj[a]]
        if far:
            break
    else:
        return None
    order = [a, far[int(rng.integers(len(far)))]]
    seen = set(order)
    while len(order) < n:
        nxt = [w for w in range(n) if w not in seen and adj[w] & seen]
        w = min(nxt) if nxt else min(w for w in range(n) if w not in seen)
        order.append(w)
        seen.add(w)
    return order


def test_cartesian_orders_keep_counts(ctx):
    """Explicit orders with cartesian steps (asdm_plan_create and asdm_plan_create_ordered, which
    refused them before) count exactly the reference's numbers: the step's candidates come from a
    cartesian arc built from the query path to the cartesian parent."""
    gold = T.yeast_golden()["queries"]
    g = A.Graph.load(T.yeast_data_path())
    rng = np.random.default_rng(5)
    names = [n for n in sorted(gold) if not gold[n]["empty"]][::6][:14]
    assert len(names) >= 10
    n = 0
    for name in names:
        q = A.Graph.load(T.yeast_query_path(name))
        order = _cartesian_order(q, rng)
        if order is None:
            continue
        cs = ctx.filter(q, g)
        want = gold[name]["count"]
        assert ctx.count(cs.plan(q, order=order), cs) == want, name
        assert ctx.count(cs.plan(q, order=order, num_enumerated=q.n - 1), cs) == want, name
        if want < 10 ** 7:
            assert ctx.count(cs.plan(q, order=order, factor=False), cs) == want, name
        n += 1
    assert n >= 8


def test_cartesian_orders_keep_local_counts(ctx):
    """The reference's LOCAL_COUNT matrices of the synthetic golden cases, through plans whose
    orders have cartesian steps."""
    rng = np.random.default_rng(9)
    n = 0
    for case in T.synth_golden():
        if not case.get("local"):
            continue
        g = T.make_graph(case)
        for r in case["queries"]:
            if r.get("local_npz") is None:
                continue
            q = g.random_walk_query(r["k"], r["density"], r["max_extra"], r["seed"])
            order = _cartesian_order(q, rng)
            if order is None:
                continue
            cs = ctx.filter(q, g)
            mat, c = ctx.local_count(cs.plan(q, order=order, local=True), cs)
            want = np.load(os.path.join(T.GOLDEN, r["local_npz"]))["local"]
            assert c == r["count"]
            assert np.array_equal(mat, want), (case["name"], r["seed"])
            n += 1
    assert n >= 5
