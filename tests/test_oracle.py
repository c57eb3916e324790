"""Pins the CPU oracle (oracle/asdm_oracle.c) before it is trusted as the checker:
against the reference's own outputs (golden fixtures made by tests/golden/make_golden.py from the
reference built from its sources), against brute force, closed forms and invariants."""
import itertools
import os
import math
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests import common as T


@pytest.fixture(scope="module")
def yeast():
    return O.OGraph.load(T.yeast_data_path())


def test_yeast_counts_match_reference(yeast):
    gold = T.yeast_golden()["queries"]
    assert len(gold) >= 100
    for name, rec in gold.items():
        q = O.OGraph.load(T.yeast_query_path(name))
        cs = O.OCandSpace(q, yeast)
        assert cs.empty == rec["empty"], name
        got = 0 if cs.empty else cs.count()
        assert got == rec["count"], name


def test_yeast_candidate_sets_match_reference(yeast):
    gold = T.yeast_golden()["queries"]
    n = 0
    for name, rec in gold.items():
        if "cs" not in rec or rec["cs"] is None:
            continue
        q = O.OGraph.load(T.yeast_query_path(name))
        cs = O.OCandSpace(q, yeast)
        assert not cs.empty
        for u, ref in enumerate(rec["cs"]["cand"]):
            assert np.array_equal(cs.candidates(u), np.array(ref, np.uint32)), (name, u)
        for a, (u, u2, w) in enumerate(rec["cs"]["arc_weight"]):
            assert cs.arc_weight(a) == w, (name, a)
        n += 1
    assert n >= 30


def test_synthetic_counts_and_local_counts_match_reference():
    for case in T.synth_golden():
        g = T.make_graph(case)
        assert T.text_sha256(g) == case["data_sha256"], case["name"]
        og = T.to_oracle(g)
        for rec in case["queries"]:
            q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
            assert T.text_sha256(q) == rec["query_sha256"]
            oq = T.to_oracle(q)
            cs = O.OCandSpace(oq, og)
            assert [len(cs.candidates(u)) for u in range(q.n)] == rec["cand_sizes"]
            assert [cs.arc_weight(a) for a in range(len(rec["arc_weight"]))] == [w for _, _, w in rec["arc_weight"]]
            if rec.get("local_npz"):
                mat, c = cs.local_count()
                ref = np.load(f"{T.GOLDEN}/{rec['local_npz']}")["local"]
                assert c == rec["count"]
                assert np.array_equal(mat, ref), (case["name"], rec["seed"])
            elif rec["count"] is not None and rec["count"] < 10**7:
                assert cs.count() == rec["count"]


def _random_graph(rng, n, p, nl):
    labels = [rng.randrange(nl) for _ in range(n)]
    edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
    return labels, edges


def _random_connected_query(rng, k, nl, extra):
    labels = [rng.randrange(nl) for _ in range(k)]
    edges = [(rng.randrange(i), i) for i in range(1, k)]
    cand = [(a, b) for a in range(k) for b in range(a + 1, k) if (a, b) not in edges]
    rng.shuffle(cand)
    return labels, edges + cand[:extra]


def test_oracle_equals_bruteforce_random():
    rng = random.Random(7)
    for trial in range(60):
        gl, ge = _random_graph(rng, rng.randrange(8, 30), rng.choice([0.1, 0.2, 0.4]), rng.randrange(1, 4))
        ql, qe = _random_connected_query(rng, rng.randrange(2, 6), max(gl) + 1, rng.randrange(0, 3))
        g = O.OGraph.from_edges(gl, ge)
        q = O.OGraph.from_edges(ql, qe)
        assert O.count(q, g) == O.bruteforce_count(q, g), trial


def _falling(x, k):
    r = 1
    for i in range(k):
        r *= x - i
    return r


def test_closed_forms():
    rng = random.Random(3)
    gl, ge = _random_graph(rng, 40, 0.2, 2)
    g = O.OGraph.from_edges(gl, ge)
    adj = [set() for _ in gl]
    for a, b in ge:
        adj[a].add(b)
        adj[b].add(a)
    # star: centre label 0, k leaves of label 1 -> sum_v [L(v)=0] (d_1(v))_k
    for k in (1, 2, 3):
        q = O.OGraph.from_edges([0] + [1] * k, [(0, i) for i in range(1, k + 1)])
        want = sum(_falling(sum(1 for w in adj[v] if gl[w] == 1), k) for v in range(len(gl)) if gl[v] == 0)
        assert O.count(q, g) == want
    # unlabeled triangle -> 6 x triangles
    g1 = O.OGraph.from_edges([0] * len(gl), ge)
    tri = sum(1 for a, b, c in itertools.combinations(range(len(gl)), 3) if b in adj[a] and c in adj[a] and c in adj[b])
    assert O.count(O.OGraph.from_edges([0, 0, 0], [(0, 1), (1, 2), (0, 2)]), g1) == 6 * tri
    # path of two edges, one label -> sum_v d(v)(d(v)-1)
    assert O.count(O.OGraph.from_edges([0, 0, 0], [(0, 1), (1, 2)]), g1) == sum(len(s) * (len(s) - 1) for s in adj)
    # K_k in K_n, unlabeled -> n!/(n-k)!
    n = 7
    kn = O.OGraph.from_edges([0] * n, list(itertools.combinations(range(n), 2)))
    for k in (2, 3, 4):
        kk = O.OGraph.from_edges([0] * k, list(itertools.combinations(range(k), 2)))
        assert O.count(kk, kn) == math.factorial(n) // math.factorial(n - k)


def test_count_is_order_independent():
    rng = random.Random(11)
    gl, ge = _random_graph(rng, 25, 0.3, 2)
    g = O.OGraph.from_edges(gl, ge)
    ql, qe = [0, 1, 0, 1, 0], [(0, 1), (1, 2), (2, 3), (3, 4), (0, 2)]
    q = O.OGraph.from_edges(ql, qe)
    cs = O.OCandSpace(q, g)
    base = cs.count()
    assert base == O.bruteforce_count(q, g)
    for order in ([0, 1, 2, 3, 4], [2, 0, 1, 3, 4], [4, 3, 2, 1, 0], [1, 0, 2, 3, 4]):
        assert cs.count(order) == base


def test_local_count_invariants():
    rng = random.Random(5)
    gl, ge = _random_graph(rng, 30, 0.25, 2)
    g = O.OGraph.from_edges(gl, ge)
    q = O.OGraph.from_edges([0, 1, 0, 1], [(0, 1), (1, 2), (2, 3)])
    cs = O.OCandSpace(q, g)
    mat, c = cs.local_count()
    assert c == O.bruteforce_count(q, g)
    assert all(int(s) == c for s in mat.sum(axis=0))
    for u, lu in enumerate([0, 1, 0, 1]):
        assert all(mat[v, u] == 0 for v in range(len(gl)) if gl[v] != lu)


def test_unique_combine_matches_bruteforce_twin():
    rng = random.Random(2)
    for _ in range(200):
        k = rng.randrange(1, 6)
        sets = [sorted(rng.sample(range(12), rng.randrange(0, 8))) for _ in range(k)]
        assert O.unique_combine(sets) == O.bf_unique_combine(sets)
    a, b = [1, 2, 3, 5], [2, 5, 7]
    assert O.unique_combine([a, b]) == len(a) * len(b) - 2


def _paths_cases():
    import json

    with open(os.path.join(T.GOLDEN, "paths.json")) as f:
        return json.load(f)


def _weights32(case, off, nbr):
    wmap = {(u, u2): w for u, u2, w in case["arc_weight"]}
    return [wmap[(u, int(nbr[e]))] & 0xFFFFFFFF for u in range(case["nq"]) for e in range(off[u], off[u + 1])]


def test_query_paths_match_reference():
    """setQueryGraphWeights + allSourcesShortestPaths + reconstructPath, restated, against the
    reference's own distances, successor matrices and paths (ref_cs_dump --paths) on yeast,
    synthetic and config-2 queries; the config-2 ones have products that wrap mod 2^64 and pairs
    whose reference reconstruction would not end (None in the fixture)."""
    cases = _paths_cases()
    assert len(cases) >= 16
    wrapped = loops = 0
    for c in cases:
        q = O.OGraph.from_edges(c["labels"], c["edges"])
        off, nbr, _ = q.csr()
        dist, nxt = O.paths_from_weights(c["nq"], off, nbr, _weights32(c, off, nbr))
        assert dist.tolist() == c["dist"], (c["source"], c["query"])
        assert nxt.tolist() == c["next"], (c["source"], c["query"])
        for i in range(c["nq"]):
            for j in range(c["nq"]):
                assert O.reconstruct_path(nxt, i, j) == c["paths"][i][j], (c["source"], c["query"], i, j)
        wrapped += sum(1 for row in c["dist"] for x in row if x >= 1 << 63 and x != (1 << 64) - 1)
        loops += sum(1 for row in c["paths"] for p in row if p is None)
    assert wrapped > 0 and loops > 0  # the fixtures reach the wrapping behaviour


def test_query_paths_invariants():
    """What the definition fixes, independently of the fixtures: every reconstructed path walks
    query edges from i to j, and without wrap-around its product of weights is dist[i][j] and no
    single edge beats it."""
    for c in _paths_cases()[:16]:
        q = O.OGraph.from_edges(c["labels"], c["edges"])
        off, nbr, _ = q.csr()
        w = _weights32(c, off, nbr)
        dist, nxt = O.paths_from_weights(c["nq"], off, nbr, w)
        wmap = {(u, int(nbr[e])): w[e] for u in range(c["nq"]) for e in range(off[u], off[u + 1])}
        for i in range(c["nq"]):
            for j in range(c["nq"]):
                p = O.reconstruct_path(nxt, i, j)
                if i == j or not p:
                    continue
                assert p[0] == i and p[-1] == j
                prod = 1
                for a, b in zip(p, p[1:]):
                    assert (a, b) in wmap
                    prod *= wmap[(a, b)]
                if prod < 1 << 64:
                    assert int(dist[i, j]) == prod
                    if (i, j) in wmap:
                        assert prod <= wmap[(i, j)]


def test_cartesian_candidates_match_reference():
    """optimizedCartesianProduct, restated layer by layer, against the reference's own sets on the
    sampled (i, j, v) of the yeast and synthetic queries."""
    import hashlib

    synth = {c["name"]: c for c in T.synth_golden()}
    gy = O.OGraph.load(T.yeast_data_path())
    n = 0
    for c in _paths_cases():
        if c["source"] == "c2":
            continue
        q = O.OGraph.from_edges(c["labels"], c["edges"])
        if c["source"] == "yeast":
            g = gy
        else:
            gg = T.make_graph(synth[c["source"]])
            _, _, lab = gg.csr()
            g = O.OGraph.from_edges(lab, gg.edges())
        cs = O.OCandSpace(q, g)
        for r in c["cart"]:
            got = cs.cartesian(r["v"], c["paths"][r["i"]][r["j"]]).tolist()
            assert len(got) == r["n"]
            assert hashlib.sha256(" ".join(map(str, got)).encode()).hexdigest() == r["sha256"]
            if "ids" in r:
                assert got == r["ids"]
            n += 1
    assert n >= 500


def test_cartesian_candidates_small_cases():
    """Closed cases: a path whose first vertex is not matched to a candidate reaches nothing; on a
    path graph the two-hop set is the data vertices two steps away with the right labels."""
    # data: 0-1-2-3 path, labels 0,1,0,1; query: a-b-c path, labels 0,1,0
    g = O.OGraph.from_edges([0, 1, 0, 1], [[0, 1], [1, 2], [2, 3]])
    q = O.OGraph.from_edges([0, 1, 0], [[0, 1], [1, 2]])
    cs = O.OCandSpace(q, g)
    assert cs.candidates(0).tolist() == [0, 2] and cs.candidates(2).tolist() == [0, 2]
    assert cs.cartesian(0, [0, 1, 2]).tolist() == [2]   # 0 -> 1 -> {0, 2}, without 0 itself
    assert cs.cartesian(2, [0, 1, 2]).tolist() == [0]   # 2 -> {1, 3}: 1 -> {0, 2}, 3 -> {2}; without 2
    assert cs.cartesian(1, [0, 1, 2]).tolist() == []    # 1 is not a candidate of query vertex 0
