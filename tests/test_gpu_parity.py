"""GPU parity: the HIP path (through the C ABI) against the reference's golden outputs and the
oracle, bit-exact.  Every test here runs kernels on the MI355X."""
import random
import sys

import numpy as np
import pytest

import asdmatch_amd as A
from oracle import oracle as O
from tests import common as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def yeast():
    return A.Graph.load(T.yeast_data_path())


def test_intersect_k_matches_numpy(ctx):
    rng = np.random.default_rng(1)
    for ratio in (1, 3, 10, 100, 1000, 10000):
        for k in (1, 2, 3, 5):
            small = 50
            lists = []
            for j in range(k):
                n = small if j == 0 else small * ratio
                lists.append(np.unique(rng.integers(0, 4 * small * ratio, size=n, dtype=np.uint64)).astype(np.uint32))
            want = lists[0]
            for l in lists[1:]:
                want = np.intersect1d(want, l)
            got = ctx.intersect(lists)
            assert np.array_equal(got, want), (ratio, k)


def test_intersect_k_merge_path_tiles(ctx):
    """Long balanced lists, intersected by merge path over many 2,048-position tiles: dense overlaps
    (so equal elements straddle tile boundaries), size ratios on both sides of the merge-path
    cutoff (16), every list order, an empty list, a single list, and identical lists."""
    rng = np.random.default_rng(7)
    def rand_list(n, hi):
        return np.unique(rng.integers(0, hi, size=n, dtype=np.uint64)).astype(np.uint32)
    a, b = rand_list(200_000, 400_000), rand_list(200_000, 400_000)
    cases = [
        [a, b], [b, a], [a, a], [a, b, rand_list(150_000, 400_000)],
        [rand_list(10_000, 1_000_000), rand_list(160_000, 1_000_000)],  # ratio just under 16
        [rand_list(10_000, 1_000_000), rand_list(170_000, 1_000_000)],  # just over
        [np.arange(0, 300_000, 2, dtype=np.uint32), np.arange(0, 300_000, 3, dtype=np.uint32)],
        [a, np.zeros(0, np.uint32)], [a],
    ]
    for lists in cases:
        want = lists[0]
        for l in lists[1:]:
            want = np.intersect1d(want, l)
        got = ctx.intersect(lists)
        assert np.array_equal(got, want), [len(l) for l in lists]
        assert np.array_equal(ctx.intersect_noalloc(lists), want) if sum(len(l) for l in lists) <= 1 << 21 else True


def test_filter_matches_reference_on_yeast(ctx, yeast):
    gold = T.yeast_golden()["queries"]
    n = 0
    for name, rec in gold.items():
        if "cs" not in rec:
            continue
        q = A.Graph.load(T.yeast_query_path(name))
        cs = ctx.filter(q, yeast)
        if rec["cs"] is None:
            assert cs is None, name
            continue
        assert cs is not None, name
        for u, ref in enumerate(rec["cs"]["cand"]):
            assert np.array_equal(cs.candidates(u), np.array(ref, np.uint32)), (name, u)
        for a, (_, _, w) in enumerate(rec["cs"]["arc_weight"]):
            assert cs.arc_weight(a) == w, (name, a)
        n += 1
    assert n >= 30


def test_counts_match_reference_on_yeast(ctx, yeast):
    gold = T.yeast_golden()["queries"]
    for name, rec in gold.items():
        q = A.Graph.load(T.yeast_query_path(name))
        got = A.count_embeddings(ctx, q, yeast)
        assert got == rec["count"], name


def test_root_matches_oracle(ctx, yeast):
    og = O.OGraph.load(T.yeast_data_path())
    for name in list(T.yeast_golden()["queries"])[::5]:
        q = A.Graph.load(T.yeast_query_path(name))
        cs = ctx.filter(q, yeast)
        ocs = O.OCandSpace(O.OGraph.load(T.yeast_query_path(name)), og)
        if cs is None:
            assert ocs.empty
            continue
        assert cs.root == ocs.root, name


def test_synthetic_counts_match_reference(ctx):
    for case in T.synth_golden():
        g = T.make_graph(case)
        for rec in case["queries"]:
            if rec["count"] is None:
                continue
            q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
            cs = ctx.filter(q, g)
            assert list(cs.sizes()) == rec["cand_sizes"], (case["name"], rec["seed"])
            plan = cs.plan(q)
            say(case["name"], rec["seed"], plan.order())
            assert ctx.count(plan, cs) == rec["count"], (case["name"], rec["seed"])


def test_budget_does_not_change_counts(ctx):
    case = [c for c in T.synth_golden() if c["name"] == "pl_small_l4"][0]
    g = T.make_graph(case)
    for rec in case["queries"]:
        q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
        cs = ctx.filter(q, g)
        plan = cs.plan(q, bags=False)
        for budget in (1 << 12, 1 << 16, 1 << 20, 0):
            assert ctx.count(plan, cs, budget) == rec["count"], budget


def test_order_does_not_change_counts(ctx):
    case = [c for c in T.synth_golden() if c["name"] == "pl_small_l4"][0]
    g = T.make_graph(case)
    rng = random.Random(4)
    for rec in case["queries"]:
        q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
        e = q.edges().tolist()
        adj = {i: set() for i in range(q.n)}
        for a, b in e:
            adj[a].add(b)
            adj[b].add(a)
        cs = ctx.filter(q, g)
        for _ in range(4):
            start = rng.randrange(q.n)
            order, seen = [start], {start}
            while len(order) < q.n:
                fr = sorted({w for u in order for w in adj[u]} - seen)
                w = rng.choice(fr)
                order.append(w)
                seen.add(w)
            plan = cs.plan(q, order=order)
            assert ctx.count(plan, cs) == rec["count"], order


def _rand_graph(rng, n, p, nl):
    labels = [rng.randrange(nl) for _ in range(n)]
    edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
    return labels, edges


def test_same_label_leaf_groups_against_bruteforce(ctx):
    """Stars and trees whose leaves share labels exercise the inclusion-exclusion leaf counting;
    triangles and diamonds exercise the per-candidate-edge triangle tables."""
    rng = random.Random(9)
    for trial in range(12):
        gl, ge = _rand_graph(rng, 60, 0.15, 2)
        g, og = T.graph_from_lists(gl, ge)
        for ql, qe in (
            ([0, 1, 1, 1], [(0, 1), (0, 2), (0, 3)]),
            ([0, 1, 1, 1, 1], [(0, 1), (0, 2), (0, 3), (0, 4)]),
            ([0, 1, 0, 1, 1], [(0, 1), (1, 2), (2, 3), (2, 4), (0, 4)]),
            ([1, 0, 1, 0, 1, 1], [(0, 1), (1, 2), (2, 3), (1, 4), (3, 5)]),
            # triangle and diamond leaves (edge-table path)
            ([0, 1, 1], [(0, 1), (1, 2), (0, 2)]),
            ([0, 1, 0, 1], [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]),
            ([0, 0, 1, 1, 0], [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4), (2, 4)]),
        ):
            q, oq = T.graph_from_lists(ql, qe)
            want = O.bruteforce_count(oq, og)
            assert A.count_embeddings(ctx, q, g) == want, (trial, ql, qe)


def test_row_sum_table_against_bruteforce(ctx):
    """A fused level whose last enumerated vertex has one backward neighbour and no same-label
    constraint finishes each row from a per-position table (k_tail_sums) instead of scanning the
    row's candidates.  Paths, stars, tadpoles and triangle fans with distinct labels near that
    vertex take the table; the others (shared labels) take the per-candidate path.  Both engines
    must give the brute-force count, and the table must have been used."""
    rng = random.Random(21)
    queries = (
        ([0, 1, 2, 3], [(0, 1), (1, 2), (2, 3)]),
        ([0, 1, 2, 3], [(0, 1), (0, 2), (0, 3)]),
        ([0, 1, 2, 3], [(0, 1), (1, 2), (0, 2), (2, 3)]),
        ([0, 1, 2, 3, 0], [(0, 1), (1, 2), (2, 3), (3, 4)]),
        ([0, 1, 2, 3], [(0, 1), (1, 2), (0, 2), (1, 3), (2, 3)]),
        ([0, 1, 2, 3, 1], [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4)]),
        ([0, 1, 1, 2, 3], [(0, 1), (0, 2), (1, 3), (2, 4)]),
    )
    used = 0
    for trial in range(6):
        gl, ge = _rand_graph(rng, 80, 0.12, 4)
        g, og = T.graph_from_lists(gl, ge)
        qs, want = [], []
        for ql, qe in queries:
            q, oq = T.graph_from_lists(ql, qe)
            qs.append(q)
            want.append(O.bruteforce_count(oq, og))
        ctx.set_profiling(True)
        try:
            got = [A.count_embeddings(ctx, q, g) for q in qs]
            batch, _ = ctx.batch_count(g, qs)
            used += sum(k["launches"] for k in ctx.kernel_stats() if k["name"] == "tail_sums")
        finally:
            ctx.set_profiling(False)
        assert got == want, trial
        assert [int(c) for c in batch] == want, trial
    assert used > 0


def test_row_sum_table_corrections_against_bruteforce(ctx):
    """Forced plans with same-label constraints around the last enumerated vertex d, all on the
    row-sum table: a same-label earlier depth of d (query 1, corrected at that vertex); one
    correction factor -- a hoisted leaf with d's label (2), a same-label matched vertex of a pendant
    leaf (3) or of a triangle leaf (4) of d, a hoisted triangle leaf with d's label (7) -- alone or
    with a same-label depth (5).  Query 6 has two correction factors and takes the per-candidate
    path.  Every count must be the brute force's, and the table must have been used."""
    A_, B_, C_ = 0, 1, 2
    cases = (  # labels, edges, order, enumerated depths
        ([A_, B_, A_, C_], [(0, 1), (1, 2), (2, 3)], [1, 0, 2, 3], 3),
        ([B_, A_, C_, A_], [(0, 1), (0, 2), (2, 3)], [0, 2, 3, 1], 3),
        ([A_, B_, C_, A_], [(0, 1), (1, 2), (2, 3)], [1, 0, 2, 3], 3),
        ([A_, B_, C_, A_], [(0, 1), (1, 2), (1, 3), (2, 3)], [1, 0, 2, 3], 3),
        ([A_, B_, A_, A_], [(0, 1), (1, 2), (2, 3)], [1, 0, 2, 3], 3),
        ([B_, A_, C_, A_, C_, A_], [(0, 1), (0, 2), (2, 3), (3, 4), (0, 5)], [0, 2, 5, 3, 1, 4], 4),
        ([B_, C_, A_, A_], [(0, 1), (0, 2), (1, 2), (1, 3)], [0, 1, 3, 2], 3),
    )
    rng = random.Random(33)
    used = 0
    for trial in range(8):
        gl, ge = _rand_graph(rng, 60, 0.15, 3)
        g, og = T.graph_from_lists(gl, ge)
        for ql, qe, order, D in cases:
            q, oq = T.graph_from_lists(ql, qe)
            want = O.bruteforce_count(oq, og)
            cs = ctx.filter(q, g)
            if cs is None:
                assert want == 0
                continue
            ctx.set_profiling(True)
            try:
                got = ctx.count(cs.plan(q, order=order, num_enumerated=D), cs)
                used += sum(k["launches"] for k in ctx.kernel_stats() if k["name"] == "tail_sums")
            finally:
                ctx.set_profiling(False)
            assert got == want, (trial, ql, qe, order, D)
    assert used > 0


def test_second_row_sum_table_against_bruteforce(ctx):
    """Forced plans where depth c = D-2 has one backward neighbour: rows of width D-2 are finished
    from a second table (two table launches) -- a path (1), a path whose depth c has a same-label
    earlier depth (2), a triangle leaf on (c, D-1) (4), a hoisted leaf on depth 0 (5); with the last
    vertex's backward neighbour before c, alone (6) or with a hoisted leaf pendant on c (7); with
    both a pendant leaf on c and the last vertex hanging on c (8) -- but not when a hoisted leaf
    reads c through its same-label set (3)."""
    cases = (  # labels, edges, order, enumerated depths, second table expected
        ([0, 1, 2, 3], [(0, 1), (1, 2), (2, 3)], [0, 1, 2, 3], 3, True),
        ([1, 0, 1, 2, 3], [(0, 1), (1, 2), (2, 3), (3, 4)], [0, 1, 2, 3, 4], 4, True),
        ([0, 1, 2, 3, 1], [(0, 1), (1, 2), (2, 3), (0, 4)], [0, 1, 2, 3, 4], 3, False),
        ([0, 1, 2, 3], [(0, 1), (1, 2), (1, 3), (2, 3)], [0, 1, 2, 3], 3, True),
        ([0, 1, 2, 3, 0], [(0, 1), (1, 2), (2, 3), (0, 4)], [0, 1, 2, 3, 4], 3, True),
        ([0, 1, 2, 3, 4], [(0, 1), (1, 2), (1, 3), (3, 4)], [0, 1, 2, 3, 4], 4, True),
        ([0, 1, 2, 3, 4], [(0, 1), (1, 2), (1, 3), (2, 4)], [0, 1, 2, 3, 4], 4, True),
        ([0, 1, 2, 3, 4], [(0, 1), (1, 2), (2, 3), (1, 4)], [0, 1, 2, 3, 4], 3, True),
    )
    rng = random.Random(44)
    for trial in range(6):
        gl, ge = _rand_graph(rng, 70, 0.16, 5)
        g, og = T.graph_from_lists(gl, ge)
        for ql, qe, order, D, second in cases:
            q, oq = T.graph_from_lists(ql, qe)
            want = O.bruteforce_count(oq, og)
            cs = ctx.filter(q, g)
            if cs is None:
                assert want == 0
                continue
            ctx.set_profiling(True)
            try:
                got = ctx.count(cs.plan(q, order=order, num_enumerated=D), cs)
                tables = sum(k["launches"] for k in ctx.kernel_stats() if k["name"] == "tail_sums")
            finally:
                ctx.set_profiling(False)
            assert got == want, (trial, ql, qe, order, D)
            assert tables == (2 if second else 1), (trial, ql, tables)


def test_edge_row_sum_table_against_bruteforce(ctx):
    """Forced plans whose last enumerated vertex d closes a triangle with its two backward
    neighbours b1 < b2: its candidates are a function of the candidate edge (row[b1], row[b2]), and
    rows are finished from an edge-indexed table (one edge_tail_sums launch) -- with a pendant leaf
    (1), a same-label earlier depth of d (2), a triangle leaf on (b1, d) (3) or on (b2, d) (4) --
    but not when a leaf of d has a same-label matched vertex to avoid (5)."""
    cases = (  # labels, edges, order, enumerated depths, edge table expected
        ([0, 1, 2, 3], [(0, 1), (0, 2), (1, 2), (2, 3)], [0, 1, 2, 3], 3, True),
        ([1, 0, 2, 1, 3], [(0, 1), (1, 2), (1, 3), (2, 3), (3, 4)], [0, 1, 2, 3, 4], 4, True),
        ([0, 1, 2, 3], [(0, 1), (0, 2), (1, 2), (0, 3), (2, 3)], [0, 1, 2, 3], 3, True),
        ([0, 1, 2, 3], [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)], [0, 1, 2, 3], 3, True),
        ([0, 1, 2, 0], [(0, 1), (0, 2), (1, 2), (2, 3)], [0, 1, 2, 3], 3, False),
    )
    rng = random.Random(55)
    for trial in range(6):
        gl, ge = _rand_graph(rng, 70, 0.16, 4)
        g, og = T.graph_from_lists(gl, ge)
        for ql, qe, order, D, edge in cases:
            q, oq = T.graph_from_lists(ql, qe)
            want = O.bruteforce_count(oq, og)
            cs = ctx.filter(q, g)
            if cs is None:
                assert want == 0
                continue
            ctx.set_profiling(True)
            try:
                got = ctx.count(cs.plan(q, order=order, num_enumerated=D), cs)
                tables = sum(k["launches"] for k in ctx.kernel_stats() if k["name"] == "edge_tail_sums")
            finally:
                ctx.set_profiling(False)
            assert got == want, (trial, ql, qe, order, D)
            assert tables == (1 if edge else 0), (trial, ql, tables)


def _embeddings_bruteforce(ql, qe, gl, ge):
    """All injective label- and edge-preserving maps, as tuples of data ids by query vertex."""
    gadj = {v: set() for v in range(len(gl))}
    for a, b in ge:
        gadj[a].add(b)
        gadj[b].add(a)
    out, f = [], []

    def rec(u):
        if u == len(ql):
            out.append(tuple(f))
            return
        for v in range(len(gl)):
            if gl[v] != ql[u] or v in f:
                continue
            if all(f[w] in gadj[v] for a, b in qe for w in ((a,) if b == u else (b,) if a == u else ()) if w < u):
                f.append(v)
                rec(u + 1)
                f.pop()
    rec(0)
    return sorted(out)


def test_enumerate_matches_bruteforce(ctx):
    """COLLECT_RESULT: every embedding, not just the count.  Paths, stars, triangles and a 4-cycle,
    with shared labels, against a brute-force enumeration; the count equals asdm_count; a buffer
    too small is an error."""
    rng = random.Random(66)
    queries = (
        ([0, 1, 0], [(0, 1), (1, 2)]),
        ([0, 1, 1, 1], [(0, 1), (0, 2), (0, 3)]),
        ([0, 1, 2], [(0, 1), (1, 2), (0, 2)]),
        ([0, 1, 0, 1], [(0, 1), (1, 2), (2, 3), (3, 0)]),
        ([1, 0, 2, 0, 1], [(0, 1), (1, 2), (2, 3), (3, 4), (1, 3)]),
    )
    for trial in range(4):
        gl, ge = _rand_graph(rng, 24, 0.25, 3)
        g, _ = T.graph_from_lists(gl, ge)
        for ql, qe in queries:
            want = _embeddings_bruteforce(ql, qe, gl, ge)
            q, _ = T.graph_from_lists(ql, qe)
            cs = ctx.filter(q, g)
            if cs is None:
                assert want == []
                continue
            plan = cs.plan(q, factor=False)
            rows = ctx.enumerate(plan, cs, max_rows=len(want) + 5)
            assert sorted(map(tuple, rows.tolist())) == want, (trial, ql, qe)
            assert ctx.count(cs.plan(q), cs) == len(want)
            if len(want) > 1:
                with pytest.raises(A.AsdmError) as e:
                    ctx.enumerate(plan, cs, max_rows=len(want) - 1)
                assert e.value.code == -6
    # a real data graph: every row is an embedding, and there are as many as the reference counts
    yeast = A.Graph.load(T.yeast_data_path())
    goff, gnbr, glab = yeast.csr()
    gold = T.yeast_golden()["queries"]
    names = [n for n in gold if 0 < gold[n]["count"] <= 200_000][:3]
    assert names
    for name in names:
        q = A.Graph.load(T.yeast_query_path(name))
        cs = ctx.filter(q, yeast)
        rows = ctx.enumerate(cs.plan(q, factor=False), cs, max_rows=gold[name]["count"])
        assert len(rows) == gold[name]["count"], name
        assert len({tuple(r) for r in rows.tolist()}) == len(rows)
        qoff, qnbr, qlab = q.csr()
        assert np.array_equal(glab[rows], np.broadcast_to(qlab, rows.shape))
        assert all(len(set(r)) == q.n for r in rows[:2000].tolist())
        for u in range(q.n):
            for w in qnbr[qoff[u]:qoff[u + 1]]:
                a, b = rows[:, u], rows[:, int(w)]
                ok = [int(bb) in set(gnbr[goff[aa]:goff[aa + 1]].tolist()) for aa, bb in zip(a[:500], b[:500])]
                assert all(ok), (name, u, int(w))


def test_tiny_queries_in_every_engine(ctx):
    """Corner shapes -- one vertex, one edge (distinct and equal labels), a same-label path, a
    triangle and a star of one label -- through the single-query engine, the batch engine and
    enumeration, against brute force."""
    rng = random.Random(77)
    queries = (
        ([0], []),
        ([0, 1], [(0, 1)]),
        ([0, 0], [(0, 1)]),
        ([0, 0, 0], [(0, 1), (1, 2)]),
        ([0, 0, 0], [(0, 1), (1, 2), (0, 2)]),
        ([1, 0, 0, 0], [(0, 1), (0, 2), (0, 3)]),
    )
    for trial in range(3):
        gl, ge = _rand_graph(rng, 40, 0.2, 2)
        g, og = T.graph_from_lists(gl, ge)
        qs, want = [], []
        for ql, qe in queries:
            q, oq = T.graph_from_lists(ql, qe)
            qs.append(q)
            want.append(len(_embeddings_bruteforce(ql, qe, gl, ge)))
        assert [A.count_embeddings(ctx, q, g) for q in qs] == want, trial
        batch, _ = ctx.batch_count(g, qs)
        assert [int(c) for c in batch] == want, trial
        for q, w in zip(qs, want):
            cs = ctx.filter(q, g)
            if cs is None:
                assert w == 0
                continue
            assert len(ctx.enumerate(cs.plan(q, factor=False), cs, max_rows=w + 1)) == w


def test_local_counts_match_reference(ctx):
    for case in T.synth_golden():
        if not case["local"]:
            continue
        g = T.make_graph(case)
        for rec in case["queries"]:
            if not rec.get("local_npz"):
                continue
            q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
            ref = np.load(f"{T.GOLDEN}/{rec['local_npz']}")["local"]
            for bags in (None, False):  # the planner's choice, and the frontier engine
                mat, c = A.local_count(ctx, q, g, bags=bags)
                assert c == rec["count"]
                assert np.array_equal(mat, ref), (case["name"], rec["seed"], bags)
                assert all(int(s) == c for s in mat.sum(axis=0))


def test_local_counts_same_under_many_pair_batches(ctx):
    """The atomic-free aggregation sorts and reduces (cell, value) pairs in batches bounded by the
    frontier budget; a tiny budget splits every leaf chunk into many batches and must give the
    same matrices as the reference."""
    done = 0
    for case in T.synth_golden():
        if not case["local"]:
            continue
        g = T.make_graph(case)
        for rec in case["queries"]:
            if not rec.get("local_npz") or done >= 6:
                continue
            q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
            mat, c = A.local_count(ctx, q, g, budget_bytes=1 << 16, bags=False)
            ref = np.load(f"{T.GOLDEN}/{rec['local_npz']}")["local"]
            assert c == rec["count"]
            assert np.array_equal(mat, ref), (case["name"], rec["seed"])
            done += 1
    assert done > 0


def test_config2_count_matches_reference(ctx):
    """Config-2 scale (1M vertices, 15M edges): the reference's exact build counted this admitted
    query (2,265,571,276 embeddings, tests/golden/c2_reference.json)."""
    import json

    from asdmatch_amd import workloads as W

    ref = json.load(open(f"{T.GOLDEN}/c2_reference.json"))
    g = W.graph("c2")
    s = W.slots("c2")[ref["slot"]]
    q = W.query(g, s, ref["seed"])
    assert A.count_embeddings(ctx, q, g) == ref["count"]
